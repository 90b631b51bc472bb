// Host decoding of PIK's DC groups section (pikgpu_decode_dc_groups): the
// size table of the 256x256-block DC groups, then per group the quantized DC
// image (ANS tokens with zero runs, then the adaptive DC predictor undone)
// and the control fields -- AC strategy, quant field and AR sigma LUT ids
// under one ANS state. This is DecodeDCGroups / DecodeDCGroup
// (pik/compressed_dc.cc:276-345, 436-507) without DequantDC and the DC
// smoothing, which run on the GPU (KDcDequant, KEpf<true>); the int16 DC
// and the side information land in the frame-wide arrays pikgpu_frame
// takes.
#include <smmintrin.h>
#include <stdarg.h>
#include <stdio.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pikgpu.h"
#include "entropy.h"

namespace pikgpu {
namespace ent {
namespace {

constexpr int kDcGroupBlocks = 256;
constexpr int kDcContexts = 3;        // one histogram context per channel
constexpr int kDcAlphabet = 78;       // 39 value tokens + 39 zero-run tokens
constexpr int kRunTokens = 39;
constexpr int kFieldContexts = 3;     // strategy, AR parameters, quant field
constexpr int kStrategyContext = 0;
constexpr int kArContext = 1;
constexpr int kQuantContext = 2;
constexpr uint8_t kUnset = 0xFF;

// Token -> value of the hybrid integer code (pik/entropy_coder.h:336-400):
// tokens 0..15 are the value; token t >= 16 carries n - 1 extra bits,
// n = 4 + (t - 16) / 2, value 2^n + (t odd) * 2^(n - 1) + extra.
inline uint32_t ReadHybrid(int token, BitStream* bs) {
  if (token < 16) return uint32_t(token);
  const unsigned n = 4 + unsigned(token - 16) / 2;
  const uint32_t extra = bs->Read(n - 1);
  return (1u << n) + (uint32_t(token & 1) << (n - 1)) + extra;
}

inline int32_t UnpackSigned(uint32_t u) {
  return int32_t((u >> 1) ^ (((~u) & 1) - 1));
}

// ---------------------------------------------------------------------------
// The adaptive DC predictor (pik/dc_predictor.cc), decoding side. Values
// are int16 with the arithmetic of 16-bit SIMD lanes: saturating adds and
// subtracts inside the predictors, wrapping adds for residuals and costs.
// The predictor with the smallest cost is taken, costs compared as
// unsigned 16-bit values with the lowest index winning ties (the minpos
// instruction of the reference's default build).

inline int16_t SatAdd(int a, int b) {
  return int16_t(std::min(32767, std::max(-32768, a + b)));
}
inline int16_t SatSub(int a, int b) {
  return int16_t(std::min(32767, std::max(-32768, a - b)));
}
inline int16_t WrapAdd(int a, int b) { return int16_t(uint16_t(a + b)); }
// Average rounding towards minus infinity of the saturated sum.
inline int16_t Avg(int16_t a, int16_t b) { return int16_t(SatAdd(a, b) >> 1); }
inline int16_t ClampedGradient(int16_t n, int16_t w, int16_t l) {
  const int16_t g = SatSub(SatAdd(n, w), l);
  const int16_t lo = std::min(n, std::min(w, l));
  const int16_t hi = std::max(n, std::max(w, l));
  return std::min(std::max(lo, g), hi);
}
// |c - p| of the saturated difference, as the 16-bit lane holds it
// (|-32768| stays 0x8000).
inline uint16_t AbsResidual(int16_t c, int16_t p) {
  const int d = SatSub(c, p);
  return uint16_t(d < 0 ? -d : d);
}

// Chroma predictors from the neighbours n (above), w (left), l (above
// left) and r (above right).
inline void PredictXB(int16_t n, int16_t w, int16_t l, int16_t r,
                      int16_t p[8]) {
  p[0] = ClampedGradient(n, w, l);
  p[1] = Avg(n, w);
  p[2] = n;
  p[3] = Avg(n, r);
  p[4] = w;
  p[5] = Avg(w, l);
  p[6] = r;
  p[7] = Avg(Avg(w, r), n);
}

// Predictor i of PredictXB alone.
inline int16_t PredictOneXB(int i, int16_t n, int16_t w, int16_t l,
                            int16_t r) {
  switch (i) {
    case 0: return ClampedGradient(n, w, l);
    case 1: return Avg(n, w);
    case 2: return n;
    case 3: return Avg(n, r);
    case 4: return w;
    case 5: return Avg(w, l);
    case 6: return r;
    default: return Avg(Avg(w, r), n);
  }
}

// The eight luma predictions as one vector of 16-bit lanes, lane i =
// predictor i of PredictY: lanes 0..4 are averages of two operands (the
// saturating sum halved), 5..7 the operand itself.
__attribute__((target("sse4.1"))) inline __m128i PredictYv(int16_t n,
                                                          int16_t w,
                                                          int16_t l,
                                                          int16_t r) {
  const int16_t nw = Avg(n, w), g = ClampedGradient(n, w, l);
  const __m128i a = _mm_setr_epi16(nw, w, n, w, n, w, g, n);
  const __m128i b = _mm_setr_epi16(r, n, r, l, l, 0, 0, 0);
  const __m128i avg = _mm_srai_epi16(_mm_adds_epi16(a, b), 1);
  return _mm_blend_epi16(avg, a, 0xE0);
}

// AbsResidual of every lane: |c - p| of the saturated difference (the
// 16-bit absolute value keeps -32768 as 0x8000).
__attribute__((target("sse4.1"))) inline __m128i AbsResidualv(__m128i c,
                                                             __m128i p) {
  return _mm_abs_epi16(_mm_subs_epi16(c, p));
}

// One row of luma: `t` and `m` the two rows above (equal for row 1), `b`
// the row being decoded from `res`. x = 0 is predicted by the pixel above,
// x = 1 and the last x by the pixel to the left; in between the predictor
// with the smallest error on the pixel above (predicted from its own
// neighbours) plus on the pixel to the left, costs summed with 16-bit
// wrap-around and compared as unsigned values, the lowest index winning
// ties (the minpos instruction the reference's default build uses).
__attribute__((target("sse4.1"))) void ExpandRowY(int xs, const int16_t* res,
                                                  const int16_t* t,
                                                  const int16_t* m,
                                                  int16_t* b) {
  b[0] = WrapAdd(m[0], res[0]);
  if (xs >= 2) b[1] = WrapAdd(b[0], res[1]);
  if (xs >= 4) {
    // predictions for b[x - 1]
    __m128i pw = PredictYv(m[1], b[0], m[0], m[2]);
    alignas(16) int16_t pc[8];
    for (int x = 2; x < xs - 1; ++x) {
      const __m128i pn = PredictYv(t[x], m[x - 1], t[x - 1], t[x + 1]);
      const __m128i cost =
          _mm_add_epi16(AbsResidualv(_mm_set1_epi16(m[x]), pn),
                        AbsResidualv(_mm_set1_epi16(b[x - 1]), pw));
      const int best = _mm_extract_epi16(_mm_minpos_epu16(cost), 1);
      pw = PredictYv(m[x], b[x - 1], m[x - 1], m[x + 1]);
      _mm_store_si128(reinterpret_cast<__m128i*>(pc), pw);
      b[x] = WrapAdd(pc[best], res[x]);
    }
  }
  if (xs >= 2) b[xs - 1] = WrapAdd(b[xs - 2], res[xs - 1]);
}

// The chroma predictor of every x of a row (both chroma channels use it):
// the one with the smallest error on the co-located luma pixel, `ym` / `yb`
// the decoded luma rows above and at the row.
__attribute__((target("sse4.1"))) void ChromaPredictors(int xs,
                                                        const int16_t* ym,
                                                        const int16_t* yb,
                                                        uint8_t* best) {
  for (int x = 2; x < xs - 1; ++x) {
    int16_t p[8];
    PredictXB(ym[x], yb[x - 1], ym[x - 1], ym[x + 1], p);
    const __m128i pv = _mm_loadu_si128(reinterpret_cast<const __m128i*>(p));
    const __m128i cost = AbsResidualv(_mm_set1_epi16(yb[x]), pv);
    best[x] = uint8_t(_mm_extract_epi16(_mm_minpos_epu16(cost), 1));
  }
}

// One row of chroma (X and B): predictor best[x] of each pixel.
void ExpandRowXB(int xs, const uint8_t* best, const int16_t* const res[2],
                 const int16_t* const m[2], int16_t* const b[2]) {
  for (int k = 0; k < 2; ++k) {
    const int16_t* mk = m[k];
    int16_t* bk = b[k];
    bk[0] = WrapAdd(mk[0], res[k][0]);
    if (xs >= 2) bk[1] = WrapAdd(bk[0], res[k][1]);
    for (int x = 2; x < xs - 1; ++x) {
      const int16_t p =
          PredictOneXB(best[x], mk[x], bk[x - 1], mk[x - 1], mk[x + 1]);
      bk[x] = WrapAdd(p, res[k][x]);
    }
    if (xs >= 2) bk[xs - 1] = WrapAdd(bk[xs - 2], res[k][xs - 1]);
  }
}

// ExpandDC (pik/entropy_coder.cc:156-197 with ExpandY / ExpandXB,
// pik/dc_predictor.cc:561-646): residuals in `q` (3 planes of xs x ys,
// row stride xs) become the quantized DC in `dc`. Row 0 is coded as
// differences to the left neighbour.
void ExpandDc(int xs, int ys, const std::vector<int16_t>& q,
              std::vector<int16_t>* dc) {
  const size_t plane = size_t(xs) * ys;
  dc->assign(3 * plane, 0);
  int16_t* out = dc->data();
  for (int c = 0; c < 3; ++c) {
    const int16_t* r = q.data() + c * plane;
    int16_t* o = out + c * plane;
    o[0] = r[0];
    for (int x = 1; x < xs; ++x) o[x] = WrapAdd(o[x - 1], r[x]);
  }
  const int16_t* yq = q.data() + plane;
  int16_t* yo = out + plane;
  for (int y = 1; y < ys; ++y) {
    const int16_t* t = yo + size_t(y >= 2 ? y - 2 : 0) * xs;
    ExpandRowY(xs, yq + size_t(y) * xs, t, yo + size_t(y - 1) * xs,
               yo + size_t(y) * xs);
  }
  std::vector<uint8_t> best(size_t(std::max(xs, 1)), 0);
  for (int y = 1; y < ys; ++y) {
    const size_t my = size_t(y - 1) * xs, by = size_t(y) * xs;
    const int16_t* res[2] = {q.data() + by, q.data() + 2 * plane + by};
    const int16_t* m[2] = {out + my, out + 2 * plane + my};
    int16_t* b[2] = {out + by, out + 2 * plane + by};
    ChromaPredictors(xs, yo + my, yo + by, best.data());
    ExpandRowXB(xs, best.data(), res, m, b);
  }
}

inline int Covered(int type) {
  return type == PIKGPU_ACS_DCT16X16 ? 2 : type == PIKGPU_ACS_DCT32X32 ? 4 : 1;
}

struct DcJob {
  const uint8_t* data;
  size_t size;
  int x0, y0, xs, ys;  // group rect in blocks
};

struct DcOut {
  int16_t* dc[3];
  size_t dc_stride;
  uint8_t* acs;
  size_t acs_stride;
  int32_t* qf;
  size_t qf_stride;
  uint8_t* lut;
  size_t lut_stride;
};

bool DecodeDcGroup(const DcJob& g, const DcOut& o, std::string* err) {
  BitStream bs(g.data, g.size);
  const size_t n = size_t(g.xs) * g.ys;
  // DC residuals (DecodeImage, pik/entropy_coder.cc:839-887, 926-939).
  std::vector<int16_t> res(3 * n, 0);
  {
    AnsTables tables;
    std::vector<uint8_t> map;
    if (!ReadHistograms(&bs, kDcContexts, kDcAlphabet, &tables, &map, err)) {
      return false;
    }
    AnsReader ans(&tables);
    for (int c = 0; c < 3; ++c) {
      const size_t h = map[c];
      int16_t* p = res.data() + c * n;
      uint32_t run = 0;
      for (size_t i = 0; i < n; ++i) {
        if (run > 0) {  // already zero
          --run;
          continue;
        }
        const int token = ans.Read(h, &bs);
        if (token >= kRunTokens) {
          run = ReadHybrid(token - kRunTokens, &bs);
          continue;
        }
        p[i] = int16_t(UnpackSigned(ReadHybrid(token, &bs)));
      }
      if (run != 0) {
        *err = "DC zero run past the channel";
        return false;
      }
    }
    if (!bs.JumpToByteBoundary()) {
      *err = "non-zero padding after the DC";
      return false;
    }
    if (!ans.FinalStateOk()) {
      *err = "DC: ANS final state mismatch";
      return false;
    }
  }
  std::vector<int16_t> dc;
  ExpandDc(g.xs, g.ys, res, &dc);

  // Control fields under one ANS state.
  AnsTables tables;
  std::vector<uint8_t> map;
  if (!ReadHistograms(&bs, kFieldContexts, 256, &tables, &map, err)) {
    return false;
  }
  AnsReader ans(&tables);
  // AC strategy (DecodeAcStrategy, pik/entropy_coder.cc:1045-1108): raw
  // types in raster order of the blocks not yet covered, then the combined
  // bytes as AcStrategyImage::SetFromRaw builds them (pik/ac_strategy.cc:
  // 1066-1099).
  std::vector<uint8_t> raw(n, kUnset), acs(n, kUnset);
  const size_t hs = map[kStrategyContext];
  for (int y = 0; y < g.ys; ++y) {
    for (int x = 0; x < g.xs; ++x) {
      if (raw[size_t(y) * g.xs + x] != kUnset) continue;
      const int t = ans.Read(hs, &bs);
      if (t > PIKGPU_ACS_DCT4X4_NOHF) {
        *err = "invalid AC strategy";
        return false;
      }
      const int s = Covered(t);
      if (y + s > g.ys || x + s > g.xs) {
        *err = "AC strategy overflows the DC group";
        return false;
      }
      for (int iy = 0; iy < s; ++iy) {
        for (int ix = 0; ix < s; ++ix) {
          raw[size_t(y + iy) * g.xs + x + ix] = uint8_t(t);
        }
      }
    }
  }
  for (int y = 0; y < g.ys; ++y) {
    for (int x = 0; x < g.xs; ++x) {
      if (acs[size_t(y) * g.xs + x] != kUnset) continue;
      const int t = raw[size_t(y) * g.xs + x];
      const int s = Covered(t);
      // Only reachable with overlapping transforms in the stream; the
      // reference writes past the group there.
      if (y + s > g.ys || x + s > g.xs) {
        *err = "inconsistent AC strategy";
        return false;
      }
      for (int iy = 0; iy < s; ++iy) {
        for (int ix = 0; ix < s; ++ix) {
          acs[size_t(y + iy) * g.xs + x + ix] = uint8_t(t << 4 | (iy * s + ix));
        }
      }
    }
  }
  // Quant field (DecodeQuantField, pik/entropy_coder.cc:944-1010), per
  // transform: 255 escapes to an absolute value - 1, any other symbol is
  // the signed difference to the mean of the values above and to the left
  // (either alone at an edge, 32 at the corner).
  std::vector<int32_t> qf(n, 0);
  std::vector<uint8_t> lut(n, 0);
  const size_t hq = map[kQuantContext], ha = map[kArContext];
  for (int y = 0; y < g.ys; ++y) {
    int32_t* row = qf.data() + size_t(y) * g.xs;
    const int32_t* top = y > 0 ? row - g.xs : nullptr;
    for (int x = 0; x < g.xs; ++x) {
      const uint8_t a = acs[size_t(y) * g.xs + x];
      if (a & 15) continue;
      int32_t pred;
      if (x == 0) pred = top ? top[0] : 32;
      else if (!top) pred = row[x - 1];
      else pred = (top[x] + row[x - 1] + 1) / 2;
      const int q = ans.Read(hq, &bs);
      const int32_t v = q == 255 ? ans.Read(hq, &bs) + 1
                                 : UnpackSigned(uint32_t(q)) + pred;
      const int s = Covered(a >> 4);
      for (int iy = 0; iy < s; ++iy) {
        for (int ix = 0; ix < s; ++ix) row[size_t(iy) * g.xs + x + ix] = v;
      }
    }
  }
  // AR sigma LUT ids (DecodeARParameters, pik/entropy_coder.cc:1125-1147),
  // one per transform.
  for (int y = 0; y < g.ys; ++y) {
    for (int x = 0; x < g.xs; ++x) {
      const uint8_t a = acs[size_t(y) * g.xs + x];
      if (a & 15) continue;
      const uint8_t v = uint8_t(ans.Read(ha, &bs));
      const int s = Covered(a >> 4);
      for (int iy = 0; iy < s; ++iy) {
        for (int ix = 0; ix < s; ++ix) {
          lut[size_t(y + iy) * g.xs + x + ix] = v;
        }
      }
    }
  }
  if (!ans.FinalStateOk()) {
    *err = "control fields: ANS final state mismatch";
    return false;
  }
  if (!bs.JumpToByteBoundary()) {
    *err = "non-zero padding after the control fields";
    return false;
  }
  for (int y = 0; y < g.ys; ++y) {
    const size_t fy = size_t(g.y0 + y), i = size_t(y) * g.xs;
    for (int c = 0; c < 3; ++c) {
      std::copy(dc.begin() + c * n + i, dc.begin() + c * n + i + g.xs,
                o.dc[c] + fy * o.dc_stride + g.x0);
    }
    std::copy(acs.begin() + i, acs.begin() + i + g.xs,
              o.acs + fy * o.acs_stride + g.x0);
    std::copy(qf.begin() + i, qf.begin() + i + g.xs,
              o.qf + fy * o.qf_stride + g.x0);
    std::copy(lut.begin() + i, lut.begin() + i + g.xs,
              o.lut + fy * o.lut_stride + g.x0);
  }
  return true;
}

void SetMsg(char* msg, size_t n, const char* fmt, ...) {
  if (msg == nullptr || n == 0) return;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(msg, n, fmt, ap);
  va_end(ap);
}

}  // namespace
}  // namespace ent
}  // namespace pikgpu

namespace pikgpu {
namespace ent {

int DecodeDcSection(pikgpu_dc_section* s, size_t skip_bits, int threads,
                    char* msg, size_t msg_size) {
  if (msg != nullptr && msg_size > 0) msg[0] = '\0';
  if (s == nullptr || (s->data == nullptr && s->size != 0) ||
      s->dc[0] == nullptr || s->dc[1] == nullptr || s->dc[2] == nullptr ||
      s->ac_strategy == nullptr || s->quant_field == nullptr ||
      s->sigma_lut_ids == nullptr) {
    SetMsg(msg, msg_size, "null argument");
    return PIKGPU_EINVAL;
  }
  if (s->xsize == 0 || s->ysize == 0) {
    SetMsg(msg, msg_size, "empty frame");
    return PIKGPU_EINVAL;
  }
  const int bx = int((s->xsize + 7) / 8), by = int((s->ysize + 7) / 8);
  if (s->dc_stride < size_t(bx) || s->ac_strategy_stride < size_t(bx) ||
      s->quant_field_stride < size_t(bx) || s->sigma_lut_stride < size_t(bx)) {
    SetMsg(msg, msg_size, "stride smaller than a row");
    return PIKGPU_EINVAL;
  }
  s->consumed = 0;
  const int gxs = (bx + kDcGroupBlocks - 1) / kDcGroupBlocks;
  const int gys = (by + kDcGroupBlocks - 1) / kDcGroupBlocks;
  const size_t ng = size_t(gxs) * gys;
  // Size table: per group a 2-bit selector of 12, 14, 15 or 21 bits
  // (SizeCoderT<0x150F0E0C>, pik/size_coder.h:17-33, U32Coder,
  // pik/fields.h:100-115).
  BitStream bs(s->data, s->size);
  bs.SkipBits(skip_bits);
  static const unsigned kSizeBits[4] = {12, 14, 15, 21};
  std::vector<size_t> off(ng + 1, 0);
  for (size_t i = 0; i < ng; ++i) {
    off[i + 1] = off[i] + bs.Read(kSizeBits[bs.Read(2)]);
  }
  if (!bs.JumpToByteBoundary()) {
    SetMsg(msg, msg_size, "non-zero padding after the DC group sizes");
    return PIKGPU_EDATA;
  }
  const size_t begin = bs.Position();
  if (begin + off[ng] > s->size) {
    SetMsg(msg, msg_size, "DC groups extend past the section (%zu > %zu)",
           begin + off[ng], s->size);
    return PIKGPU_EDATA;
  }
  std::vector<DcJob> jobs(ng);
  for (int gy = 0; gy < gys; ++gy) {
    for (int gx = 0; gx < gxs; ++gx) {
      const size_t i = size_t(gy) * gxs + gx;
      DcJob& j = jobs[i];
      j.data = s->data + begin + off[i];
      j.size = off[i + 1] - off[i];
      j.x0 = gx * kDcGroupBlocks;
      j.y0 = gy * kDcGroupBlocks;
      j.xs = std::min(kDcGroupBlocks, bx - j.x0);
      j.ys = std::min(kDcGroupBlocks, by - j.y0);
    }
  }
  const DcOut o = {{s->dc[0], s->dc[1], s->dc[2]}, s->dc_stride,
                   s->ac_strategy, s->ac_strategy_stride, s->quant_field,
                   s->quant_field_stride, s->sigma_lut_ids,
                   s->sigma_lut_stride};
  int nt = threads > 0 ? threads
                       : int(std::max(1u, std::thread::hardware_concurrency()));
  nt = std::min<int>(nt, int(ng));
  std::atomic<size_t> next{0};
  std::mutex lock;
  size_t bad = ng;
  std::string bad_err;
  auto work = [&]() {
    std::string err;
    for (;;) {
      const size_t i = next.fetch_add(1);
      if (i >= ng) return;
      if (!DecodeDcGroup(jobs[i], o, &err)) {
        std::lock_guard<std::mutex> l(lock);
        if (i < bad) {
          bad = i;
          bad_err = err;
        }
      }
    }
  };
  if (nt <= 1) {
    work();
  } else {
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t) pool.emplace_back(work);
    for (std::thread& t : pool) t.join();
  }
  if (bad < ng) {
    SetMsg(msg, msg_size, "DC group %zu: %s", bad, bad_err.c_str());
    return PIKGPU_EDATA;
  }
  s->consumed = begin + off[ng];
  return PIKGPU_OK;
}

}  // namespace ent
}  // namespace pikgpu

extern "C" PIKGPU_API int pikgpu_decode_dc_groups(pikgpu_dc_section* s,
                                                  int threads, char* msg,
                                                  size_t msg_size) {
  return pikgpu::ent::DecodeDcSection(s, 0, threads, msg, msg_size);
}
