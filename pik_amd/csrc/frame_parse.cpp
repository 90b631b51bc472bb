// Host parsing of a whole PIK frame (pikgpu_parse_frame): the FrameHeader,
// the pass-global sections (quantization tables, quantizer, colour maps, DC
// groups, dot dictionary, dequant control field and map) and, per pass, the
// group size table and the groups (GroupHeader, then the AC data). The
// result is a pikgpu_frame plus the AC of every further pass, ready for
// pikgpu_frame_upload / pikgpu_frame_add_pass.
//
// Order and formats follow PikPassToPixels (pik/pik_frame.cc:935-1187) and
// PikGroupToPixels (:846-931) with the default DecompressParams (no
// overrides, full size, every pass). The bundles are read as the
// reference's ReadVisitor reads them (pik/fields.h:67-300, 663-760,
// 853-935): U32 fields through a 4-way selector distribution, U64 through
// a 2-bit selector and varint, Bool as one raw bit, nested bundles in
// place, AllDefault as one bit that ends the bundle when set, and
// extensions as a U64 mask followed, when non-zero, by a U64 bit count
// that is skipped (no extension is known).
#include <stdarg.h>
#include <stdio.h>

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pikgpu.h"
#include "common.h"
#include "entropy.h"

struct pikgpu_parsed {
  pikgpu_frame frame;
  pikgpu_parsed_info info;
  uint32_t bx = 0, by = 0;
  std::vector<int16_t> dc;
  // passes x 3 planes; every entry is written by the AC decoding, so the
  // array is not cleared first (and is kept when a parse reuses the object)
  std::unique_ptr<int16_t[]> ac;
  size_t ac_cap = 0;
  std::vector<int32_t> qf, ytox, ytob;
  std::vector<uint8_t> acs, lut, dcf, dqmap;
  std::vector<float> table, noise;
  std::vector<uint32_t> offsets, downsampling;
  std::vector<pikgpu_dot> dots;
};

namespace pikgpu {
namespace ent {
namespace {

// Field coders (pik/field_encodings.h:19-32, pik/fields.h:100-170, 252-278).
constexpr uint32_t kU32RawBits = ~32u;
constexpr uint32_t kDirect0To3 = 0x83828180u;
constexpr uint32_t kDirect3Plus4 = 0x51828180u;
constexpr uint32_t kDirect3Plus8 = 0x52828180u;
constexpr uint32_t kDirect1248 = 0x88848281u;
constexpr uint32_t kSizeDistribution = 0x150F0E0Cu;  // group size tables

// A distribution above kU32RawBits is that many raw bits; otherwise byte
// `selector` of it is a direct value (0x80 | v), an offset code
// (0x40 | (offset - 1) << 3 | (bits - 1)) or a plain bit count.
uint32_t ReadU32(BitStream* bs, uint32_t dist) {
  if (dist > kU32RawBits) return bs->Read(dist - kU32RawBits);
  const uint32_t b = (dist >> (8 * bs->Read(2))) & 0xFF;
  if (b & 0x80) return b & 0x7F;
  if (b & 0x40) return bs->Read((b & 7) + 1) + ((b >> 3) & 7) + 1;
  return bs->Read(b);
}

// 0; 1 + 4 bits; 17 + 8 bits; or 12 bits followed by 8-bit groups while a
// continuation bit is set (the group at bit 60 has 4 bits).
uint64_t ReadU64(BitStream* bs) {
  const uint32_t sel = bs->Read(2);
  if (sel == 0) return 0;
  if (sel == 1) return 1 + bs->Read(4);
  if (sel == 2) return 17 + bs->Read(8);
  uint64_t v = bs->Read(12);
  for (unsigned shift = 12; bs->Read(1);) {
    if (shift == 60) {
      v |= uint64_t(bs->Read(4)) << shift;
      break;
    }
    v |= uint64_t(bs->Read(8)) << shift;
    shift += 8;
  }
  return v;
}

bool ReadBool(BitStream* bs) { return bs->Read(1) != 0; }

// A bundle's extensions (BeginExtensions / EndExtensions, pik/fields.h:
// 899-928): none is known, so all of their bits are skipped.
bool ReadExtensions(BitStream* bs, std::string* err) {
  if (ReadU64(bs) == 0) return true;
  const uint64_t bits = ReadU64(bs);
  const uint64_t at = bs->BitsRead();
  if (at + bits < at) {
    *err = "extension size overflows";
    return false;
  }
  bs->SkipBits(bits);
  return true;
}

struct FrameHeader {
  uint64_t size = 0;
  bool has_alpha = false;
  uint32_t encoding = 0;
  uint32_t flags = 0;
  uint32_t gaborish = 2;  // GaborishStrength::k750
  bool predict_lf = true, predict_hf = true;
  bool adaptive_reconstruction = false;
  bool epf_adaptive = true, epf_sharpened = false;
  uint32_t epf_sigma = 0;
  uint32_t num_passes = 1;
  std::vector<uint32_t> downsampling;
};

// FrameHeader::VisitFields (pik/headers.h:201-251) with FrameInfo (:149-178)
// and EpfParams (pik/epf.h:19-40) nested. nonserialized_num_groups is never
// set by the decoder, so the group size list it sizes is empty.
bool ReadFrameHeader(BitStream* bs, FrameHeader* h, std::string* err) {
  h->size = ReadU64(bs);
  h->has_alpha = ReadBool(bs);
  if (!ReadBool(bs)) {           // FrameInfo::all_default
    ReadU32(bs, 0x20088180u);    // duration
    if (ReadBool(bs)) ReadU32(bs, kU32RawBits + 32);  // timecode
    ReadBool(bs);                // is_keyframe
  }
  h->encoding = ReadU32(bs, kDirect3Plus4);
  if (h->encoding == 0) {        // ImageEncoding::kPasses
    h->flags = ReadU32(bs, 0x20181008u);
    h->gaborish = ReadU32(bs, kDirect3Plus4);
    h->predict_lf = ReadBool(bs);
    h->predict_hf = ReadBool(bs);
    h->adaptive_reconstruction = ReadBool(bs);
    if (h->adaptive_reconstruction) {
      h->epf_adaptive = ReadBool(bs);
      if (!h->epf_adaptive) h->epf_sigma = ReadU32(bs, 0x0A090880u);
      h->epf_sharpened = ReadBool(bs);
    }
    h->num_passes = ReadU32(bs, 0x20181008u);
    const uint32_t n = ReadU32(bs, kDirect0To3);
    for (uint32_t i = 0; i < n; ++i) {
      h->downsampling.push_back(ReadU32(bs, kDirect1248));
      h->downsampling.push_back(ReadU32(bs, kDirect3Plus8));
    }
  }
  if (h->encoding == 2) {        // kLossless: grayscale, 16 bits
    ReadBool(bs);
    ReadBool(bs);
  }
  return ReadExtensions(bs, err);
}

// GroupHeader::VisitFields (pik/headers.h:106-136) without alpha: 64
// TileHeaders (:52-76), each possibly with 8 raw bytes of projective
// transform parameters (:36-50) that the decoder does not use.
bool ReadGroupHeader(BitStream* bs, std::string* err) {
  if (ReadBool(bs)) return true;  // all_default
  for (int t = 0; t < 64; ++t) {
    if (ReadBool(bs)) continue;   // TileHeader::all_default
    if (ReadBool(bs)) {
      for (int i = 0; i < 8; ++i) bs->Read(8);
    }
    if (!ReadExtensions(bs, err)) return false;
  }
  return ReadExtensions(bs, err);
}

inline int32_t UnpackSigned(uint32_t u) {
  return int32_t((u >> 1) ^ (((~u) & 1) - 1));
}

// The dot dictionary's one static patch (pik/block_dictionary.cc:26-29).
constexpr int kStaticXsize = 1, kStaticYsize = 3;
constexpr int8_t kStaticPixels[3][3] = {{0, 0, 0}, {-11, -8, -11},
                                        {-10, -7, -10}};
constexpr size_t kMaxDictionary = size_t(1) << 20;
constexpr size_t kMaxPlacements = size_t(1) << 22;

struct Patch {
  size_t xsize, ysize;
  int8_t px[3][64];
};

// BlockDictionary::Decode (pik/block_dictionary.cc:199-307): nine contexts
// of an alphabet of 32 under one ANS state; every number is a token s and
// s raw bits, (1 << s) + bits - 1. Sizes and positions after the first of
// an id are signed deltas; patch pixels are signed residuals against the
// int8 mean of the left and upper neighbours.
int ReadDots(BitStream* bs, size_t padded_x, size_t padded_y,
             std::vector<pikgpu_dot>* dots, std::string* err) {
  AnsTables tables;
  std::vector<uint8_t> map;
  if (!ReadHistograms(bs, 9, 32, &tables, &map, err)) return PIKGPU_EDATA;
  AnsReader ans(&tables);
  auto num = [&](int ctx) -> uint32_t {
    const int s = ans.Read(map[ctx], bs);
    const uint32_t bits = bs->Read(unsigned(s));
    return (1u << s) + bits - 1;
  };
  const size_t dict_size = num(0);
  if (dict_size > (size_t(1) << 24)) {
    *err = "too many blocks in the dot dictionary";
    return PIKGPU_EDATA;
  }
  if (dict_size > kMaxDictionary) {
    *err = "dot dictionary larger than 2^20 patches";
    return PIKGPU_EUNSUPPORTED;
  }
  std::vector<Patch> dict(dict_size);
  for (size_t i = 0; i < dict_size; ++i) {
    Patch& p = dict[i];
    p.xsize = num(1);
    p.ysize = num(1);
    if (i != 0) {
      // size_t arithmetic as the reference's: a negative size wraps to a
      // huge one and fails the check below
      p.xsize = size_t(int64_t(UnpackSigned(uint32_t(p.xsize)))) +
                dict[i - 1].xsize;
      p.ysize = size_t(int64_t(UnpackSigned(uint32_t(p.ysize)))) +
                dict[i - 1].ysize;
    }
    if (p.xsize > 8 || p.ysize > 8) {
      *err = "dot patch larger than 8x8";
      return PIKGPU_EDATA;
    }
    for (int c = 0; c < 3; ++c) {
      for (size_t iy = 0; iy < p.ysize; ++iy) {
        for (size_t ix = 0; ix < p.xsize; ++ix) {
          int8_t pred = 0;
          if (ix != 0) pred = int8_t(pred + p.px[c][iy * p.xsize + ix - 1]);
          if (iy != 0) pred = int8_t(pred + p.px[c][(iy - 1) * p.xsize + ix]);
          if (ix != 0 && iy != 0) pred = int8_t(pred >> 1);
          p.px[c][iy * p.xsize + ix] =
              int8_t(UnpackSigned(num(2 + c)) + pred);
        }
      }
    }
  }
  size_t placed = 0;
  for (int transform = 0; transform < 2; ++transform) {
    size_t to_skip = 0;
    for (size_t id = 0; id < dict_size + 1; ++id) {
      if (to_skip > 0) {
        --to_skip;
        continue;
      }
      const size_t count = num(7);
      if (count > (size_t(1) << 24)) {
        *err = "too many dot placements";
        return PIKGPU_EDATA;
      }
      if (count == 0) {
        to_skip = num(8);
        continue;
      }
      placed += count;
      if (placed > kMaxPlacements) {
        *err = "more than 2^22 dot placements";
        return PIKGPU_EUNSUPPORTED;
      }
      const bool is_static = id == dict_size;
      const size_t pxs = is_static ? kStaticXsize : dict[id].xsize;
      const size_t pys = is_static ? kStaticYsize : dict[id].ysize;
      size_t x = 0, y = 0;
      for (size_t k = 0; k < count; ++k) {
        size_t nx = num(5), ny = num(5);
        if (k != 0) {
          nx = size_t(int64_t(UnpackSigned(uint32_t(nx)))) + x;
          ny = size_t(int64_t(UnpackSigned(uint32_t(ny)))) + y;
        }
        x = nx;
        y = ny;
        // The reference checks x + xsize > padded size only; a position
        // that wrapped below zero would pass it and write outside its
        // image, so it is refused here as well.
        if (x + pxs > padded_x || x >= padded_x) {
          *err = "dot outside the image (x)";
          return PIKGPU_EDATA;
        }
        if (y + pys > padded_y || y >= padded_y) {
          *err = "dot outside the image (y)";
          return PIKGPU_EDATA;
        }
        pikgpu_dot d;
        memset(&d, 0, sizeof(d));
        d.x = uint32_t(x);
        d.y = uint32_t(y);
        d.xsize = uint32_t(pxs);
        d.ysize = uint32_t(pys);
        d.transform = transform;
        if (transform) {
          d.dx = UnpackSigned(num(5));
          d.dy = UnpackSigned(num(5));
          d.width = UnpackSigned(num(6));
          // The reference divides by the patch's length and width and
          // clamps source indices into the patch: zero of either is
          // undefined there.
          if (pxs == 0 || pys == 0 || (d.dx == 0 && d.dy == 0) ||
              d.width <= 0) {
            *err = "degenerate dot transform";
            return PIKGPU_EDATA;
          }
        }
        for (int c = 0; c < 3; ++c) {
          for (size_t i = 0; i < pxs * pys; ++i) {
            d.pixels[c][i] = is_static ? kStaticPixels[c][i] : dict[id].px[c][i];
          }
        }
        // an empty patch placed as-is changes nothing
        if (pxs != 0 && pys != 0) dots->push_back(d);
      }
    }
    if (to_skip > 0) {
      *err = "dot dictionary skips past its last patch";
      return PIKGPU_EDATA;
    }
  }
  if (!ans.FinalStateOk()) {
    *err = "dot dictionary: ANS final state mismatch";
    return PIKGPU_EDATA;
  }
  if (!bs->JumpToByteBoundary()) {
    *err = "non-zero padding after the dot dictionary";
    return PIKGPU_EDATA;
  }
  return PIKGPU_OK;
}

// DecodeColorMap (pik/color_correlation.cc:244-262): a prefix code, the DC
// factor, then one factor per colour tile, zero bits to the byte boundary.
// The caller ignores the result, so a code that does not parse leaves the
// map and factor untouched and the reader where the code stopped.
void ReadColorMap(BitStream* bs, int32_t* map, size_t n, int32_t* dc) {
  PrefixCode code;
  std::string ignored;
  if (!code.Read(bs, &ignored)) return;
  *dc = code.Decode(bs);
  for (size_t i = 0; i < n; ++i) map[i] = code.Decode(bs);
  bs->JumpToByteBoundary();
}

void SetMsg(char* msg, size_t n, const char* fmt, ...) {
  if (msg == nullptr || n == 0) return;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(msg, n, fmt, ap);
  va_end(ap);
}

int Parse(const uint8_t* data, size_t size, uint32_t xsize, uint32_t ysize,
          int threads, pikgpu_parsed* p, char* msg, size_t msg_size) {
  std::string err;
  BitStream bs(data, size);
  FrameHeader h;
  if (!ReadFrameHeader(&bs, &h, &err)) {
    SetMsg(msg, msg_size, "frame header: %s", err.c_str());
    return PIKGPU_EDATA;
  }
  if (!bs.JumpToByteBoundary()) {
    SetMsg(msg, msg_size, "non-zero padding after the frame header");
    return PIKGPU_EDATA;
  }
  if (h.encoding == 2) {
    SetMsg(msg, msg_size, "lossless frames are not supported");
    return PIKGPU_EUNSUPPORTED;
  }
  if (h.encoding != 0) {
    SetMsg(msg, msg_size, "unsupported frame encoding %u", h.encoding);
    return PIKGPU_EDATA;
  }
  if (h.has_alpha) {
    SetMsg(msg, msg_size, "alpha is not supported");
    return PIKGPU_EUNSUPPORTED;
  }
  if (h.flags & PIKGPU_FLAG_GRADIENT_MAP) {
    SetMsg(msg, msg_size, "the gradient map (FrameHeader::kGradientMap) is "
                          "not supported");
    return PIKGPU_EUNSUPPORTED;
  }
  if (h.gaborish > 4) {
    SetMsg(msg, msg_size, "invalid Gaborish strength %u", h.gaborish);
    return PIKGPU_EDATA;
  }
  if (h.num_passes == 0) {
    SetMsg(msg, msg_size, "frame without passes");
    return PIKGPU_EDATA;
  }
  if (h.num_passes > 16) {
    SetMsg(msg, msg_size, "%u passes (at most 16 are supported)",
           h.num_passes);
    return PIKGPU_EUNSUPPORTED;
  }
  const uint32_t bx = (xsize + 7) / 8, by = (ysize + 7) / 8;
  const uint32_t tx = (xsize + 63) / 64, ty = (ysize + 63) / 64;
  const size_t nb = size_t(bx) * by;
  p->bx = bx;
  p->by = by;

  // DequantMatrices::Decode: the number of encoded tables, then per table
  // its mode byte. Library tables (the only predefined one) and copies of
  // earlier tables resolve to the library tables; every other mode carries
  // custom weights.
  const uint32_t num_tables = bs.Read(8);
  const size_t full = std::max<size_t>(1, (num_tables + 6) / 7);
  for (uint32_t i = 0; i < num_tables; ++i) {
    const uint32_t mode = bs.Read(8);
    if (mode == 0) {
      if (bs.Read(8) >= 1) {
        SetMsg(msg, msg_size, "quantization table %u: invalid predefined "
                              "table", i);
        return PIKGPU_EDATA;
      }
    } else if (mode == 7) {
      if (bs.Read(8) >= i) {
        SetMsg(msg, msg_size, "quantization table %u: invalid source", i);
        return PIKGPU_EDATA;
      }
    } else if (mode <= 6) {
      SetMsg(msg, msg_size, "custom quantization tables (mode %u) are not "
                            "supported", mode);
      return PIKGPU_EUNSUPPORTED;
    } else {
      SetMsg(msg, msg_size, "quantization table %u: invalid mode %u", i,
             mode);
      return PIKGPU_EDATA;
    }
    if (bs.Position() > size) {
      SetMsg(msg, msg_size, "quantization tables past the end");
      return PIKGPU_EDATA;
    }
  }
  {
    size_t noff = 0;
    const size_t n = DefaultDequantTables(nullptr, 0, nullptr, 0, &noff);
    std::vector<float> one(n);
    std::vector<uint32_t> off(noff);
    DefaultDequantTables(one.data(), n, off.data(), noff, &noff);
    p->table.resize(full * n);
    p->offsets.resize(full * noff);
    for (size_t t = 0; t < full; ++t) {
      std::copy(one.begin(), one.end(), p->table.begin() + t * n);
      for (size_t k = 0; k < noff; ++k) {
        p->offsets[t * noff + k] = uint32_t(off[k] + t * n);
      }
    }
  }
  // Quantizer::Decode: global_scale - 1 (15 of 16 bits) and quant_dc - 1.
  const uint32_t gs_hi = bs.Read(8), gs_lo = bs.Read(8);
  const int32_t global_scale = int32_t(((gs_hi << 8 | gs_lo) & 0x7FFF) + 1);
  const int32_t quant_dc = int32_t(bs.Read(8)) + 1;
  if (!bs.JumpToByteBoundary()) {
    SetMsg(msg, msg_size, "non-zero padding after the quantizer");
    return PIKGPU_EDATA;
  }
  // Colour maps, Y-to-B first (defaults of ColorCorrelationMap,
  // pik/color_correlation.h:47-54).
  p->ytob.assign(size_t(tx) * ty, 120);
  p->ytox.assign(size_t(tx) * ty, 128);
  int32_t ytob_dc = 120, ytox_dc = 128;
  ReadColorMap(&bs, p->ytob.data(), p->ytob.size(), &ytob_dc);
  ReadColorMap(&bs, p->ytox.data(), p->ytox.size(), &ytox_dc);

  // DC groups.
  p->dc.assign(3 * nb, 0);
  p->acs.assign(nb, 0);
  p->qf.assign(nb, 0);
  p->lut.assign(nb, 0);
  // The size table is read from wherever the colour maps left the reader
  // (not byte aligned when a map's code failed to parse).
  size_t pos = 0;
  {
    pikgpu_dc_section s;
    memset(&s, 0, sizeof(s));
    s.xsize = xsize;
    s.ysize = ysize;
    s.data = data;
    s.size = size;
    for (int c = 0; c < 3; ++c) s.dc[c] = p->dc.data() + c * nb;
    s.dc_stride = bx;
    s.ac_strategy = p->acs.data();
    s.ac_strategy_stride = bx;
    s.quant_field = p->qf.data();
    s.quant_field_stride = bx;
    s.sigma_lut_ids = p->lut.data();
    s.sigma_lut_stride = bx;
    char m[200];
    const int rc = DecodeDcSection(&s, bs.BitsRead(), threads, m, sizeof(m));
    if (rc != PIKGPU_OK) {
      SetMsg(msg, msg_size, "%s", m);
      return rc;
    }
    pos = s.consumed;
  }
  // The rest of the pass-global data continues at the end of the DC groups.
  BitStream rest(data + std::min(pos, size), size - std::min(pos, size));
  p->dots.clear();
  int rc = ReadDots(&rest, size_t(bx) * 8, size_t(by) * 8, &p->dots, &err);
  if (rc != PIKGPU_OK) {
    SetMsg(msg, msg_size, "%s", err.c_str());
    return rc;
  }
  // DecodeDequantControlField: one uint8 per colour tile under a prefix
  // code, byte aligned.
  p->dcf.assign(size_t(tx) * ty, 0);
  {
    PrefixCode code;
    if (!code.Read(&rest, &err)) {
      SetMsg(msg, msg_size, "dequant control field: %s", err.c_str());
      return PIKGPU_EDATA;
    }
    for (size_t i = 0; i < p->dcf.size(); ++i) {
      p->dcf[i] = uint8_t(code.Decode(&rest));
    }
    if (!rest.JumpToByteBoundary()) {
      SetMsg(msg, msg_size, "non-zero padding after the dequant control "
                            "field");
      return PIKGPU_EDATA;
    }
    for (uint8_t v : p->dcf) {
      if (v >= 16) {
        SetMsg(msg, msg_size, "dequant control field value %u >= 16", v);
        return PIKGPU_EDATA;
      }
    }
  }
  // DecodeDequantControlFieldMap: the (control value, quant - 1) pairs the
  // frame uses, in row-major order of that 16 x 256 table, under a prefix
  // code.
  p->dqmap.assign(16 * 256, 0);
  {
    std::vector<uint8_t> mask(16 * 256, 0);
    for (uint32_t y = 0; y < by; ++y) {
      for (uint32_t x = 0; x < bx; ++x) {
        const int32_t q = p->qf[size_t(y) * bx + x];
        if (q < 1 || q > 256) {
          SetMsg(msg, msg_size, "quant field value %d outside [1, 256]", q);
          return PIKGPU_EDATA;
        }
        mask[p->dcf[(y / 8) * tx + x / 8] * 256 + (q - 1)] = 1;
      }
    }
    PrefixCode code;
    if (!code.Read(&rest, &err)) {
      SetMsg(msg, msg_size, "dequant control field map: %s", err.c_str());
      return PIKGPU_EDATA;
    }
    for (size_t i = 0; i < mask.size(); ++i) {
      if (mask[i]) p->dqmap[i] = uint8_t(code.Decode(&rest));
    }
    if (!rest.JumpToByteBoundary()) {
      SetMsg(msg, msg_size, "non-zero padding after the dequant control "
                            "field map");
      return PIKGPU_EDATA;
    }
    for (uint8_t v : p->dqmap) {
      if (v >= full) {
        SetMsg(msg, msg_size, "dequant map names table %u of %zu", v, full);
        return PIKGPU_EDATA;
      }
    }
  }

  // Group size tables and groups, pass by pass.
  const uint32_t gxs = (xsize + 511) / 512, gys = (ysize + 511) / 512;
  const size_t ng = size_t(gxs) * gys;
  std::vector<size_t> begin(h.num_passes);
  std::vector<std::vector<size_t>> off(h.num_passes);
  for (uint32_t i = 0; i < h.num_passes; ++i) {
    off[i].assign(ng + 1, 0);
    for (size_t g = 0; g < ng; ++g) {
      off[i][g + 1] = off[i][g] + ReadU32(&rest, kSizeDistribution);
    }
    if (!rest.JumpToByteBoundary()) {
      SetMsg(msg, msg_size, "pass %u: non-zero padding after the group "
                            "sizes", i);
      return PIKGPU_EDATA;
    }
    begin[i] = pos + rest.Position();
    rest.SkipBits(uint64_t(off[i][ng]) * 8);
    if (pos + rest.Position() > size) {
      SetMsg(msg, msg_size, "pass %u: groups extend past the end (%zu > %zu)",
             i, pos + rest.Position(), size);
      return PIKGPU_EDATA;
    }
  }
  const size_t consumed = pos + rest.Position();

  const size_t ac_need = size_t(h.num_passes) * 3 * nb * 64;
  if (p->ac_cap < ac_need) {
    p->ac.reset(new int16_t[ac_need]);
    p->ac_cap = ac_need;
  }
  p->noise.assign(size_t(h.num_passes) * ng * 3, 0.0f);
  std::vector<const uint8_t*> gdata(ng);
  std::vector<size_t> gsize(ng);
  for (uint32_t i = 0; i < h.num_passes; ++i) {
    for (size_t g = 0; g < ng; ++g) {
      const uint8_t* d = data + begin[i] + off[i][g];
      size_t n = off[i][g + 1] - off[i][g];
      if (i == 0) {
        // The group header sits in front of the first pass's data.
        BitStream gb(d, n);
        if (!ReadGroupHeader(&gb, &err)) {
          SetMsg(msg, msg_size, "group %zu header: %s", g, err.c_str());
          return PIKGPU_EDATA;
        }
        if (!gb.JumpToByteBoundary()) {
          SetMsg(msg, msg_size, "group %zu: non-zero padding after the "
                                "group header", g);
          return PIKGPU_EDATA;
        }
        const size_t hdr = std::min(gb.Position(), n);
        d += hdr;
        n -= hdr;
      }
      gdata[g] = d;
      gsize[g] = n;
    }
    pikgpu_ac_pass a;
    memset(&a, 0, sizeof(a));
    a.xsize = xsize;
    a.ysize = ysize;
    a.ac_strategy = p->acs.data();
    a.ac_strategy_stride = bx;
    a.num_groups = ng;
    a.group_data = gdata.data();
    a.group_size = gsize.data();
    int16_t* base = p->ac.get() + size_t(i) * 3 * nb * 64;
    for (int c = 0; c < 3; ++c) a.ac[c] = base + c * nb * 64;
    a.ac_stride = size_t(bx) * 64;
    a.noise = p->noise.data() + size_t(i) * ng * 3;
    char m[200];
    rc = pikgpu_decode_ac_pass(&a, threads, m, sizeof(m));
    if (rc != PIKGPU_OK) {
      SetMsg(msg, msg_size, "pass %u: %s", i, m);
      return rc;
    }
  }

  p->downsampling = h.downsampling;
  pikgpu_frame& f = p->frame;
  memset(&f, 0, sizeof(f));
  f.xsize = xsize;
  f.ysize = ysize;
  f.gaborish = int32_t(h.gaborish);
  f.predict_lf = h.predict_lf;
  f.predict_hf = h.predict_hf;
  f.adaptive_reconstruction = h.adaptive_reconstruction;
  f.epf_enable_adaptive = h.epf_adaptive;
  f.epf_use_sharpened = h.epf_sharpened;
  f.global_scale = global_scale;
  f.quant_dc = quant_dc;
  for (int c = 0; c < 3; ++c) {
    f.ac[c] = p->ac.get() + c * nb * 64;
    f.dc[c] = p->dc.data() + c * nb;
  }
  f.ac_stride = size_t(bx) * 64;
  f.dc_stride = bx;
  f.quant_field = p->qf.data();
  f.quant_field_stride = bx;
  f.ac_strategy = p->acs.data();
  f.ac_strategy_stride = bx;
  f.sigma_lut_ids = p->lut.data();
  f.sigma_lut_stride = bx;
  f.dequant_cf = p->dcf.data();
  f.dequant_cf_stride = tx;
  f.ytox_map = p->ytox.data();
  f.ytox_stride = tx;
  f.ytob_map = p->ytob.data();
  f.ytob_stride = tx;
  f.ytox_dc = ytox_dc;
  f.ytob_dc = ytob_dc;
  f.dequant_map = p->dqmap.data();
  f.dequant_table = p->table.data();
  f.dequant_table_size = p->table.size();
  f.dequant_offsets = p->offsets.data();
  f.num_dequant_offsets = p->offsets.size();
  f.dots = p->dots.empty() ? nullptr : p->dots.data();
  f.num_dots = p->dots.size();
  f.grayscale = (h.flags & PIKGPU_FLAG_GRAYSCALE_OPT) ? 1 : 0;
  f.flags = h.flags;

  pikgpu_parsed_info& in = p->info;
  memset(&in, 0, sizeof(in));
  in.frame_size = h.size;
  in.num_passes = h.num_passes;
  in.num_groups = uint32_t(ng);
  in.epf_sigma = h.epf_sigma;
  in.num_downsampling = uint32_t(p->downsampling.size() / 2);
  in.downsampling = p->downsampling.empty() ? nullptr : p->downsampling.data();
  in.noise = p->noise.data();
  in.consumed = consumed;
  return PIKGPU_OK;
}

}  // namespace
}  // namespace ent
}  // namespace pikgpu

extern "C" PIKGPU_API int pikgpu_parse_frame(const uint8_t* data, size_t size,
                                             uint32_t xsize, uint32_t ysize,
                                             int threads, pikgpu_parsed** out,
                                             char* msg, size_t msg_size) {
  using namespace pikgpu::ent;
  if (msg != nullptr && msg_size > 0) msg[0] = '\0';
  if (out == nullptr || (data == nullptr && size != 0)) {
    SetMsg(msg, msg_size, "null argument");
    return PIKGPU_EINVAL;
  }
  if (xsize == 0 || ysize == 0) {
    SetMsg(msg, msg_size, "empty image");
    return PIKGPU_EINVAL;
  }
  static const uint8_t kNone[1] = {0};
  // A parsed frame passed in is reused: its storage (above all the AC
  // planes, which need not be cleared) is kept where it is large enough.
  pikgpu_parsed* p = *out != nullptr ? *out : new pikgpu_parsed;
  const int rc = Parse(data != nullptr ? data : kNone, size, xsize, ysize,
                       threads, p, msg, msg_size);
  if (rc != PIKGPU_OK) {
    if (*out == nullptr) delete p;
    return rc;
  }
  *out = p;
  return PIKGPU_OK;
}

extern "C" PIKGPU_API int pikgpu_parsed_get(const pikgpu_parsed* p,
                                            pikgpu_frame* frame,
                                            pikgpu_parsed_info* info) {
  if (p == nullptr) return PIKGPU_EINVAL;
  if (frame != nullptr) *frame = p->frame;
  if (info != nullptr) *info = p->info;
  return PIKGPU_OK;
}

extern "C" PIKGPU_API int pikgpu_parsed_pass_ac(const pikgpu_parsed* p,
                                                uint32_t pass,
                                                const int16_t* ac[3],
                                                size_t* ac_stride) {
  if (p == nullptr || ac == nullptr || pass >= p->info.num_passes) {
    return PIKGPU_EINVAL;
  }
  const size_t nb = size_t(p->bx) * p->by;
  const int16_t* base = p->ac.get() + size_t(pass) * 3 * nb * 64;
  for (int c = 0; c < 3; ++c) ac[c] = base + c * nb * 64;
  if (ac_stride != nullptr) *ac_stride = size_t(p->bx) * 64;
  return PIKGPU_OK;
}

extern "C" PIKGPU_API void pikgpu_parsed_free(pikgpu_parsed* p) { delete p; }
