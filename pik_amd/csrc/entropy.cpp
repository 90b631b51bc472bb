// Host entropy decoding of PIK's AC groups (pikgpu_decode_ac_pass): each
// group's pass bitstream -- noise parameters, three coefficient orders, the
// context map and ANS histograms, then the ANS-coded coefficients of every
// 64x64 tile -- into the frame-wide int16 AC planes the HIP reconstruction
// takes (pikgpu_frame.ac). This is DecodeFromBitstream
// (pik/compressed_image.cc:719-813) without its per-tile dequantization,
// which runs on the GPU for the whole frame.
//
// Decoding is sequential within a group (one ANS state threads through all
// its tiles) and independent across groups; groups are spread over host
// threads.
#include <stdarg.h>
#include <stdio.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <thread>

#include "entropy.h"
#include "../../include/pikgpu.h"

namespace pikgpu {
namespace ent {
namespace {

// 1-11 bits: 0, 1, or 2^n + n extra bits for n in 1..7
// (pik/context_map_decode.cc:40-50).
uint32_t ReadVarLenU8(BitStream* bs) {
  if (!bs->Read(1)) return 0;
  const uint32_t n = bs->Read(3);
  return n == 0 ? 1 : bs->Read(n) + (1u << n);
}

// 1-20 bits: 0, 1, or 2^n + n extra bits for n in 1..15
// (pik/ans_decode.cc:17-27, pik/huffman_decode.cc:170-180).
uint32_t ReadVarLenU16(BitStream* bs) {
  if (!bs->Read(1)) return 0;
  const uint32_t n = bs->Read(4);
  return n == 0 ? 1 : bs->Read(n) + (1u << n);
}

// A fixed prefix code given by its code words (symbol, length, value of the
// bits in stream order), decoded by a lookup of `bits` peeked bits.
struct CodeWord {
  uint8_t symbol, length, bits;
};

template <int kBits, int kWords>
struct FixedCode {
  explicit FixedCode(const CodeWord (&w)[kWords]) {
    for (const CodeWord& c : w) {
      for (int j = c.bits; j < (1 << kBits); j += 1 << c.length) {
        entry[j] = uint16_t(c.length << 8 | c.symbol);
      }
    }
  }
  int Decode(BitStream* bs) const {
    bs->Fill();
    const uint16_t e = entry[bs->Peek(kBits)];
    bs->Skip(e >> 8);
    return e & 0xFF;
  }
  uint16_t entry[1 << kBits] = {};
};

// Code of the code-length code lengths 0..5 (pik/huffman_decode.cc:285-288,
// RFC 7932 3.5).
const CodeWord kLengthCodeWords[6] = {
    {0, 2, 0}, {4, 2, 1}, {3, 2, 2}, {2, 3, 3}, {1, 4, 7}, {5, 4, 15}};

// Code of the ANS histogram log-counts 0..10 and 11 = run of the previous
// count (pik/ans_decode.cc:67-84).
const CodeWord kLogCountWords[12] = {
    {6, 2, 0},  {7, 3, 1},  {4, 3, 2},   {8, 3, 5},   {5, 3, 6},   {1, 4, 3},
    {3, 4, 7},  {2, 4, 11}, {0, 5, 15},  {9, 6, 31},  {10, 7, 63}, {11, 7, 127}};

// Order in which the code-length code lengths are sent
// (pik/huffman_decode.cc:17-19, RFC 7932 3.5).
const uint8_t kLengthCodeOrder[18] = {1, 2,  3, 4,  0, 5, 17, 6,  16,
                                      7, 8, 9, 10, 11, 12, 13, 14, 15};

uint32_t Reverse(uint32_t v, unsigned n) {
  uint32_t r = 0;
  for (unsigned i = 0; i < n; ++i) r |= ((v >> i) & 1) << (n - 1 - i);
  return r;
}

}  // namespace

bool PrefixCode::Build(const std::vector<uint8_t>& lengths,
                       std::string* err) {
  // Canonical code (RFC 7932 3.2): shorter codes first, increasing symbol
  // order within a length; the first bit in the stream is the code's most
  // significant one.
  uint32_t count[16] = {0};
  size_t used = 0;
  int last = -1;
  for (size_t s = 0; s < lengths.size(); ++s) {
    if (lengths[s] == 0) continue;
    ++count[lengths[s]];
    ++used;
    last = int(s);
  }
  if (used == 0) {
    *err = "empty prefix code";
    return false;
  }
  if (used == 1) {  // one symbol takes no bits
    max_len_ = 0;
    table_.assign(1, uint32_t(last));
    return true;
  }
  unsigned max_len = 0;
  for (unsigned l = 1; l < 16; ++l) {
    if (count[l]) max_len = l;
  }
  uint32_t next[16] = {0};
  uint32_t code = 0;
  for (unsigned l = 1; l < 16; ++l) {
    code = (code + count[l - 1]) << 1;
    next[l] = code;
  }
  next[0] = 0;
  max_len_ = max_len;
  table_.assign(size_t(1) << max_len, 0);
  for (size_t s = 0; s < lengths.size(); ++s) {
    const unsigned l = lengths[s];
    if (l == 0) continue;
    const uint32_t c = next[l]++;
    if (c >= (1u << l)) {
      *err = "over-subscribed prefix code";
      return false;
    }
    for (uint32_t j = Reverse(c, l); j < table_.size(); j += 1u << l) {
      table_[j] = (l << 16) | uint32_t(s);
    }
  }
  return true;
}

bool PrefixCode::Read(BitStream* bs, std::string* err) {
  std::vector<uint8_t> lengths;
  const uint32_t kind = bs->Read(2);
  if (kind == 1) {
    // Simple code: 1-4 symbols with fixed length patterns
    // (pik/huffman_decode.cc:244-283).
    uint32_t sym[4] = {0};
    const uint32_t n = bs->Read(2) + 1;
    uint32_t top = 0;
    for (uint32_t i = 0; i < n; ++i) {
      sym[i] = ReadVarLenU16(bs);
      top = std::max(top, sym[i]);
    }
    lengths.assign(top + 1, 0);
    lengths[sym[0]] = 1;
    for (uint32_t i = 1; i < n; ++i) lengths[sym[i]] = 2;
    bool distinct = true;
    for (uint32_t i = 0; i < n; ++i) {
      for (uint32_t j = i + 1; j < n; ++j) distinct &= sym[i] != sym[j];
    }
    if (n == 2) lengths[sym[1]] = 1;
    if (n == 4) {
      if (bs->Read(1)) {
        lengths[sym[2]] = 3;
        lengths[sym[3]] = 3;
      } else {
        lengths[sym[0]] = 2;
      }
    }
    if (!distinct) {
      *err = "repeated symbol in a simple prefix code";
      return false;
    }
    return Build(lengths, err);
  }
  // Complex code: the code-length code's lengths (skipping the first
  // `kind` of them), then the symbol code lengths with repeat codes 16
  // (previous non-zero length) and 17 (zeros) (pik/huffman_decode.cc:
  // 178-241, 284-310).
  static const FixedCode<4, 6> length_code(kLengthCodeWords);
  std::vector<uint8_t> cl(18, 0);
  int space = 32, nonzero = 0;
  for (uint32_t i = kind; i < 18 && space > 0; ++i) {
    const int v = length_code.Decode(bs);
    cl[kLengthCodeOrder[i]] = uint8_t(v);
    if (v != 0) {
      space -= 32 >> v;
      ++nonzero;
    }
  }
  if (!(nonzero == 1 || space == 0)) {
    *err = "invalid code-length code";
    return false;
  }
  PrefixCode clc;
  if (!clc.Build(cl, err)) return false;
  uint8_t prev = 8, rep_len = 0;
  int repeat = 0;
  space = 32768;
  const size_t kMaxSymbols = 1u << 16;
  while (lengths.size() < kMaxSymbols && space > 0) {
    const int len = clc.Decode(bs);
    if (len < 16) {
      repeat = 0;
      lengths.push_back(uint8_t(len));
      if (len != 0) {
        prev = uint8_t(len);
        space -= 32768 >> len;
      }
      continue;
    }
    const unsigned extra = unsigned(len - 14);
    const uint8_t new_len = len == 16 ? prev : 0;
    if (rep_len != new_len) {
      repeat = 0;
      rep_len = new_len;
    }
    const int old = repeat;
    if (repeat > 0) repeat = (repeat - 2) << extra;
    repeat += int(bs->Read(extra)) + 3;
    const int delta = repeat - old;
    if (lengths.size() + size_t(delta) > kMaxSymbols) {
      *err = "prefix code lengths overflow";
      return false;
    }
    lengths.insert(lengths.end(), size_t(delta), rep_len);
    if (rep_len != 0) space -= delta << (15 - rep_len);
  }
  if (space != 0) {
    *err = "prefix code lengths do not fill the code space";
    return false;
  }
  return Build(lengths, err);
}

bool ReadContextMap(BitStream* bs, size_t num_contexts,
                    std::vector<uint8_t>* map, size_t* num_histograms,
                    std::string* err) {
  map->assign(num_contexts, 0);
  *num_histograms = ReadVarLenU8(bs) + 1;
  if (*num_histograms <= 1) return true;
  uint32_t max_run_prefix = 0;
  if (bs->Read(1)) max_run_prefix = bs->Read(4) + 1;
  PrefixCode code;
  if (!code.Read(bs, err)) return false;
  for (size_t i = 0; i < num_contexts;) {
    const uint32_t v = uint32_t(code.Decode(bs));
    if (v == 0) {
      (*map)[i++] = 0;
    } else if (v <= max_run_prefix) {
      // A run of 2^v + extra zeros.
      const uint32_t run = (1u << v) + bs->Read(v);
      if (run > num_contexts - i) {
        *err = "context map zero run too long";
        return false;
      }
      i += run;  // already zero
    } else {
      (*map)[i++] = uint8_t(v - max_run_prefix);
    }
  }
  if (bs->Read(1)) {
    // Inverse move-to-front.
    uint8_t mtf[256];
    for (int i = 0; i < 256; ++i) mtf[i] = uint8_t(i);
    for (size_t i = 0; i < num_contexts; ++i) {
      const uint8_t idx = (*map)[i];
      const uint8_t v = mtf[idx];
      (*map)[i] = v;
      memmove(mtf + 1, mtf, idx);
      mtf[0] = v;
    }
  }
  // Every histogram is used and no index is out of range.
  std::vector<uint8_t> seen(*num_histograms, 0);
  size_t found = 0;
  for (uint8_t h : *map) {
    if (h >= *num_histograms) {
      *err = "context map names a missing histogram";
      return false;
    }
    found += seen[h] == 0;
    seen[h] = 1;
  }
  if (found != *num_histograms) {
    *err = "context map leaves a histogram unused";
    return false;
  }
  return true;
}

namespace {

// One histogram's counts, summing to 1024 (ReadHistogram,
// pik/ans_decode.cc:29-140).
bool ReadCounts(BitStream* bs, std::vector<int>* counts, std::string* err) {
  const int kTotal = int(kAnsTab);
  if (bs->Read(1)) {
    // One or two symbols.
    const uint32_t n = bs->Read(1) + 1;
    uint32_t sym[2] = {0, 0};
    uint32_t top = 0;
    for (uint32_t i = 0; i < n; ++i) {
      sym[i] = ReadVarLenU16(bs);
      top = std::max(top, sym[i]);
    }
    counts->assign(top + 1, 0);
    if (n == 1) {
      (*counts)[sym[0]] = kTotal;
      return true;
    }
    if (sym[0] == sym[1]) {
      *err = "two-symbol histogram repeats its symbol";
      return false;
    }
    (*counts)[sym[0]] = int(bs->Read(kAnsLogTab));
    (*counts)[sym[1]] = kTotal - (*counts)[sym[0]];
    return true;
  }
  if (bs->Read(1)) {
    // Flat over an alphabet of n symbols, the remainder on the first ones.
    const int n = int(bs->Read(kAnsLogTab));
    if (n == 0) {
      *err = "flat histogram of zero symbols";
      return false;
    }
    counts->assign(n, kTotal / n);
    for (int i = 0; i < kTotal % n; ++i) ++(*counts)[i];
    return true;
  }
  static const FixedCode<7, 12> log_code(kLogCountWords);
  const int len = int(ReadVarLenU16(bs)) + 3;
  counts->assign(len, 0);
  std::vector<int> logc(len, 0), run(len, 0);
  int omit_log = -1, omit_pos = -1;
  // Log-counts; code 11 starts a run whose 8-bit length counts the run
  // position itself plus the positions it repeats into. (The position
  // arithmetic is kept exactly as the format defines it, including for
  // lengths 0 and 1.)
  for (int i = 0; i >= 0 && i < len; ++i) {
    logc[i] = log_code.Decode(bs);
    if (logc[i] == 11) {
      bs->Fill();
      const int l = int(bs->Peek(8));
      bs->Skip(8);
      run[i] = l;
      i += l - 2;
      continue;
    }
    if (logc[i] > omit_log) {
      omit_log = logc[i];
      omit_pos = i;
    }
  }
  if (omit_pos < 0) {
    *err = "histogram without a largest count";
    return false;
  }
  int prev = 0, left = 0, total = 0;
  for (int i = 0; i < len; ++i) {
    if (run[i]) {
      left = run[i] - 1;
      prev = i > 0 ? (*counts)[i - 1] : 0;
    }
    if (left > 0) {
      (*counts)[i] = prev;
      --left;
    } else {
      const int code = logc[i];
      if (i == omit_pos || code == 0) continue;
      if (code == 1) {
        (*counts)[i] = 1;
      } else {
        // 2^(code-1) plus the top code/2 bits of the remainder.
        const int keep = code >> 1;
        (*counts)[i] = (1 << (code - 1)) +
                       (int(bs->Read(unsigned(keep))) << (code - 1 - keep));
      }
    }
    total += (*counts)[i];
  }
  (*counts)[omit_pos] = kTotal - total;
  if ((*counts)[omit_pos] <= 0) {
    *err = "histogram counts exceed the table size";
    return false;
  }
  return true;
}

}  // namespace

bool ReadAnsTables(BitStream* bs, size_t num_histograms,
                   size_t max_alphabet, AnsTables* t, std::string* err) {
  t->num = num_histograms;
  t->slot.assign(num_histograms << kAnsLogTab, 0);
  std::vector<int> counts;
  for (size_t h = 0; h < num_histograms; ++h) {
    if (!ReadCounts(bs, &counts, err)) return false;
    if (counts.size() > max_alphabet) {
      *err = "histogram alphabet too large";
      return false;
    }
    uint32_t* slot = t->slot.data() + (h << kAnsLogTab);
    // Slots no symbol claims decode as symbol 0 at its offset 0.
    const uint32_t f0 = uint32_t(counts[0]);
    for (uint32_t r = 0; r < kAnsTab; ++r) slot[r] = (r << 8) | (f0 << 18);
    uint32_t pos = 0;
    for (size_t s = 0; s < counts.size(); ++s) {
      const uint32_t f = uint32_t(counts[s]);
      if (pos + f > kAnsTab) {
        *err = "histogram overflows the ANS table";
        return false;
      }
      for (uint32_t j = 0; j < f; ++j) {
        slot[pos + j] = uint32_t(s) | (j << 8) | (f << 18);
      }
      pos += f;
    }
  }
  return true;
}

bool ReadHistograms(BitStream* bs, size_t num_contexts, size_t max_alphabet,
                    AnsTables* t, std::vector<uint8_t>* map,
                    std::string* err) {
  size_t num = 1;
  map->assign(num_contexts, 0);
  if (num_contexts > 1 && !ReadContextMap(bs, num_contexts, map, &num, err)) {
    return false;
  }
  if (!ReadAnsTables(bs, num, max_alphabet, t, err)) return false;
  if (!bs->JumpToByteBoundary()) {
    *err = "non-zero padding after the histograms";
    return false;
  }
  return true;
}

const int32_t* ZigZag8() {
  static const struct Scan {
    Scan() {
      // Anti-diagonals s = x + y, alternating direction: even ones run
      // from bottom-left to top-right.
      int n = 0;
      for (int s = 0; s < 15; ++s) {
        for (int i = 0; i <= s; ++i) {
          const int y = (s & 1) ? i : s - i;
          const int x = s - y;
          if (x < 8 && y < 8) pos[n++] = y * 8 + x;
        }
      }
    }
    int32_t pos[64];
  } scan;
  return scan.pos;
}

void ReadCoeffOrder(BitStream* bs, int32_t order[64]) {
  // Lehmer code of the order's permutation of the zig-zag scan, in four
  // spans of 16 that are either all zero (one 0 bit) or sent as 3-bit
  // groups (7 = add 7 and continue); codes up to the last non-zero one
  // carry +1.
  int32_t code[64] = {0};
  for (int i = 0; i < 64; i += 16) {
    if (!bs->Read(1)) continue;
    for (int j = i > 0 ? i : 1; j < i + 16; ++j) {
      int32_t v = 0;
      while (v <= 64) {
        const int32_t b = int32_t(bs->Read(3));
        v += b;
        if (b < 7) break;
      }
      code[j] = std::min(v, 64);
    }
  }
  int end = 63;
  while (end > 0 && code[end] == 0) --end;
  for (int i = 1; i <= end; ++i) --code[i];
  // Element i is the code[i]-th smallest index not yet taken; a code
  // outside the remaining range gives 0 and takes nothing
  // (pik/lehmer_code.cc:36-62).
  int32_t left[64];
  int n = 64;
  for (int i = 0; i < 64; ++i) left[i] = i;
  const int32_t* zz = ZigZag8();
  for (int i = 0; i < 64; ++i) {
    const int32_t c = code[i];
    int32_t v = 0;
    if (c >= 0 && c < n) {
      v = left[c];
      memmove(left + c, left + c + 1, size_t(n - c - 1) * sizeof(int32_t));
      --n;
    }
    order[i] = zz[v];
  }
}

namespace {

// Context layout of the AC histograms (pik/entropy_coder.h:80-150): 3 block
// contexts (the channel) x 32 buckets of the predicted non-zero count, then
// per block context 105 zero-density contexts indexed by the non-zeros left
// and the scan position.
constexpr int kBlockContexts = 3;
constexpr int kNonZeroBuckets = 32;
constexpr int kZeroDensityContexts = 105;
constexpr size_t kAcContexts =
    kBlockContexts * (kNonZeroBuckets + kZeroDensityContexts);

// Zero-density context of `left` non-zeros still to come (1..64) at scan
// position k (0..63).
inline int ZeroDensityContext(int left, int k) {
  int base;
  if (left <= 2) base = 0;
  else if (left <= 5) base = 16;
  else if (left <= 9) base = 32;
  else if (left <= 15) base = 48;
  else if (left <= 23) base = 64;
  else if (left <= 39) base = 79;
  else base = 93;
  int f;
  if (k < 4) f = k;
  else if (k < 12) f = 4 + ((k - 4) >> 1);
  else if (k < 32) f = 8 + ((k - 12) >> 2);
  else if (k < 48) f = 13 + ((k - 32) >> 3);
  else f = 15;
  return base + f;
}

// Coefficient token -> (zeros skipped << 4) | magnitude bits: the inverse of
// the format's token numbering of (skip, bits) pairs
// (pik/entropy_coder.cc:43-62).
const uint8_t kTokenSkipBits[256] = {
    0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x21, 0x12, 0x31, 0x41, 0x05, 0x51,
    0x13, 0x61, 0x22, 0x71, 0x81, 0x06, 0x91, 0x32, 0xa1, 0xf0, 0x14, 0xb1,
    0xc1, 0xd1, 0x23, 0x42, 0x52, 0xe1, 0xf1, 0x15, 0x07, 0x62, 0x33, 0x72,
    0x24, 0x82, 0x92, 0x43, 0xa2, 0x53, 0xb2, 0x16, 0x34, 0xc2, 0x63, 0x73,
    0x25, 0xd2, 0x93, 0x83, 0x44, 0x54, 0xa3, 0xb3, 0xc3, 0x35, 0xd3, 0x94,
    0x17, 0x45, 0x84, 0x64, 0x55, 0x74, 0x26, 0xe2, 0x08, 0x75, 0xc4, 0xd4,
    0x65, 0xf2, 0x85, 0x95, 0xa4, 0xb4, 0x36, 0x46, 0x56, 0xe3, 0xa5, 0x09,
    0x0a, 0x0b, 0x0c, 0x0d, 0x0e, 0x0f, 0x10, 0x18, 0x19, 0x1a, 0x1b, 0x1c,
    0x1d, 0x1e, 0x1f, 0x20, 0x27, 0x28, 0x29, 0x2a, 0x2b, 0x2c, 0x2d, 0x2e,
    0x2f, 0x30, 0x37, 0x38, 0x39, 0x3a, 0x3b, 0x3c, 0x3d, 0x3e, 0x3f, 0x40,
    0x47, 0x48, 0x49, 0x4a, 0x4b, 0x4c, 0x4d, 0x4e, 0x4f, 0x50, 0x57, 0x58,
    0x59, 0x5a, 0x5b, 0x5c, 0x5d, 0x5e, 0x5f, 0x60, 0x66, 0x67, 0x68, 0x69,
    0x6a, 0x6b, 0x6c, 0x6d, 0x6e, 0x6f, 0x70, 0x76, 0x77, 0x78, 0x79, 0x7a,
    0x7b, 0x7c, 0x7d, 0x7e, 0x7f, 0x80, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x8b,
    0x8c, 0x8d, 0x8e, 0x8f, 0x90, 0x96, 0x97, 0x98, 0x99, 0x9a, 0x9b, 0x9c,
    0x9d, 0x9e, 0x9f, 0xa0, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xab, 0xac, 0xad,
    0xae, 0xaf, 0xb0, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xbb, 0xbc, 0xbd,
    0xbe, 0xbf, 0xc0, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xcb, 0xcc, 0xcd,
    0xce, 0xcf, 0xd0, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xdb, 0xdc, 0xdd,
    0xde, 0xdf, 0xe0, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xeb, 0xec,
    0xed, 0xee, 0xef, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa, 0xfb,
    0xfc, 0xfd, 0xfe, 0xff};

constexpr int kGroupBlocks = 64;  // 512x512-pixel groups
constexpr int kTileBlocks = 8;    // 64x64-pixel tiles

inline int Covered(int type) {
  return type == PIKGPU_ACS_DCT16X16 ? 2 : type == PIKGPU_ACS_DCT32X32 ? 4 : 1;
}

struct GroupJob {
  const uint8_t* data;
  size_t size;
  const uint8_t* acs;  // group origin
  size_t acs_stride;
  int bxs, bys;        // group size in blocks
  int16_t* ac[3];      // group origin
  size_t ac_stride;
  float noise[3];
  size_t consumed;
};

// Every transform lies in one tile of the group and its blocks carry its
// type with indices 0, 1, ... in row-major order, so that every block is
// written exactly once: a first block checks its whole region, any other
// block that its first block is one.
bool CheckStrategies(const GroupJob& g, std::string* err) {
  for (int y = 0; y < g.bys; ++y) {
    const uint8_t* row = g.acs + size_t(y) * g.acs_stride;
    for (int x = 0; x < g.bxs; ++x) {
      const int type = row[x] >> 4, idx = row[x] & 15;
      if (type > PIKGPU_ACS_DCT4X4_NOHF) {
        *err = "unknown AC strategy";
        return false;
      }
      const int sc = Covered(type);
      if (idx >= sc * sc) {
        *err = "AC strategy block index out of range";
        return false;
      }
      if (sc == 1) continue;
      const int fy = y - idx / sc, fx = x - idx % sc;
      if (fy < 0 || fx < 0 || fy + sc > g.bys || fx + sc > g.bxs ||
          fy / kTileBlocks != (fy + sc - 1) / kTileBlocks ||
          fx / kTileBlocks != (fx + sc - 1) / kTileBlocks) {
        *err = "AC strategy transform outside its tile";
        return false;
      }
      if (idx != 0) {
        if (g.acs[size_t(fy) * g.acs_stride + fx] != uint8_t(type << 4)) {
          *err = "inconsistent AC strategy";
          return false;
        }
        continue;
      }
      for (int i = 1; i < sc * sc; ++i) {
        const uint8_t want = uint8_t(type << 4 | i);
        if (g.acs[size_t(y + i / sc) * g.acs_stride + x + i % sc] != want) {
          *err = "inconsistent AC strategy";
          return false;
        }
      }
    }
  }
  return true;
}

// The histogram of every AC context under one group's context map:
// count[c][b] for channel c's non-zero count with predicted-count bucket b
// (0..32), coef[c][left][k] for a coefficient token with `left` non-zeros
// still to come at scan position k (left + k <= 64).
struct AcHistograms {
  explicit AcHistograms(const std::vector<uint8_t>& map) {
    memset(coef, 0, sizeof(coef));
    for (int c = 0; c < 3; ++c) {
      for (int b = 0; b <= kNonZeroBuckets; ++b) {
        count[c][b] = map[kBlockContexts * b + c];
      }
      const int base = kBlockContexts * kNonZeroBuckets +
                       kZeroDensityContexts * c;
      for (int left = 1; left <= 64; ++left) {
        for (int k = 0; left + k <= 64; ++k) {
          coef[c][left][k] = map[base + ZeroDensityContext(left, k)];
        }
      }
    }
  }
  uint8_t count[3][kNonZeroBuckets + 1];
  uint8_t coef[3][65][64];
};

// Where each transmitted coefficient lands in the frame layout. A 1x1
// transform's block is stored as is. An S x S-block transform's (8S)^2
// coefficients are stored row-major over S rows of S * 64 (the layout
// GatherCoefficients makes, pik/ac_strategy.h:254-272); the bitstream
// carries them interleaved over the transform's blocks: block (dy, dx)
// holds coefficient (y, x), y = (q / 8) * S + dy, x = (q % 8) * S + dx, at
// its position q (pik/dct_util.h:21-55). For each
// (S, dy, dx, q) the table holds the offset from the transform's first
// slot in the frame: row * ac_stride + column.
struct Placement {
  explicit Placement(size_t stride) {
    for (int q = 0; q < 64; ++q) one[q] = q;
    Fill(2, stride, s2);
    Fill(4, stride, s4);
  }
  static void Fill(int s, size_t stride, int32_t* t) {
    for (int d = 0; d < s * s; ++d) {
      const int dy = d / s, dx = d % s;
      for (int q = 0; q < 64; ++q) {
        const int y = (q / 8) * s + dy, x = (q % 8) * s + dx;
        const int p = y * 8 * s + x;
        const int row = p / (s * 64), col = p % (s * 64);
        t[d * 64 + q] = int32_t(size_t(row) * stride) + col;
      }
    }
  }
  int32_t one[64];
  int32_t s2[4 * 64];
  int32_t s4[16 * 64];
};

// One tile's coefficients in scan order (DecodeAC,
// pik/entropy_coder.cc:1149-1212), written straight to their places in the
// frame layout (the reference decodes a tile into a scratch buffer and
// gathers the larger transforms afterwards, pik/compressed_image.cc:
// 783-801).
bool DecodeTile(const GroupJob& g, const AcHistograms& hist,
                const Placement& place, const int32_t* order, BitStream* bs,
                AnsReader* ans, int tx0, int ty0, int txs, int tys,
                std::string* err) {
  int32_t nz[kTileBlocks][kTileBlocks];
  for (int c = 0; c < 3; ++c) {
    const int32_t* ord = order + 64 * c;
    for (int by = 0; by < tys; ++by) {
      const uint8_t* acs = g.acs + size_t(ty0 + by) * g.acs_stride + tx0;
      for (int bx = 0; bx < txs; ++bx) {
        const int type = acs[bx] >> 4, idx = acs[bx] & 15;
        const int s = Covered(type);
        // The transform's first block and this block's place in it.
        const int fy = ty0 + by - idx / s, fx = tx0 + bx - idx % s;
        int16_t* base = g.ac[c] + size_t(fy) * g.ac_stride + size_t(fx) * 64;
        const int32_t* pos =
            s == 1 ? place.one : (s == 2 ? place.s2 : place.s4) + idx * 64;
        if (idx == 0) {
          for (int i = 0; i < s; ++i) {
            memset(base + size_t(i) * g.ac_stride, 0,
                   size_t(s) * 64 * sizeof(int16_t));
          }
        }
        // Predicted count: mean of top and left, either alone at an edge,
        // 32 at the tile's corner.
        int pred;
        if (bx == 0) pred = by == 0 ? 32 : nz[by - 1][bx];
        else if (by == 0) pred = nz[by][bx - 1];
        else pred = (nz[by - 1][bx] + nz[by][bx - 1] + 1) / 2;
        const int n = ans->Read(hist.count[c][pred >> 1], bs);
        nz[by][bx] = n;
        if (n > 64) {
          *err = "AC block with more than 64 non-zeros";
          return false;
        }
        int left = n;
        for (int k = 1; k < 64 && left > 0; ++k) {
          // A valid stream keeps left + k <= 64 (the reference asserts
          // left + k - 1 < 65 here, pik/entropy_coder.h:141-146).
          if (left + k - 1 >= 65) {
            *err = "AC non-zeros exceed the positions left";
            return false;
          }
          bs->Fill();
          const uint8_t sb =
              kTokenSkipBits[ans->ReadFilled(hist.coef[c][left][k - 1], bs)];
          k += sb >> 4;
          const unsigned nbits = sb & 15;
          if (nbits == 0) continue;
          if (k + left > 64) {
            *err = "AC non-zeros run past the block";
            return false;
          }
          const uint32_t u = (1u << nbits) + bs->Peek(nbits) - 1;
          bs->Skip(nbits);
          // Even u -> u/2, odd u -> -(u+1)/2.
          const int32_t v = int32_t((u >> 1) ^ (((~u) & 1) - 1));
          --left;
          base[pos[ord[k]]] = int16_t(v);
        }
        if (left != 0) {
          *err = "AC block ends with non-zeros left";
          return false;
        }
      }
    }
  }
  return true;
}

bool DecodeGroup(GroupJob* g, std::string* err) {
  if (!CheckStrategies(*g, err)) return false;
  BitStream bs(g->data, g->size);
  // Noise parameters (DecodeNoise, pik/noise.cc:529-539): sign bit, then
  // 16 bits of |value| * 1000.
  g->noise[0] = g->noise[1] = g->noise[2] = 0.0f;
  if (bs.Read(1)) {
    for (int i = 0; i < 3; ++i) {
      const int sign = 2 * int(bs.Read(1)) - 1;
      const int mag = int(bs.Read(16));
      g->noise[i] = float(sign * mag) / 1000.0f;
    }
  }
  if (!bs.JumpToByteBoundary()) {
    *err = "non-zero padding after the noise parameters";
    return false;
  }
  int32_t order[3 * 64];
  for (int c = 0; c < 3; ++c) ReadCoeffOrder(&bs, order + 64 * c);
  if (!bs.JumpToByteBoundary()) {
    *err = "non-zero padding after the coefficient orders";
    return false;
  }
  AnsTables tables;
  std::vector<uint8_t> map;
  if (!ReadHistograms(&bs, kAcContexts, 256, &tables, &map, err)) {
    return false;
  }
  AnsReader ans(&tables);
  const AcHistograms hist(map);
  const Placement place(g->ac_stride);
  const int ntx = (g->bxs + kTileBlocks - 1) / kTileBlocks;
  const int nty = (g->bys + kTileBlocks - 1) / kTileBlocks;
  for (int ty = 0; ty < nty; ++ty) {
    for (int tx = 0; tx < ntx; ++tx) {
      const int x0 = tx * kTileBlocks, y0 = ty * kTileBlocks;
      const int txs = std::min(kTileBlocks, g->bxs - x0);
      const int tys = std::min(kTileBlocks, g->bys - y0);
      if (!DecodeTile(*g, hist, place, order, &bs, &ans, x0, y0, txs, tys,
                      err)) {
        return false;
      }
    }
  }
  if (!ans.FinalStateOk()) {
    *err = "ANS final state mismatch";
    return false;
  }
  if (!bs.JumpToByteBoundary()) {
    *err = "non-zero padding after the AC data";
    return false;
  }
  g->consumed = bs.Position();
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

extern "C" PIKGPU_API int pikgpu_decode_ac_pass(const pikgpu_ac_pass* p,
                                                int threads, char* msg,
                                                size_t msg_size) {
  using namespace pikgpu::ent;
  if (msg != nullptr && msg_size > 0) msg[0] = '\0';
  if (p == nullptr || p->ac_strategy == nullptr || p->group_data == nullptr ||
      p->group_size == nullptr || p->ac[0] == nullptr ||
      p->ac[1] == nullptr || p->ac[2] == nullptr) {
    SetMsg(msg, msg_size, "null argument");
    return PIKGPU_EINVAL;
  }
  if (p->xsize == 0 || p->ysize == 0) {
    SetMsg(msg, msg_size, "empty frame");
    return PIKGPU_EINVAL;
  }
  const int bx = int((p->xsize + 7) / 8), by = int((p->ysize + 7) / 8);
  const int gxs = (bx + kGroupBlocks - 1) / kGroupBlocks;
  const int gys = (by + kGroupBlocks - 1) / kGroupBlocks;
  if (p->num_groups != size_t(gxs) * size_t(gys)) {
    SetMsg(msg, msg_size, "%zu groups given, the frame has %d",
           p->num_groups, gxs * gys);
    return PIKGPU_EINVAL;
  }
  if (p->ac_strategy_stride < size_t(bx) || p->ac_stride < size_t(bx) * 64) {
    SetMsg(msg, msg_size, "stride smaller than a row");
    return PIKGPU_EINVAL;
  }
  std::vector<GroupJob> jobs(p->num_groups);
  for (int gy = 0; gy < gys; ++gy) {
    for (int gx = 0; gx < gxs; ++gx) {
      GroupJob& g = jobs[size_t(gy) * gxs + gx];
      const size_t i = size_t(gy) * gxs + gx;
      if (p->group_data[i] == nullptr && p->group_size[i] != 0) {
        SetMsg(msg, msg_size, "group %zu: null data", i);
        return PIKGPU_EINVAL;
      }
      g.data = p->group_data[i];
      g.size = p->group_size[i];
      const int x0 = gx * kGroupBlocks, y0 = gy * kGroupBlocks;
      g.bxs = std::min(kGroupBlocks, bx - x0);
      g.bys = std::min(kGroupBlocks, by - y0);
      g.acs = p->ac_strategy + size_t(y0) * p->ac_strategy_stride + x0;
      g.acs_stride = p->ac_strategy_stride;
      for (int c = 0; c < 3; ++c) {
        g.ac[c] = p->ac[c] + size_t(y0) * p->ac_stride + size_t(x0) * 64;
      }
      g.ac_stride = p->ac_stride;
      g.consumed = 0;
    }
  }
  int nt = threads > 0 ? threads
                       : int(std::max(1u, std::thread::hardware_concurrency()));
  nt = std::min<int>(nt, int(jobs.size()));
  std::atomic<size_t> next{0};
  std::mutex lock;
  size_t bad = jobs.size();
  std::string bad_err;
  auto work = [&]() {
    std::string err;
    for (;;) {
      const size_t i = next.fetch_add(1);
      if (i >= jobs.size()) return;
      if (!DecodeGroup(&jobs[i], &err)) {
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
  for (size_t i = 0; i < jobs.size(); ++i) {
    if (p->noise != nullptr) {
      for (int k = 0; k < 3; ++k) p->noise[3 * i + k] = jobs[i].noise[k];
    }
    if (p->group_consumed != nullptr) p->group_consumed[i] = jobs[i].consumed;
  }
  if (bad < jobs.size()) {
    SetMsg(msg, msg_size, "group %zu: %s", bad, bad_err.c_str());
    return PIKGPU_EDATA;
  }
  return PIKGPU_OK;
}
