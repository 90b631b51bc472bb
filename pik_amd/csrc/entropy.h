// Host entropy decoding of PIK bitstreams: the bit reader, the prefix
// (Huffman) codes of the context map, the ANS histograms and symbol reader,
// and the coefficient orders. The AC group decoder (entropy.cpp) is built
// from these; they stay on the host, ahead of the HIP reconstruction.
//
// The formats follow the reference decoder:
//   bit reader             pik/bit_reader.h:22-135 (LSB-first, zeros past
//                          the end)
//   prefix codes           pik/huffman_decode.cc:178-336 (RFC 7932 3.4-3.5)
//   context map            pik/context_map_decode.cc:74-117
//   ANS histograms         pik/ans_decode.cc:31-184
//   ANS symbol reader      pik/ans_decode.h:36-78
//   coefficient orders     pik/entropy_coder.cc:889-924
#ifndef PIKGPU_ENTROPY_H_
#define PIKGPU_ENTROPY_H_

#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

namespace pikgpu {
namespace ent {

// Bits are taken from the least significant end of each byte first; bytes
// past the end read as zero (the reference reader does the same and lets
// the caller judge the result). A 64-bit buffer holds at least 32 unread
// bits after Fill().
class BitStream {
 public:
  BitStream(const uint8_t* data, size_t size) : data_(data), size_(size) {}

  inline void Fill() {
    if (avail_ >= 32) return;
    if (next_ + 8 <= size_) {
      // Whole-word refill: the bits above `avail_` that do not fit are the
      // following bytes, which a later refill ORs in again unchanged.
      uint64_t w;
      memcpy(&w, data_ + next_, 8);
      buf_ |= w << avail_;
      const unsigned take = (63 - avail_) >> 3;
      next_ += take;
      avail_ += take * 8;
    } else {
      while (avail_ <= 56) {
        const uint64_t b = next_ < size_ ? data_[next_] : 0;
        buf_ |= b << avail_;
        ++next_;
        avail_ += 8;
      }
    }
  }
  // n <= 32, with n bits in the buffer.
  inline uint32_t Peek(unsigned n) const {
    return uint32_t(buf_ & ((uint64_t(1) << n) - 1));
  }
  inline void Skip(unsigned n) {
    buf_ >>= n;
    avail_ -= n;
  }
  inline uint32_t Read(unsigned n) {
    Fill();
    const uint32_t v = Peek(n);
    Skip(n);
    return v;
  }
  size_t BitsRead() const { return next_ * 8 - avail_; }
  // Bytes consumed, rounded up.
  size_t Position() const { return (BitsRead() + 7) / 8; }
  bool Overrun() const { return BitsRead() > size_ * 8; }
  // Skips to the next byte boundary; the skipped bits must be zero.
  bool JumpToByteBoundary() {
    const unsigned r = unsigned(BitsRead() & 7);
    return r == 0 || Read(8 - r) == 0;
  }
  // Skips n bits. A skip past the end lands 8 bytes beyond it (everything
  // there reads as zeros, and Position() reports more than the size).
  void SkipBits(uint64_t n) {
    if (n <= avail_) {
      Skip(unsigned(n));
      return;
    }
    const uint64_t limit = uint64_t(size_) * 8 + 64;
    const uint64_t at = BitsRead();
    const uint64_t to = n > limit - std::min(at, limit) ? limit : at + n;
    next_ = size_t(to >> 3);
    buf_ = 0;
    avail_ = 0;
    if (to & 7) Read(unsigned(to & 7));
  }

 private:
  const uint8_t* data_;
  size_t size_;
  size_t next_ = 0;  // next byte to load
  uint64_t buf_ = 0;
  unsigned avail_ = 0;
};

// A canonical prefix code (RFC 7932 3.2) read as PIK's context map writes
// it; Decode() looks up `max_len` peeked bits in one table.
class PrefixCode {
 public:
  bool Read(BitStream* bs, std::string* err);
  inline int Decode(BitStream* bs) const {
    bs->Fill();
    const uint32_t e = table_[bs->Peek(max_len_)];
    bs->Skip(e >> 16);
    return int(e & 0xFFFF);
  }

 private:
  bool Build(const std::vector<uint8_t>& lengths, std::string* err);
  std::vector<uint32_t> table_;  // (length << 16) | symbol
  unsigned max_len_ = 0;
};

constexpr int kAnsLogTab = 10;
constexpr uint32_t kAnsTab = 1u << kAnsLogTab;
constexpr uint32_t kAnsSignature = 0x13;
constexpr uint32_t kAnsChunk = 1u << 16;

// Decoding tables of a set of ANS histograms: per histogram, one entry per
// state slot r in [0, 1024): the symbol s whose range holds r, its frequency
// and r minus the range start, packed as sym | delta << 8 | freq << 18.
struct AnsTables {
  std::vector<uint32_t> slot;
  size_t num = 0;
};

// DecodeANSCodes (pik/ans_decode.cc:144-183).
bool ReadAnsTables(BitStream* bs, size_t num_histograms,
                   size_t max_alphabet, AnsTables* t, std::string* err);

class AnsReader {
 public:
  explicit AnsReader(const AnsTables* t) : slot_(t->slot.data()) {}
  inline int Read(size_t histogram, BitStream* bs) {
    bs->Fill();
    return ReadFilled(histogram, bs);
  }
  // With at least 32 bits in the buffer; at least 16 remain afterwards.
  inline int ReadFilled(size_t histogram, BitStream* bs) {
    if (left_ == 0) {
      // The state is (re)loaded every kAnsChunk symbols.
      const uint32_t hi = bs->Read(16);
      state_ = (hi << 16) | bs->Read(16);
      bs->Fill();
      left_ = kAnsChunk;
    }
    const uint32_t r = state_ & (kAnsTab - 1);
    const uint32_t e = slot_[(histogram << kAnsLogTab) + r];
    state_ = (e >> 18) * (state_ >> kAnsLogTab) + ((e >> 8) & 1023);
    --left_;
    if (state_ < (1u << 16)) {
      state_ = (state_ << 16) | bs->Peek(16);
      bs->Skip(16);
    }
    return int(e & 0xFF);
  }
  bool FinalStateOk() const { return state_ == (kAnsSignature << 16); }

 private:
  const uint32_t* slot_;
  uint32_t state_ = kAnsSignature << 16;
  uint32_t left_ = 0;
};

// DecodeContextMap (pik/context_map_decode.cc:74-117): `map` receives
// num_contexts histogram indices.
bool ReadContextMap(BitStream* bs, size_t num_contexts,
                    std::vector<uint8_t>* map, size_t* num_histograms,
                    std::string* err);

// DecodeHistograms (pik/entropy_coder.cc:822-836): context map, then the
// histograms, then zero bits to the byte boundary.
bool ReadHistograms(BitStream* bs, size_t num_contexts, size_t max_alphabet,
                    AnsTables* t, std::vector<uint8_t>* map,
                    std::string* err);

// DecodeCoeffOrder (pik/entropy_coder.cc:889-924): the scan order of one
// block context as 64 positions of the 8x8 block.
void ReadCoeffOrder(BitStream* bs, int32_t order[64]);

// The zig-zag scan of an 8x8 block (row-major positions).
const int32_t* ZigZag8();

}  // namespace ent
}  // namespace pikgpu

struct pikgpu_dc_section;

namespace pikgpu {
namespace ent {

// pikgpu_decode_dc_groups with the size table starting skip_bits into
// s->data (the frame parser reads it from wherever the colour maps left
// the reader); s->consumed counts from s->data.
int DecodeDcSection(pikgpu_dc_section* s, size_t skip_bits, int threads,
                    char* msg, size_t msg_size);

}  // namespace ent
}  // namespace pikgpu

#endif  // PIKGPU_ENTROPY_H_
