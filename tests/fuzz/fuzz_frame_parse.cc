// Mutation fuzz of the host frame parser (pikgpu_parse_frame, with the DC
// group, dot dictionary and AC group decoders it calls), built with
// AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_parse_sanitized.py. Arguments: a directory holding NAME.bin
// bitstreams (the fixture frames), the mutations per frame, then
// NAME:WIDTH:HEIGHT for each frame. Every mutated stream must either parse
// or be rejected with an error code; the sanitizers abort on any invalid
// access.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <random>
#include <string>
#include <vector>
#include "pikgpu.h"
int main(int argc, char** argv) {
  if (argc < 4) return 2;
  const std::string dir = argv[1];
  const int iters = atoi(argv[2]);
  std::mt19937 rng(12345);
  long ok = 0, bad = 0;
  pikgpu_parsed* reuse = nullptr;
  for (int a = 3; a < argc; ++a) {
    char name[64];
    unsigned w = 0, h = 0;
    if (sscanf(argv[a], "%63[^:]:%u:%u", name, &w, &h) != 3) return 2;
    const std::string path = dir + "/" + name + ".bin";
    FILE* fp = fopen(path.c_str(), "rb");
    if (fp == nullptr) return 2;
    std::vector<uint8_t> d;
    int c;
    while ((c = fgetc(fp)) != EOF) d.push_back(uint8_t(c));
    fclose(fp);
    char msg[256];
    pikgpu_parsed* p = nullptr;
    if (pikgpu_parse_frame(d.data(), d.size(), w, h, 2, &p, msg,
                           sizeof msg) != PIKGPU_OK) {
      printf("clean %s failed: %s\n", name, msg);
      return 1;
    }
    pikgpu_parsed_free(p);
    for (int it = 0; it < iters; ++it) {
      std::vector<uint8_t> m = d;
      const int kind = rng() % 4;
      if (kind == 0) {  // bit flips
        const int n = 1 + rng() % 8;
        for (int k = 0; k < n; ++k) {
          m[rng() % m.size()] ^= uint8_t(1u << (rng() % 8));
        }
      } else if (kind == 1) {  // truncation (to at least one byte)
        m.resize(1 + rng() % (m.size() - 1));
      } else if (kind == 2) {  // random bytes in a window
        const size_t at = rng() % m.size();
        const size_t end = std::min(m.size(), at + 1 + rng() % 16);
        for (size_t k = at; k < end; ++k) m[k] = uint8_t(rng());
      } else {  // a header byte
        m[rng() % std::min<size_t>(m.size(), 64)] = uint8_t(rng());
      }
      const int rc = pikgpu_parse_frame(m.data(), m.size(), w, h, 1 + it % 3,
                                        &reuse, msg, sizeof msg);
      if (rc == PIKGPU_OK) {
        ++ok;
        pikgpu_parsed_info info;
        pikgpu_frame fr;
        if (pikgpu_parsed_get(reuse, &fr, &info) != PIKGPU_OK) return 1;
      } else if (rc == PIKGPU_EDATA || rc == PIKGPU_EUNSUPPORTED) {
        ++bad;
      } else {
        printf("unexpected code %d: %s\n", rc, msg);
        return 1;
      }
    }
  }
  pikgpu_parsed_free(reuse);
  printf("fuzz done: %ld parsed, %ld rejected\n", ok, bad);
  return 0;
}
