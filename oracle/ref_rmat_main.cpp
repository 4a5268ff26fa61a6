// Driver that compiles the reference's own RMAT generator and hashes
// (include/SSTGraph/internal/rmat_util.h, which depends on the standard
// library only) straight from the reference tree, to print golden vectors.
// Test infrastructure: built by oracle/Makefile into oracle/_ref/ only where
// the reference tree is present; never linked into the product.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <limits>
#include <utility>

#include "SSTGraph/internal/rmat_util.h"

int main(int argc, char **argv) {
  if (argc >= 2 && std::strcmp(argv[1], "hash") == 0) {
    // hash KATs: "h32 x value" and "h64 x value" lines
    const uint32_t xs32[] = {0u, 1u, 2u, 3u, 7u, 255u, 256u, 65535u, 123456789u,
                             0x7fffffffu, 0x80000000u, 0xfffffffeu, 0xffffffffu};
    for (uint32_t x : xs32)
      std::printf("h32 %u %u\n", x, SSTGraph::hash32(x));
    const uint64_t xs64[] = {0ull, 1ull, 2ull, 42ull, 1000000007ull,
                             0xffffffffull, 0x100000000ull,
                             0x7fffffffffffffffull, 0xffffffffffffffffull};
    for (uint64_t x : xs64)
      std::printf("h64 %llu %llu\n", (unsigned long long)x,
                  (unsigned long long)SSTGraph::hash64(x));
    SSTGraph::random_aspen r;
    for (uint64_t i = 0; i < 4; i++)
      std::printf("aspen %llu %llu\n", (unsigned long long)i,
                  (unsigned long long)r.ith_rand(i));
    const uint64_t ls[] = {1, 2, 3, 4, 5, 1000, 65536, 65537, 16777216};
    for (uint64_t l : ls)
      std::printf("log2_up %llu %zu\n", (unsigned long long)l,
                  SSTGraph::log2_up(l));
    return 0;
  }
  if (argc < 5) {
    std::fprintf(stderr,
                 "usage: %s hash | %s n seed first count [a b c]\n", argv[0],
                 argv[0]);
    return 2;
  }
  uint32_t n = (uint32_t)std::strtoul(argv[1], nullptr, 10);
  uint32_t seed = (uint32_t)std::strtoul(argv[2], nullptr, 10);
  uint64_t first = std::strtoull(argv[3], nullptr, 10);
  uint64_t count = std::strtoull(argv[4], nullptr, 10);
  double a = 0.5, b = 0.1, c = 0.1;
  if (argc >= 8) {
    a = std::atof(argv[5]);
    b = std::atof(argv[6]);
    c = std::atof(argv[7]);
  }
  SSTGraph::rMat<uint32_t> rm(n, seed, a, b, c);
  for (uint64_t i = first; i < first + count; i++) {
    std::pair<uint32_t, uint32_t> e = rm((uint32_t)i);
    std::printf("%u %u\n", e.first, e.second);
  }
  return 0;
}
