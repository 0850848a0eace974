// Prints Philox4x32-10 blocks from rocRAND's engine header (host build) for the
// counters/keys given on stdin as "c0 c1 c2 c3 k0 k1" lines (hex).  Test
// infrastructure: pins oracle/oracle.cpp:orc_philox4x32_10 against an independent
// implementation.  engine(seed = k1:k0, subsequence = c3:c2, offset = 4 * (c1:c0))
// starts at counter {c0, c1, c2, c3}.
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>

int main() {
  unsigned c0, c1, c2, c3, k0, k1;
  while (std::scanf("%x %x %x %x %x %x", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
    const unsigned long long seed = ((unsigned long long)k1 << 32) | k0;
    const unsigned long long subseq = ((unsigned long long)c3 << 32) | c2;
    const unsigned long long offset = 4ull * (((unsigned long long)c1 << 32) | c0);
    rocrand_device::philox4x32_10_engine eng(seed, subseq, offset);
    unsigned a = eng(), b = eng(), c = eng(), d = eng();
    std::printf("%08x %08x %08x %08x\n", a, b, c, d);
  }
  return 0;
}
