// Host check of the binary64 base-2 strong probable-prime test of rs_mont.h
// (fp_sprp2, the prime generator's test below 2^51) against the integer
// Montgomery one (sprp2, the reference's steps): random odd N below 2^51,
// N = 2^t + 1 (the masked-shift case), small base-2 strong pseudoprimes and
// N next to 2^51.  Prints the counts (mismatches 0 expected) and exits
// non-zero on any mismatch.
#include <stdio.h>
#include <stdlib.h>

#include "../../rssieve_amd/csrc/rs_mont.h"

static uint64_t s_state;
static uint64_t next_u64() {  // splitmix64
  uint64_t z = (s_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static uint64_t bad = 0, checked = 0, passing = 0;

static void check(uint64_t N) {
  const bool want = rs::sprp2(N), got = rs::fp_sprp2(N);
  checked++;
  passing += want;
  if (want != got) {
    if (bad < 10) printf("mismatch N=%llu integer %d binary64 %d\n", (unsigned long long)N, want, got);
    bad++;
  }
}

int main(int argc, char** argv) {
  const uint64_t n = argc > 1 ? strtoull(argv[1], 0, 10) : 1000000;
  s_state = argc > 2 ? strtoull(argv[2], 0, 10) : 1;
  for (uint64_t i = 0; i < n; i++) {
    const int bits = 3 + (int)(next_u64() % 49);  // 3..51 bits
    uint64_t N = (next_u64() & ((1ull << bits) - 1)) | 1ull | (1ull << (bits - 1));
    if (N >= (1ull << 51)) N = (1ull << 51) - 1;
    if (N >= 3) check(N);
  }
  for (int t = 1; t < 51; t++) check((1ull << t) + 1);
  const uint64_t spsp[] = {2047, 3277, 4033, 4681, 8321, 15841, 29341, 42799, 49141, 52633, 65281, 74665, 80581,
                           85489, 88357, 90751, 1373653, 25326001, 3215031751ull, 2152302898747ull};
  for (uint64_t N : spsp) check(N);
  for (uint64_t N = (1ull << 51) - 4001; N < (1ull << 51); N += 2) check(N);
  for (uint64_t N = 3; N < 200001; N += 2) check(N);
  printf("checked %llu passing %llu mismatches %llu\n", (unsigned long long)checked, (unsigned long long)passing,
         (unsigned long long)bad);
  return bad ? 1 : 0;
}
