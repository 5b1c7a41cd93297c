// Host check of the binary64 modular product of rs_mont.h against 128-bit
// integer arithmetic: random and extreme operands for odd moduli below 2^51,
// and the Montgomery-times-plain identity the sieve relies on.  Prints the
// number of mismatches (0 expected) and exits non-zero on any.
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

static uint64_t bad = 0, checked = 0;

static void check(uint64_t a, uint64_t b, uint64_t p) {
  const double pd = (double)p, pinv = 1.0 / pd, pc = pd + rs::kFpTwo52;
  const uint64_t want = (uint64_t)(((unsigned __int128)a * b) % p);
  const double got = rs::fp_mulmod((double)a, (double)b, pd, pinv);
  const uint64_t bits = rs::fp_mulmod_bits((double)a, (double)b, pd, pinv, pc);
  checked++;
  if (got != (double)want || (bits ^ rs::kFpBias) != want || rs::fp_to_u64(got) != want) {
    if (bad < 10) printf("mismatch a=%llu b=%llu p=%llu want=%llu got=%.0f\n", (unsigned long long)a,
                         (unsigned long long)b, (unsigned long long)p, (unsigned long long)want, got);
    bad++;
  }
}

// Signed operands in (-p, p): the product stays in (-p, p) and is congruent to
// a*b; the powmods (signed chains, canonical results) equal pow(x, e) mod p.
static void check_signed(uint64_t a, uint64_t b, uint64_t p, bool na, bool nb) {
  const double pd = (double)p, pinv = 1.0 / pd;
  const double sa = na && a ? (double)a - pd : (double)a, sb = nb && b ? (double)b - pd : (double)b;
  const double t = rs::fp_mulmod_signed(sa, sb, pd, pinv);
  const uint64_t want = (uint64_t)(((unsigned __int128)a * b) % p);
  checked++;
  if (!(t > -pd && t < pd) || rs::fp_to_u64(rs::fp_canon(t, pd)) != want) {
    if (bad < 10) printf("signed mismatch a=%llu b=%llu p=%llu\n", (unsigned long long)a, (unsigned long long)b,
                         (unsigned long long)p);
    bad++;
  }
}

static uint64_t powmod128(uint64_t x, uint64_t e, uint64_t p) {
  unsigned __int128 r = 1, b = x % p;
  while (e) {
    if (e & 1) r = r * b % p;
    b = b * b % p;
    e >>= 1;
  }
  return (uint64_t)r;
}

static void check_pow(uint64_t x, uint64_t e, uint64_t p) {
  const double pd = (double)p, pinv = 1.0 / pd;
  checked++;
  // exponents >= 2 (an exponent of 1 squares 64 times, the reference's masked shift)
  if (rs::fp_to_u64(rs::fp_powmod_lg((double)x, e, pd, pinv)) != powmod128(x, e, p)) bad++;
  checked++;
  if (e < (1ull << 32) &&
      rs::fp_to_u64(rs::fp_powmod_sm((double)x, (uint32_t)e, pd, pinv)) != powmod128(x, e, p)) bad++;
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? atol(argv[1]) : 1000000;
  s_state = argc > 2 ? strtoull(argv[2], 0, 10) : 1;
  // an optional third argument moves the modulus range to [2^(B-1), 2^B), which
  // must show mismatches for B well above 51 (the check can fail)
  const int B = argc > 3 ? atoi(argv[3]) : 51;
  const uint64_t top = (1ull << B) - 1;  // largest odd modulus on the binary64 path
  for (long i = 0; i < n; i++) {
    uint64_t p;
    switch (i % 4) {
      case 0: p = top - 2 * (next_u64() % 1000); break;                      // just below 2^51
      case 1: p = (next_u64() % top) | 1; break;                             // anywhere
      case 2: p = ((next_u64() >> (13 + next_u64() % 50)) | 1) + 2; break;  // every size
      default: p = (1ull << (B - 1)) + 2 * (next_u64() % (1ull << (B - 3))) + 1; break;
    }
    if (p > top || p < 3) continue;
    const uint64_t a = next_u64() % p, b = next_u64() % p;
    check(a, b, p);
    check(p - 1, p - 1 - (next_u64() % 4 < 2 ? 0 : b % 7), p);
    check(p - 1, b, p);
    check(0, b, p);
    check(1, b, p);
    check_signed(a, b, p, i & 1, i & 2);
    check_signed(p - 1, p - 1, p, true, i & 1);
    check_signed(1, b, p, true, true);
    if (i % 16 == 0) check_pow(a, 2 + next_u64() % (p - 2), p);
    // Montgomery residue times plain residue is the Montgomery residue of the product
    const uint64_t q = rs::mont_inv(p), one = rs::r_mod(p);
    const uint64_t am = rs::mmul(a, (uint64_t)(((unsigned __int128)one * one) % p), p, q);  // a*R mod p
    const double pd = (double)p, pinv = 1.0 / pd;
    const uint64_t prod_m = rs::fp_to_u64(rs::fp_mulmod((double)am, (double)b, pd, pinv));
    const uint64_t want_m = (uint64_t)(((unsigned __int128)((unsigned __int128)a * b % p) * one) % p);
    checked++;
    if (prod_m != want_m) bad++;
  }
  printf("checked %llu mismatches %llu\n", (unsigned long long)checked, (unsigned long long)bad);
  return bad != 0;
}
