// rs_host.cpp -- host-side exact arithmetic: the CPU re-check of every
// reported factor (the reference's verify_factor.c) and the wheel offset of
// the segment generator (cl_sieve.cpp:453-512).
#include <cstdint>

#include "../../include/rssieve.h"
#include "rs_mont.h"

using rs::madd;
using rs::mmul;
using rs::mont_inv;

namespace {

struct Mont {
  uint64_t p, q, one, pmo, r2;
  explicit Mont(uint64_t p_) : p(p_) {
    q = mont_inv(p);
    one = (0 - p) % p;
    pmo = p - one;
    uint64_t t = madd(madd(one, one, p), madd(one, one, p), p);  // Montgomery 4
    for (int i = 0; i < 5; ++i) t = mmul(t, t, p, q);           // 4^(2^5) = 2^64 -> R^2 mod p
    r2 = t;
  }
  uint64_t to(uint64_t x) const { return mmul(x, r2, p, q); }
};

// strong probable-prime test of odd p to base a (a < p, a != 0)
bool sprp(const Mont& m, uint64_t a) {
  const uint64_t p = m.p;
  const int t = __builtin_ctzll(p - 1);
  const uint64_t d = (p - 1) >> t;
  const uint64_t ma = m.to(a);
  uint64_t x = m.one;
  for (int b = 63 - __builtin_clzll(d); b >= 0; --b) {
    x = mmul(x, x, p, m.q);
    if ((d >> b) & 1) x = mmul(x, ma, p, m.q);
  }
  if (x == m.one || x == m.pmo) return true;
  for (int s = 1; s < t; ++s) {
    x = mmul(x, x, p, m.q);
    if (x == m.pmo) return true;
  }
  return false;
}

}  // namespace

// Deterministic for all 64-bit p: strong tests to the seven bases
// {2, 325, 9375, 28178, 450775, 9780504, 1795265022} (Jim Sinclair's set, the
// one verify_factor.c:100-128 uses).
extern "C" int rs_is_prime(uint64_t p) {
  if (p < 2) return 0;
  if (p < 4) return 1;
  if (!(p & 1)) return 0;
  static const uint32_t bases[7] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
  Mont m(p);
  for (uint32_t b : bases) {
    uint64_t a = b % p;
    if (a == 0) continue;
    if (!sprp(m, a)) return 0;
  }
  return 1;
}

// 1: p divides k*b^n + c;  0: it does not;  -1: p is composite
// (verify_factor.c:131-181).  Computed as k * b^n mod p with a right-to-left
// ladder in exact Montgomery arithmetic, compared with -c mod p.
extern "C" int rs_verify_factor(uint64_t p, uint64_t k, uint32_t n, int32_t c, uint32_t b) {
  if (p < 3 || !(p & 1)) return -1;  // isPrime() rejects even p (verify_factor.c:164-165)
  if (!rs_is_prime(p)) return -1;
  Mont m(p);
  uint64_t acc = m.one;
  uint64_t sq = m.to(b % p);
  for (uint32_t e = n; e; e >>= 1) {
    if (e & 1) acc = mmul(acc, sq, p, m.q);
    sq = mmul(sq, sq, p, m.q);
  }
  // leave Montgomery form: acc * k * R^-1 with acc = b^n R  ->  k * b^n mod p
  const uint64_t v = mmul(acc, k % p, p, m.q);
  const uint64_t target = (c == -1) ? 1 % p : p - 1;
  if (c != 1 && c != -1) return 0;
  return v == target ? 1 : 0;
}

// cl_sieve.cpp:453-512: first N >= start coprime to 30, and the index of the
// gap that follows it in the wheel {4,2,4,2,4,6,2,6} (which starts at 7).
extern "C" void rs_wheel_offset(uint64_t* start, int32_t* index) {
  static const int32_t residues[8] = {7, 11, 13, 17, 19, 23, 29, 1};  // wheel position i starts at residues[i]
  uint64_t N = *start;
  while (true) {
    const uint32_t r = (uint32_t)(N % 30);
    for (int i = 0; i < 8; i++)
      if ((uint32_t)residues[i] == r) {
        *start = N;
        *index = i;
        return;
      }
    ++N;
  }
}
