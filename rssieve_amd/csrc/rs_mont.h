// rs_mont.h -- 64-bit Montgomery arithmetic (R = 2^64) for the sieve, usable
// from HIP device code and from host C++.
//
// Semantics follow the reference's kernels/common.cl:24-176: q = p^-1 mod 2^64
// (not -p^-1), products return canonical residues in [0, p) for odd p, and the
// left-to-right powmods reproduce OpenCL's masked shift count (an exponent of
// 1 squares 64 times, common.cl:52-53), so every value matches the reference
// bit for bit.  On gfx950 the 64x64->128 products lower to v_mad_u64_u32 /
// v_mul_hi_u32 / v_mul_lo_u32 on the VALU.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RS_HD __host__ __device__ __forceinline__
#else
#define RS_HD static inline
#endif

namespace rs {

RS_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

RS_HD int clz64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return x ? __clzll((long long)x) : 64;
#else
  return x ? __builtin_clzll(x) : 64;
#endif
}

RS_HD int clz32(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return x ? __clz((int)x) : 32;
#else
  return x ? __builtin_clz(x) : 32;
#endif
}

// p^-1 mod 2^64 by Newton's iteration; five steps suffice from the 3-bit seed
// p (p*p == 1 mod 8 for odd p), and the fixed point is the same as the
// reference's run-until-stable loop (common.cl:24-28).
RS_HD uint64_t mont_inv(uint64_t p) {
  uint64_t x = p;
  x *= 2 - p * x;
  x *= 2 - p * x;
  x *= 2 - p * x;
  x *= 2 - p * x;
  x *= 2 - p * x;
  return x;
}

// a*b*2^-64 mod p (common.cl:33-40)
RS_HD uint64_t mmul(uint64_t a, uint64_t b, uint64_t p, uint64_t q) {
  uint64_t lo = a * b;
  uint64_t hi = mulhi64(a, b);
  uint64_t m = lo * q;
  uint64_t mp = mulhi64(m, p);
  uint64_t r = hi - mp;
  return hi < mp ? r + p : r;
}

// (a + b) mod p (common.cl:42-47)
RS_HD uint64_t madd(uint64_t a, uint64_t b, uint64_t p) {
  uint64_t c = (a >= p - b) ? p : 0;
  return a + b - c;
}

// a * base for a Montgomery residue a: additions for base 2, 3, 5, otherwise a
// product with the Montgomery form of the base (common.cl:73-85).
RS_HD uint64_t mul_base(uint64_t a, uint64_t mbase, uint32_t base, uint64_t p, uint64_t q) {
  if (base == 2) return madd(a, a, p);
  if (base == 3) return madd(a, madd(a, a, p), p);
  if (base == 5) {
    uint64_t b2 = madd(a, a, p);
    b2 = madd(b2, b2, p);
    return madd(a, b2, p);
  }
  return mmul(a, mbase, p, q);
}

// mbase^exp for a 64-bit exponent (common.cl:51-63).  Shift count masked to
// 6 bits as in OpenCL.
RS_HD uint64_t powmod_lg(uint64_t mbase, uint64_t exp, uint64_t p, uint64_t q) {
  uint64_t bit = 0x8000000000000000ull >> ((clz64(exp) + 1) & 63);
  uint64_t a = mbase;
  while (bit) {
    a = mmul(a, a, p, q);
    if (exp & bit) a = mmul(a, mbase, p, q);
    bit >>= 1;
  }
  return a;
}

// base^exp with mbase the Montgomery form of base (common.cl:67-90)
RS_HD uint64_t basepow_lg(uint64_t mbase, uint64_t exp, uint32_t base, uint64_t p, uint64_t q) {
  uint64_t bit = 0x8000000000000000ull >> ((clz64(exp) + 1) & 63);
  uint64_t a = mbase;
  while (bit) {
    a = mmul(a, a, p, q);
    if (exp & bit) a = mul_base(a, mbase, base, p, q);
    bit >>= 1;
  }
  return a;
}

// mbase^exp for a 32-bit exponent (common.cl:93-107)
RS_HD uint64_t powmod_sm(uint64_t mbase, uint32_t exp, uint64_t p, uint64_t q, uint64_t one) {
  if (!exp) return one;
  if (exp == 1) return mbase;
  uint32_t bit = 0x80000000u >> ((clz32(exp) + 1) & 31);
  uint64_t a = mbase;
  while (bit) {
    a = mmul(a, a, p, q);
    if (exp & bit) a = mmul(a, mbase, p, q);
    bit >>= 1;
  }
  return a;
}

// base^exp for a 32-bit exponent (common.cl:110-135)
RS_HD uint64_t basepow_sm(uint64_t mbase, uint32_t exp, uint32_t base, uint64_t p, uint64_t q, uint64_t one) {
  if (!exp) return one;
  if (exp == 1) return mbase;
  uint32_t bit = 0x80000000u >> ((clz32(exp) + 1) & 31);
  uint64_t a = mbase;
  while (bit) {
    a = mmul(a, a, p, q);
    if (exp & bit) a = mul_base(a, mbase, base, p, q);
    bit >>= 1;
  }
  return a;
}

// 2^64 mod p for odd p >= 3 (the reference's '(-p) % p').  The quotient of
// 2^64 / p is estimated in double precision and corrected exactly, which
// avoids a 64-bit software division on the GPU.
RS_HD uint64_t r_mod(uint64_t p) {
  // x = 2^64 - p*floor(2^64/p); work with t = -p*qe (mod 2^64) = 2^64 - p*qe.
  if (p >> 63) return 0 - p;  // 2^64 < 2p
  double inv = 18446744073709551616.0 / (double)p;  // ~2^64/p, relative error < 2^-52
  uint64_t qe = (uint64_t)(inv * (1.0 - 0x1p-50));   // underestimate of floor(2^64/p)
  uint64_t r = 0 - p * qe;                           // 2^64 - p*qe, exact when qe <= floor
  // correction: r = 2^64 - p*qe >= 0 and small multiples of p
  while (r >= p) r -= p;
  return r;
}

// Jacobi symbol (a/n) for odd n > 0 by the binary algorithm: shifts and
// subtractions only, no division.  For prime n and a != 0 (mod n) it equals
// Euler's criterion a^((n-1)/2).
RS_HD int jacobi(uint64_t a, uint64_t n) {
  int t = 1;
  while (a) {
#if defined(__HIP_DEVICE_COMPILE__)
    int z = __ffsll((long long)a) - 1;
#else
    int z = __builtin_ctzll(a);
#endif
    a >>= z;
    if ((z & 1) && ((n & 7) == 3 || (n & 7) == 5)) t = -t;
    if (a < n) {
      uint64_t tmp = a;
      a = n;
      n = tmp;
      if ((a & 3) == 3 && (n & 3) == 3) t = -t;
    }
    a -= n;
  }
  return n == 1 ? t : 0;
}

}  // namespace rs
