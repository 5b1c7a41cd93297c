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
#include <math.h>
#include <string.h>
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

// ---------------------------------------------------------------------------
// Exact modular products in binary64 for odd p < 2^51 (full-rate FP64 FMA on
// gfx950, against quarter/half-rate 32x32->64 integer products).
//
// For integers a, b in [0, p):
//   h = fl(a*b), l = fma(a, b, -h)   h + l = a*b exactly (the rounding error
//                                     of a product is a double, TwoProduct)
//   s = fma(h, pinv, 1.5*2^52)       one rounding of h*pinv + 1.5*2^52 to the
//   qf = s - 1.5*2^52                 integer grid: qf = rint(h*pinv), exactly
//   r = fma(-qf, p, h)               exact: |h - qf*p| < 2^53 (below)
//   t = r + l                        exact: t = a*b - qf*p
// with pinv = fl(1/p).  Error of the quotient: h*pinv = (ab/p)(1+e1)(1+e2),
// |e1|,|e2| <= 2^-53 and ab/p < p, so |h*pinv - ab/p| <= p*(2^-52 + 2^-106),
// which is below 1/2 for p <= 2^51 - 1; rint adds at most 1/2, so
// |ab/p - qf| < 1.  Hence t = ab - qf*p lies in (-p, p), |r| <= |l| + |t| <
// 2^49 + 2^51, and one conditional add of p gives the canonical residue.
// Adding 2^52 to a canonical value c < 2^51 leaves the bits
// 0x4330000000000000 + c, so the integer residue costs one add.
//
// Montgomery residues multiplied by plain residues stay Montgomery residues
// (aR * b = (ab)R mod p), which lets the sieve keep its tables in the
// integer kernels' representation while stepping in binary64.
// ---------------------------------------------------------------------------
constexpr uint64_t kFpLimit = 1ull << 51;
constexpr double kFpTwo52 = 4503599627370496.0;   // 2^52
constexpr double kFpRound = 6755399441055744.0;   // 1.5 * 2^52
constexpr uint64_t kFpBias = 0x4330000000000000ull;  // bits of 2^52

#if defined(__HIP_DEVICE_COMPILE__)
#define RS_FMUL(a, b) __dmul_rn(a, b)
#define RS_FADD(a, b) __dadd_rn(a, b)
#define RS_FMA(a, b, c) __fma_rn(a, b, c)
#define RS_AS_U64(x) ((uint64_t)__double_as_longlong(x))
#else
RS_HD uint64_t rs_as_u64_host(double x) {
  uint64_t u;
  memcpy(&u, &x, 8);
  return u;
}
#define RS_FMUL(a, b) ((double)(a) * (double)(b))
#define RS_FADD(a, b) ((double)(a) + (double)(b))
#define RS_FMA(a, b, c) fma(a, b, c)
#define RS_AS_U64(x) rs_as_u64_host(x)
#endif

// a*b - qf*p in (-p, p) (the derivation above).  The same bounds hold for
// signed operands |a|, |b| < p (|ab| < p^2 is all they use), so chains of
// products may stay in (-p, p) and be made canonical once at the end.
RS_HD double fp_mulmod_signed(double a, double b, double p, double pinv) {
  const double h = RS_FMUL(a, b);
  const double l = RS_FMA(a, b, -h);
  const double qf = RS_FADD(RS_FMA(h, pinv, kFpRound), -kFpRound);
  const double r = RS_FMA(-qf, p, h);
  return RS_FADD(r, l);
}

// a*b mod p, canonical, as a double
RS_HD double fp_mulmod(double a, double b, double p, double pinv) {
  const double t = fp_mulmod_signed(a, b, p, pinv);
  return t < 0.0 ? RS_FADD(t, p) : t;
}

// kFpBias + (a*b mod p): the integer residue plus a constant in the exponent
// field; pc = p + 2^52 (exact for p < 2^51)
RS_HD uint64_t fp_mulmod_bits(double a, double b, double p, double pinv, double pc) {
  const double t = fp_mulmod_signed(a, b, p, pinv);
  return RS_AS_U64(RS_FADD(t, t < 0.0 ? pc : kFpTwo52));
}

// kFpBias + (t mod p) for t in (-p, p); pc = p + 2^52
RS_HD uint64_t fp_signed_bits(double t, double pc) { return RS_AS_U64(RS_FADD(t, t < 0.0 ? pc : kFpTwo52)); }

// integer residue of a canonical double
RS_HD uint64_t fp_to_u64(double c) { return RS_AS_U64(RS_FADD(c, kFpTwo52)) ^ kFpBias; }

// the canonical residue of t in (-p, p)
RS_HD double fp_canon(double t, double p) { return t < 0.0 ? RS_FADD(t, p) : t; }

// x^exp on plain residues in binary64 with the control flow of powmod_lg
// (the same masked shift, so the same effective exponent); x in (-p, p), the
// squarings and products stay signed, the result is canonical
RS_HD double fp_powmod_lg(double x, uint64_t exp, double p, double pinv) {
  uint64_t bit = 0x8000000000000000ull >> ((clz64(exp) + 1) & 63);
  double a = x;
  while (bit) {
    a = fp_mulmod_signed(a, a, p, pinv);
    if (exp & bit) a = fp_mulmod_signed(a, x, p, pinv);
    bit >>= 1;
  }
  return fp_canon(a, p);
}

// x^exp on plain residues in binary64 with the control flow of powmod_sm;
// canonical result for x in (-p, p)
RS_HD double fp_powmod_sm(double x, uint32_t exp, double p, double pinv) {
  if (!exp) return 1.0;
  if (exp == 1) return fp_canon(x, p);
  uint32_t bit = 0x80000000u >> ((clz32(exp) + 1) & 31);
  double a = x;
  while (bit) {
    a = fp_mulmod_signed(a, a, p, pinv);
    if (exp & bit) a = fp_mulmod_signed(a, x, p, pinv);
    bit >>= 1;
  }
  return fp_canon(a, p);
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

// Base-2 strong probable-prime test (getsegprimes.cl:37-66), same shift
// semantics: N = 2^t + 1 squares 64 times before the final doubling; on
// Montgomery residues, for any odd N >= 3.
RS_HD bool sprp2(uint64_t N) {
  const uint64_t q = mont_inv(N);
  const uint64_t one = r_mod(N);
  const uint64_t nmo = N - one;
  const uint64_t two = madd(one, one, N);
  const int t = __builtin_ctzll(N - 1);
  const uint64_t exp = N >> t;
  uint64_t bit = 0x8000000000000000ull >> ((clz64(exp) + 1) & 63);
  uint64_t a = two;
  while (bit) {
    a = mmul(a, a, N, q);
    if (exp & bit) a = madd(a, a, N);
    bit >>= 1;
  }
  if (a == one || a == nmo) return true;
  for (int s = 1; s < t; ++s) {
    a = mmul(a, a, N, q);
    if (a == nmo) return true;
  }
  return false;
}

// The same test on plain residues in binary64 for odd 3 <= N < 2^51: the
// same steps on the same values (squarings exact by fp_mulmod's bound, doublings of values
// in (-N, N) folded back into it), compared in canonical form; a fifth of
// the integer path's instructions per squaring.
RS_HD bool fp_sprp2(uint64_t N) {
  const double Nd = (double)N, ninv = 1.0 / Nd;
  const int t = __builtin_ctzll(N - 1);
  const uint64_t exp = N >> t;
  uint64_t bit = 0x8000000000000000ull >> ((clz64(exp) + 1) & 63);
  double a = 2.0;
  while (bit) {
    a = fp_mulmod_signed(a, a, Nd, ninv);
    if (exp & bit) {
      a = RS_FADD(a, a);
      a = a >= Nd ? RS_FADD(a, -Nd) : (a <= -Nd ? RS_FADD(a, Nd) : a);
    }
    bit >>= 1;
  }
  a = fp_canon(a, Nd);
  const double nmo = RS_FADD(Nd, -1.0);
  if (a == 1.0 || a == nmo) return true;
  for (int s = 1; s < t; ++s) {
    a = fp_mulmod(a, a, Nd, ninv);
    if (a == nmo) return true;
  }
  return false;
}

}  // namespace rs
