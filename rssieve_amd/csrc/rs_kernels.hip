// rs_kernels.hip -- HIP kernels for the k*b^n+-1 BSGS sieve on gfx950.
//
// Work is organised in rounds over a batch of jobs (one job is a batch of
// one).  A round is six launches on one stream that need no host input:
//
//  round_begin  advance each job's device-side segment cursor, emit the
//               primes <= 113 (addsmallprimes.cl:8-27), find the wheel offset
//               (cl_sieve.cpp:453-512)
//  primegen     all base-2 strong probable primes of each job's segment with
//               no prime factor <= 113 (getsegprimes.cl:128-182)
//  setup_pre    one lane per prime: Montgomery constants, the invertibility
//               test and the power-residue tests of the base, the powers of
//               the base the tables need (setup.cl:11-165)
//  setup        one wavefront per prime: batch inversion of the k by
//               wavefront product scans, the power-residue / quadratic-
//               character prefilter of every k (setup.cl:166-209) and
//               per-class compaction (the reference's sort.cl, folded in)
//  round_items  prefix sums of the work items of every job
//  bsgs         persistent workgroups pull (job, prime, class) items from one
//               device queue; the baby steps fill a blocked Bloom filter in
//               LDS, the giant steps probe it, and the rare positives are
//               resolved exactly by a discrete logarithm against every 64th
//               baby value (giant.cl:77-323)
//
// Because the cursors live on the device, the round can be captured once in
// a HIP graph and replayed until every job's range is done.
#include <atomic>
#include <cstddef>
#include <mutex>
#include <type_traits>

#include "../../include/rssieve.h"
#include "rs_internal.h"
#include "rs_mont.h"

namespace rs {

// Loads of data a kernel only reads (written by earlier launches): through
// the constant address space, so that a wavefront-uniform address is one
// scalar load, or the global address space for per-lane addresses (a generic
// pointer would make a flat load that waits on both counters).  Scalar
// types only.
template <class T>
__device__ __forceinline__ T ld_const(const T* p) {
  return *(const __attribute__((address_space(4))) T*)(uintptr_t)p;
}
template <class T>
__device__ __forceinline__ T ld_global(const T* p) {
  return *(const __attribute__((address_space(1))) T*)(uintptr_t)p;
}
__device__ __forceinline__ uint32_t rfl(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ uint64_t rfl64(uint64_t v) { return ((uint64_t)rfl((uint32_t)(v >> 32)) << 32) | rfl((uint32_t)v); }
template <class T>
__device__ __forceinline__ T* rflp(T* p) {
  const uint64_t a = (uint64_t)(uintptr_t)p;
  return (T*)(uintptr_t)(((uint64_t)rfl((uint32_t)(a >> 32)) << 32) | rfl((uint32_t)a));
}

// ---------------------------------------------------------------------------
// Small-prime bit tables: bit i of table_p[r] is set iff p divides r + 2i,
// for the 27 primes 7..113 sieved by the reference generator
// (getsegprimes.cl:74-126).  Computed at compile time from that closed form.
// ---------------------------------------------------------------------------
constexpr uint32_t kSP[27] = {7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59,
                              61, 67, 71, 73, 79, 83, 89, 97, 101, 103, 107, 109, 113};
struct BitTables {
  uint32_t w[27][113];
};
constexpr BitTables make_bit_tables() {
  BitTables t{};
  for (int s = 0; s < 27; s++)
    for (uint32_t r = 0; r < kSP[s]; r++) {
      uint32_t w = 0;
      for (uint32_t i = 0; i < 32; i++)
        if ((r + 2 * i) % kSP[s] == 0) w |= 1u << i;
      t.w[s][r] = w;
    }
  return t;
}
__constant__ BitTables g_bits = make_bit_tables();

// Half the gaps of the mod-30 wheel {4,2,4,2,4,6,2,6}, repeated three times so
// a walk of 16 steps from any start index needs no wrap (getsegprimes.cl:70).
__constant__ int32_t g_wheel3[24] = {2, 1, 2, 1, 2, 3, 1, 3, 2, 1, 2, 1, 2, 3, 1, 3, 2, 1, 2, 1, 2, 3, 1, 3};

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Start a job's range [pmin, pmax) with a given BSGS workgroup size.
__global__ void k_set_cursor(SegJob* __restrict__ sj, uint64_t pmin, uint64_t pmax, uint32_t threads, uint64_t range,
                             uint32_t gen_blocks) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    sj->p_cur = pmin;
    sj->p_end = pmax;
    sj->threads = threads;
    sj->range = range;
    sj->gen_blocks = gen_blocks;
    sj->nprimes = 0;
    sj->low = sj->high = 0;
  }
}

// First N >= p coprime to 30 and its index in the wheel that starts at 7
// (cl_sieve.cpp:453-512); returns false if N would pass 2^64.
__device__ bool wheel_start(uint64_t p, uint64_t* N, int32_t* idx) {
  // wheel position i starts at residue 7, 11, 13, 17, 19, 23, 29, 1 (mod 30)
  const int8_t pos_of[30] = {-1, 7, -1, -1, -1, -1, -1, 0, -1, -1, -1, 1, -1, 2, -1,
                             -1, -1, 3, -1, 4, -1, -1, -1, 5, -1, -1, -1, -1, -1, 6};
  uint64_t n = p;
  for (int i = 0; i < 7; i++) {
    const int k = pos_of[n % 30];
    if (k >= 0) {
      *N = n;
      *idx = k;
      return true;
    }
    if (n + 1 < n) return false;
    n++;
  }
  return false;
}

// One thread per job: next segment of each job (cl_sieve.cpp:1291-1320).
__global__ void k_round_begin(SegJob* __restrict__ jobs, BatchState* __restrict__ bs, uint32_t njobs) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j == 0) {
    bs->work_next = 0;
    bs->njobs = njobs;
  }
  if (j >= njobs) return;
  SegJob& J = jobs[j];
  J.nprimes = 0;
  J.low = J.high = 0;
  if (J.p_cur >= J.p_end) return;
  uint64_t p = J.p_cur;
  uint64_t stop = p + J.range;
  if (stop > J.p_end || stop < p) stop = J.p_end;
  if (p < 114) {
    const uint32_t small[30] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47,
                                53, 59, 61, 67, 71, 73, 79, 83, 89, 97, 101, 103, 107, 109, 113};
    const uint64_t stop_sm = stop > 114 ? 114 : stop;
    uint32_t n = 0;
    for (int i = 0; i < 30; i++)
      if (small[i] >= p && small[i] < stop_sm) {
        if (n < J.cap) J.primes[n] = small[i];
        n++;
      }
    J.nprimes = n;
    p = stop_sm;
  }
  if (p < stop) {
    uint64_t N;
    int32_t widx;
    if (wheel_start(p, &N, &widx) && N < stop) {
      J.low = N;
      J.high = stop;
      J.widx = widx;
    }
  }
  J.p_cur = stop;
}

// ---------------------------------------------------------------------------
// primegen: work-item gid covers [low + 60 gid, low + 60 gid + 60) with the
// mod-30 wheel; survivors of the small-prime bit sieve are packed in LDS, then
// every lane of the workgroup tests them; passing primes are appended with one
// global atomic per wavefront.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_primegen(SegJob* __restrict__ jobs) {
  SegJob& J = jobs[blockIdx.y];
  const uint64_t low = J.low, high = J.high;
  if (blockIdx.x >= J.gen_blocks || high <= low) return;
  const int32_t widx = J.widx;
  uint64_t* __restrict__ out = J.primes;
  const uint32_t cap = J.cap;
  __shared__ uint64_t cand[256 * 16];
  __shared__ uint32_t ncand;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) ncand = 0;
  __syncthreads();
  const uint64_t gid = (uint64_t)blockIdx.x * 256 + tid;
  uint64_t P = low + gid * 60;
  if (P < low) P = high;  // past 2^64
  uint64_t end = P + 60;
  if (end > high || end < P) end = high;
  if (P < end) {
    uint32_t bs = 0;
#pragma unroll
    for (int s = 0; s < 27; s++) bs |= g_bits.w[s][P % kSP[s]];
    int idx = widx;
    while (P < end) {
      if ((bs & 1u) == 0) {
        uint32_t pos = atomicAdd(&ncand, 1u);
        cand[pos] = P;
      }
      const int inc = g_wheel3[idx++];
      uint64_t nP = P + 2u * (uint32_t)inc;
      bs >>= inc;
      if (nP < P) nP = end;  // past 2^64
      P = nP;
    }
  }
  __syncthreads();
  const uint32_t n = ncand;
  const uint32_t lane = tid & 63;
  for (uint32_t b0 = 0; b0 < n; b0 += 256) {
    const uint32_t pos = b0 + tid;
    uint64_t N = 0;
    bool ok = false;
    if (pos < n) {
      N = cand[pos];
      ok = N < kFpLimit ? fp_sprp2(N) : sprp2(N);  // (candidates are odd)
    }
    const uint64_t m = __ballot(ok);
    if (m) {
      const uint32_t leader = __ffsll((long long)m) - 1;
      uint32_t wbase = 0;
      if (lane == leader) wbase = atomicAdd(&J.nprimes, (uint32_t)__popcll(m));
      wbase = __shfl(wbase, leader);
      if (ok) {
        const uint32_t w = wbase + lane_rank(m);
        if (w < cap) out[w] = N;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// setup: one wavefront per prime.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t shfl_up64(uint64_t v, int d) { return __shfl_up(v, d, 64); }
__device__ __forceinline__ uint64_t shfl_dn64(uint64_t v, int d) { return __shfl_down(v, d, 64); }
__device__ __forceinline__ uint64_t shfl64(uint64_t v, int l) { return __shfl(v, l, 64); }
__device__ __forceinline__ double shfl_up64(double v, int d) { return __shfl_up(v, d, 64); }
__device__ __forceinline__ double shfl_dn64(double v, int d) { return __shfl_down(v, d, 64); }
__device__ __forceinline__ double shfl64(double v, int l) { return __shfl(v, l, 64); }

// Two powmods with one shared 64-bit exponent, interleaved for ILP; same
// masked-shift semantics as powmod_lg.
__device__ __forceinline__ void powmod_lg2(uint64_t x0, uint64_t x1, uint64_t exp, uint64_t p, uint64_t q,
                                           uint64_t* r0, uint64_t* r1) {
  uint64_t bit = 0x8000000000000000ull >> ((clz64(exp) + 1) & 63);
  uint64_t a0 = x0, a1 = x1;
  while (bit) {
    a0 = mmul(a0, a0, p, q);
    a1 = mmul(a1, a1, p, q);
    if (exp & bit) {
      a0 = mmul(a0, x0, p, q);
      a1 = mmul(a1, x1, p, q);
    }
    bit >>= 1;
  }
  *r0 = a0;
  *r1 = a1;
}

// powmod_lg2 on plain residues in binary64 (signed intermediate products,
// canonical results)
__device__ __forceinline__ void fp_powmod_lg2(double x0, double x1, uint64_t exp, double p, double pinv, double* r0,
                                              double* r1) {
  uint64_t bit = 0x8000000000000000ull >> ((clz64(exp) + 1) & 63);
  double a0 = x0, a1 = x1;
  while (bit) {
    a0 = fp_mulmod_signed(a0, a0, p, pinv);
    a1 = fp_mulmod_signed(a1, a1, p, pinv);
    if (exp & bit) {
      a0 = fp_mulmod_signed(a0, x0, p, pinv);
      a1 = fp_mulmod_signed(a1, x1, p, pinv);
    }
    bit >>= 1;
  }
  *r0 = fp_canon(a0, p);
  *r1 = fp_canon(a1, p);
}

// d of the power-residue tests (setup.cl:22-48)
__constant__ uint32_t g_prd[12] = {3, 4, 5, 7, 8, 9, 11, 13, 17, 19, 23, 29};

// Residue classes of n (algo 2).  With L = gcd(p - 1, 720720) (720720 =
// 2^4 * 3^2 * 5 * 7 * 11 * 13) and z = b^((p-1)/L), every solution of b^n = h
// (mod p) satisfies z^n = y with y = h^((p-1)/L).  Split L = La * Lb * Lc
// with La = gcd(p - 1, 80), Lb = gcd(p - 1, 63) and Lc = gcd(p - 1, 143): for
// za = z^(Lb*Lc) of order sa (the first i > 0 with za^i = 1) the powers za^i,
// i < sa, are distinct, so ya = y^(Lb*Lc) = za^xa fixes n = xa (mod sa), and
// likewise n = xb (mod sb) from zb = z^(La*Lc), yb = y^(La*Lc), and n = xc
// (mod sc) from zc = z^(La*Lb), yc = y^(La*Lb); when sa, sb and sc are
// pairwise coprime (always for a prime p) the three combine by CRT into
// n mod s, s = sa*sb*sc = ord(z), and a y that is no power of z leaves no n
// at all.  This is algebra mod p, true for a composite 2-PRP as well (whose
// non-coprime orders fall back to the parity split): the reference's tests
// (parity classes, power residues, setup.cl:59-87) still decide which k are
// searched, and the classes only narrow the n they are searched on, so the
// solutions found are the same.  For a prime p with b a quadratic
// non-residue s is even and fixes the parity the reference's class gives.
// A fourth part Ld, one prime r in [17, 1021] that divides p - 1, with zd =
// b^((p-1)/Ld) != 1 and zd^Ld = 1 (of order r, always so for a prime p) and
// yd = h^((p-1)/Ld), narrows n further mod r (the largest such r whose stride
// stays below 2^24, the shifts' width in the BSGS queue; 1 when none):
// L = La*Lb*Lc*Ld.  Its discrete logarithm is a small baby-step giant-step:
// at most 128 powers zd^i in the table and ceil(Ld / 128) giant steps.
constexpr int kNRd = 166;
constexpr uint32_t kRd[kNRd] = {
    17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97, 101, 103, 107, 109, 113,
    127, 131, 137, 139, 149, 151, 157, 163, 167, 173, 179, 181, 191, 193, 197, 199, 211, 223, 227, 229,
    233, 239, 241, 251, 257, 263, 269, 271, 277, 281, 283, 293, 307, 311, 313, 317, 331, 337, 347, 349,
    353, 359, 367, 373, 379, 383, 389, 397, 401, 409, 419, 421, 431, 433, 439, 443, 449, 457, 461, 463,
    467, 479, 487, 491, 499, 503, 509, 521, 523, 541, 547, 557, 563, 569, 571, 577, 587, 593, 599, 601,
    607, 613, 617, 619, 631, 641, 643, 647, 653, 659, 661, 673, 677, 683, 691, 701, 709, 719, 727, 733,
    739, 743, 751, 757, 761, 769, 773, 787, 797, 809, 811, 821, 823, 827, 829, 839, 853, 857, 859, 863,
    877, 881, 883, 887, 907, 911, 919, 929, 937, 941, 947, 953, 967, 971, 977, 983, 991, 997, 1009, 1013,
    1019, 1021};
__constant__ uint32_t g_rd[kNRd] = {
    17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97, 101, 103, 107, 109, 113,
    127, 131, 137, 139, 149, 151, 157, 163, 167, 173, 179, 181, 191, 193, 197, 199, 211, 223, 227, 229,
    233, 239, 241, 251, 257, 263, 269, 271, 277, 281, 283, 293, 307, 311, 313, 317, 331, 337, 347, 349,
    353, 359, 367, 373, 379, 383, 389, 397, 401, 409, 419, 421, 431, 433, 439, 443, 449, 457, 461, 463,
    467, 479, 487, 491, 499, 503, 509, 521, 523, 541, 547, 557, 563, 569, 571, 577, 587, 593, 599, 601,
    607, 613, 617, 619, 631, 641, 643, 647, 653, 659, 661, 673, 677, 683, 691, 701, 709, 719, 727, 733,
    739, 743, 751, 757, 761, 769, 773, 787, 797, 809, 811, 821, 823, 827, 829, 839, 853, 857, 859, 863,
    877, 881, 883, 887, 907, 911, 919, 929, 937, 941, 947, 953, 967, 971, 977, 983, 991, 997, 1009, 1013,
    1019, 1021};
constexpr uint32_t kZV = 464;                 // za^i at [0, 81), zb^i at [128, 192), zc^i at [192, 336),
                                              // zd^i at [336, 464)
constexpr uint32_t kZH = 896;                 // their 16-bit hash slots: [0, 256), [256, 384), [384, 640),
                                              // [640, 896), two per 32-bit word
constexpr uint64_t kClassMinP = 1u << 20;     // smaller p keep the merged table's parity split; above it
                                              // the exponent (p-1)/L is at least 2 (L <= 720720)

// x mod m for x < 2^22 and 1 <= m < 2^22 without an integer division: the
// quotient from a float reciprocal (relative error below 2^-22, so off by at
// most one for quotients below 2^22) and one correction of the remainder
struct SmallMod {
  uint32_t m;
  float inv;
  __device__ explicit SmallMod(uint32_t mm) : m(mm), inv(__builtin_amdgcn_rcpf((float)mm)) {}
  __device__ uint32_t operator()(uint32_t x) const {
    const uint32_t qe = (uint32_t)((float)x * inv);
    int r = (int)x - (int)__umul24(qe, m);
    if (r < 0) r += (int)m;
    else if (r >= (int)m) r -= (int)m;
    return (uint32_t)r;
  }
};

// a^-1 mod m for gcd(a, m) = 1 and m < 2^11, by every lane of the wavefront
// trying its candidates at once (call with all lanes active); 0 for m = 1
__device__ __forceinline__ uint32_t inv_mod_lanes(uint32_t a, const SmallMod& M, uint32_t lane) {
  const uint32_t am = M(a);
  for (uint32_t t0 = 0; t0 < M.m; t0 += 64) {
    const uint32_t t = t0 + lane;
    const uint64_t hit = __ballot(t < M.m && M(am * t) == 1u);
    if (hit) return t0 + (uint32_t)__ffsll((long long)hit) - 1;
  }
  return 0;
}

// 1/d of the power-residue tests' d (setup.cl:22-48), for exact quotients of
// p - 1 in binary64
__constant__ double g_prdinv[12] = {1.0 / 3, 1.0 / 4, 1.0 / 5, 1.0 / 7, 1.0 / 8, 1.0 / 9,
                                    1.0 / 11, 1.0 / 13, 1.0 / 17, 1.0 / 19, 1.0 / 23, 1.0 / 29};

// LDS operations of one wavefront become visible to its other lanes
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// The residue arithmetic of the setup, in two representations with one
// interface.  Every value the setup computes is a canonical residue, so both
// give the same stored records bit for bit; only the cost differs.
//  MontArith: Montgomery residues (x*2^64 mod p) in 64-bit integers, any odd p
//             (the reference's own representation, common.cl:24-176).
//  FpArith:   plain residues in binary64 for odd p < 2^51 (exact products,
//             rs_mont.h), converted to Montgomery form only when stored.
struct MontArith {
  using V = uint64_t;
  uint64_t p, q, one_, r2;
  __device__ V one() const { return one_; }
  __device__ V neg_one() const { return p - one_; }
  __device__ V neg(V x) const { return x ? p - x : 0; }  // -x for a canonical x
  __device__ V mul(V a, V b) const { return mmul(a, b, p, q); }
  __device__ V from_int(uint64_t k) const { return mmul(k, r2, p, q); }  // k < 2^64
  __device__ V base(uint32_t b) const {
    const uint64_t two = madd(one_, one_, p);
    if (b == 2) return two;
    if (b == 3) return madd(one_, two, p);
    if (b == 5) return madd(madd(two, two, p), one_, p);
    return from_int(b);
  }
  __device__ V base_pow(V mb, uint32_t e, uint32_t b) const { return basepow_sm(mb, e, b, p, q, one_); }
  __device__ V pow_lg(V x, uint64_t e) const { return powmod_lg(x, e, p, q); }
  __device__ V pow_sm(V x, uint32_t e) const { return powmod_sm(x, e, p, q, one_); }
  // x^e for any 64-bit e (powmod_lg is exact from e = 2 on)
  __device__ V pow_u64(V x, uint64_t e) const { return e <= 1 ? (e ? x : one_) : powmod_lg(x, e, p, q); }
  __device__ void pow_lg2(V x0, V x1, uint64_t e, V* r0, V* r1) const { powmod_lg2(x0, x1, e, p, q, r0, r1); }
  // d | x, and the quotient (any 64-bit x)
  __device__ bool divides(uint64_t x, uint32_t d, double, uint64_t* quot) const {
    *quot = x / d;
    return *quot * d == x;
  }
  __device__ uint64_t mont(V x) const { return x; }
  __device__ uint64_t bits(V x) const { return x; }  // a canonical integer key
  __device__ uint64_t store(V x) const { return x; }  // a value as staged between the setup kernels
  __device__ V load(uint64_t u) const { return u; }
  __device__ V from_bits(uint64_t u) const { return u; }  // the inverse of bits()
};

struct FpArith {
  using V = double;
  double p, pinv, R;  // R = 2^64 mod p
  __device__ V one() const { return 1.0; }
  __device__ V neg_one() const { return p - 1.0; }
  __device__ V mul(V a, V b) const { return fp_mulmod(a, b, p, pinv); }
  // k mod p.  For k < 2^53 the product k*1 meets the bounds of fp_mulmod
  // (the quotient estimate is off by less than 1 while k/p < 2^51); larger k
  // (a sign-extended k in the k >= 2^31 compatibility mode) divide exactly.
  __device__ V from_int(uint64_t k) const {
    if (k >> 53) return (double)(k % (uint64_t)p);
    const double kd = (double)k;
    return kd < p ? kd : fp_mulmod(kd, 1.0, p, pinv);  // (k < p, nearly always: k itself)
  }
  __device__ V neg(V x) const { return x != 0.0 ? p - x : 0.0; }  // -x for a canonical x
  __device__ V base(uint32_t b) const { return from_int(b); }
  __device__ V base_pow(V mb, uint32_t e, uint32_t) const { return fp_powmod_sm(mb, e, p, pinv); }
  __device__ V pow_lg(V x, uint64_t e) const { return fp_powmod_lg(x, e, p, pinv); }
  __device__ V pow_sm(V x, uint32_t e) const { return fp_powmod_sm(x, e, p, pinv); }
  __device__ V pow_u64(V x, uint64_t e) const { return e <= 1 ? (e ? x : 1.0) : fp_powmod_lg(x, e, p, pinv); }
  __device__ void pow_lg2(V x0, V x1, uint64_t e, V* r0, V* r1) const { fp_powmod_lg2(x0, x1, e, p, pinv, r0, r1); }
  // d | x, and the quotient, for x < 2^51 (x = p - 1) and dinv = fl(1/d):
  // rint(x * dinv) is the quotient when d divides x (its relative error is
  // below 2^-52 and the quotient below 2^51), and fma(-q, d, x), exact, is
  // zero exactly when it does
  __device__ bool divides(uint64_t x, uint32_t d, double dinv, uint64_t* quot) const {
    const double xd = (double)x, qd = rint(xd * dinv);
    *quot = (uint64_t)qd;
    return __fma_rn(-qd, (double)d, xd) == 0.0;
  }
  __device__ uint64_t mont(V x) const { return fp_to_u64(fp_mulmod(x, R, p, pinv)); }
  __device__ uint64_t bits(V x) const { return fp_to_u64(x); }
  __device__ uint64_t store(V x) const { return (uint64_t)__double_as_longlong(x); }
  __device__ V load(uint64_t u) const { return __longlong_as_double((long long)u); }
  __device__ V from_bits(uint64_t u) const { return (double)u; }  // the inverse of bits() (u < 2^51)
};

// ---------------------------------------------------------------------------
// Setup, first kernel (k_setup_pre): the values that depend on the prime and
// the job's constants but on no single k, one lane per prime.  The second
// kernel (k_setup, one wavefront per prime) spends its lanes on the k; these
// per-prime exponentiations in its lanes would keep most of them idle (an
// exponent per lane for 18 of 64 lanes).  The values are staged in the
// prime's own record slot, which k_setup overwrites with the final record.
// ---------------------------------------------------------------------------
struct PreRec {
  uint64_t p, q, one;
  uint64_t v[8];             // residues as the arithmetic stages them (store / load), slots kPv*
  uint32_t nk[RS_NCLASS];    // 0, 0, 0: the final record's "no work" for a prime skipped here
  uint32_t info;             // kPreSkip | passing power-residue tests (bit i: g_prd[i]) << 1 |
                             // (quadratic character of b + 1) << 13
  uint32_t La, Lb, Lc, Ld;   // gcd(p - 1, 80), gcd(p - 1, 63), gcd(p - 1, 143) and the fourth part
                             // (residue-class primes)
  uint64_t pad;
};
static_assert(sizeof(PreRec) == sizeof(PrimeRec), "PreRec fills a PrimeRec slot");
static_assert(offsetof(PreRec, nk) == offsetof(PrimeRec, nk), "a skipped prime's staged nk is its final nk");
constexpr uint32_t kPreSkip = 1u;
// slots of PreRec::v: (prod of the k)^-1, b^-1, the first baby value (b^NMIN,
// b^-NMIN for residue classes); then for residue classes za, zb, zc, zd,
// otherwise b^-Q, b^-QT, b^-2QT, b^T, b^2T (a small prime of a residue-class
// job: b^-2Q, b^-2QT for its parity split)
enum { kPvInvK = 0, kPvBinv = 1, kPvBn = 2, kPvZa = 3, kPvZb = 4, kPvZc = 5, kPvZd = 6,
       kPvG0 = 3, kPvGT0 = 4, kPvGT1 = 5, kPvBT0 = 6, kPvBT1 = 7 };

// gcd(p - 1, 80), gcd(p - 1, 63), gcd(p - 1, 143)
template <class A>
__device__ __forceinline__ void class_moduli(const A& ar, uint64_t pm, uint32_t* La, uint32_t* Lb, uint32_t* Lc) {
  uint64_t qt;
  const uint32_t v2 = (uint32_t)__builtin_ctzll(pm);
  *La = (1u << (v2 < 4 ? v2 : 4)) * (ar.divides(pm, 5, 1.0 / 5, &qt) ? 5u : 1u);
  *Lb = (ar.divides(pm, 9, 1.0 / 9, &qt) ? 9u : ar.divides(pm, 3, 1.0 / 3, &qt) ? 3u : 1u) *
        (ar.divides(pm, 7, 1.0 / 7, &qt) ? 7u : 1u);
  *Lc = (ar.divides(pm, 11, 1.0 / 11, &qt) ? 11u : 1u) * (ar.divides(pm, 13, 1.0 / 13, &qt) ? 13u : 1u);
}

template <class A>
__device__ __forceinline__ void pre_prime(const A& ar, const SegJob& J, uint64_t p, uint64_t q, uint64_t one,
                                          PreRec* __restrict__ out) {
  using V = typename A::V;
  const JobDev* __restrict__ job = J.job;
  const uint32_t base = ld_const(&job->base);
  const int K = ld_const(&job->kcount);
  const uint32_t algo = ld_const(&job->algo);
  const uint32_t T = J.threads;
  const V vone = ar.one(), nmo = ar.neg_one();
  const V mbase = ar.base(base);
  PreRec o;
  o.p = p;
  o.q = q;
  o.one = one;
  o.nk[0] = o.nk[1] = o.nk[2] = 0;
  o.La = 2;
  o.Lb = o.Lc = o.Ld = 1;
  o.pad = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) o.v[i] = 0;
  // The invertibility test (setup.cl:137-162): total = (prod of the k) *
  // b^Qref, Qref being the reference's own split for this n-range (for a
  // composite 2-PRP the outcome depends on that exponent, so every BSGS shape
  // uses the same one), against its Fermat inverse total^(p-2).  The product
  // of the k in two chains (a canonical residue, whatever the order).
  V P0 = vone, P1 = vone;
  int i = 0;
  for (; i + 1 < K; i += 2) {
    P0 = ar.mul(P0, ar.from_int(ld_const(&job->ak[i])));
    P1 = ar.mul(P1, ar.from_int(ld_const(&job->ak[i + 1])));
  }
  if (i < K) P0 = ar.mul(P0, ar.from_int(ld_const(&job->ak[i])));
  const V P = ar.mul(P0, P1);
  const V bQ1 = ar.base_pow(mbase, ld_const(&job->Qref) - 1, base);  // b^(Qref-1), for b^-1 below
  const V gQ = ar.mul(bQ1, mbase);
  const V total = ar.mul(P, gQ);
  const V inv_total = ar.pow_lg(total, p - 2);  // setup.cl:154
  if (ar.mul(total, inv_total) != vone) {       // setup.cl:156-162
    o.info = kPreSkip;
    *out = o;
    atomicAdd((unsigned long long*)&J.cnt->skipped, 1ull);
    return;
  }
  const V binv = ar.mul(bQ1, ar.mul(inv_total, P));  // b^(Qref-1) * b^-Qref (setup.cl:164-165)
  o.v[kPvInvK] = ar.store(ar.mul(inv_total, gQ));
  o.v[kPvBinv] = ar.store(binv);
  // the first baby value: b^NMIN, or b^-NMIN for a residue-class table (which walks b^-n)
  const bool split = algo == 2 && p > kClassMinP;
  o.v[kPvBn] = ar.store(ar.pow_u64(split ? binv : mbase, ld_const(&job->nmin)));

  // The power-residue tests b^((p-1)/d) = 1 for d | p - 1 (setup.cl:22-48)
  // and the quadratic character b^((p-1)/2) (setup.cl:51-53).
  const uint64_t pm = p - 1;
  uint32_t dmask = 0;
  V rg;
  if (p > kClassMinP) {
    // With L = gcd(p - 1, 720720) and z = b^((p-1)/L), b^((p-1)/d) = z^(L/d)
    // for every d of the tests that divides 720720 (d | p - 1 <=> d | L);
    // for the primes r of kRd (17, 19, 23 and 29 among them), w = b^((p-1)/D)
    // with D the product of those that divide p - 1 gives b^((p-1)/r) =
    // w^(D/r).  The same residues as the direct exponents, in any ring:
    // (p-1)/L is at least 2 here (L <= 720720), and so is (p-1)/D (D is odd),
    // where pow_lg's masked shift is exact.
    uint32_t La, Lb, Lc;
    class_moduli(ar, pm, &La, &Lb, &Lc);
    const uint32_t L = La * Lb * Lc;
    uint64_t exL;
    ar.divides(pm, L, 1.0 / (double)L, &exL);
    const V z = ar.pow_lg(mbase, exL);
    // D: the primes r of kRd that divide p - 1 (17, 19, 23 and 29 first,
    // the reference's tests; the others while D stays below 2^32)
    uint32_t D = 1;
    uint32_t dm[(kNRd + 31) / 32] = {};  // bit j % 32 of dm[j / 32]: kRd[j] divides D
#pragma unroll
    for (int j = 0; j < kNRd; j++) {
      uint64_t qt;
      if (ar.divides(pm, kRd[j], 1.0 / kRd[j], &qt) && (j < 4 || (uint64_t)D * kRd[j] < (1ull << 32))) {
        D *= kRd[j];
        dm[j >> 5] |= 1u << (j & 31);
      }
    }
    V w = vone;
    if (D > 1) {
      uint64_t exD;
      ar.divides(pm, D, 1.0 / (double)D, &exD);
      w = ar.pow_lg(mbase, exD);
    }
    constexpr uint32_t kPrd8[8] = {3, 4, 5, 7, 8, 9, 11, 13};  // the tests' d that divide 720720
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (L % kPrd8[j] == 0 && ar.pow_u64(z, L / kPrd8[j]) == vone) dmask |= 1u << j;
    // b^((p-1)/r) = w^(D/r) for r | D: the tests of 17..29, and the fourth
    // class part (the largest r with b^((p-1)/r) != 1 and 2*L*r < 2^24)
    uint32_t Ld = 1;
    V zd = vone;
#pragma unroll
    for (int wi = 0; wi < (kNRd + 31) / 32; wi++) {
      for (uint32_t m = dm[wi]; m; m &= m - 1) {
        const int j = 32 * wi + __ffs((int)m) - 1;
        const uint32_t r = g_rd[j];
        const V t = ar.pow_u64(w, D / r);
        if (t == vone) {
          if (j < 4) dmask |= 1u << (8 + j);
        } else if (2ull * L * r < (1ull << 24)) {
          Ld = r;
          zd = t;
        }
      }
    }
    rg = ar.pow_u64(z, L / 2);
    if (split) {
      // the bases of the residue-class tables: za = z^(Lb*Lc) = b^((p-1)/La),
      // zb = z^(La*Lc), zc = z^(La*Lb) (1 when Lc = 1), zd = b^((p-1)/Ld)
      o.La = La;
      o.Lb = Lb;
      o.Lc = Lc;
      o.Ld = Ld;
      o.v[kPvZa] = ar.store(ar.pow_u64(z, Lb * Lc));
      o.v[kPvZb] = ar.store(ar.pow_u64(z, La * Lc));
      o.v[kPvZc] = ar.store(Lc > 1 ? ar.pow_u64(z, La * Lb) : vone);
      o.v[kPvZd] = ar.store(zd);
    }
  } else {
    // small p: every exponent directly, as the reference takes it (an
    // exponent of 1 gives pow_lg's masked-shift value, common.cl:52-53)
#pragma unroll
    for (int j = 0; j < 12; j++) {
      uint64_t e;
      if (ar.divides(pm, g_prd[j], g_prdinv[j], &e)) {
        const V r = e ? ar.pow_lg(mbase, e) : vone;
        if (r == vone) dmask |= 1u << j;
      }
    }
    const uint64_t e2 = pm >> 1;
    rg = e2 ? ar.pow_lg(mbase, e2) : vone;
  }
  const int resg = (rg == vone) - (rg == nmo);
  o.info = dmask << 1 | (uint32_t)(resg + 1) << 13;
  if (!split) {
    // the constants of the reference's tables and of the merged table:
    // b^T and b^2T (baby strides of a workgroup round), b^-Q (giant step)
    // and its T-th and 2T-th powers
    const uint32_t Q = ld_const(&job->Q);
    o.v[kPvBT0] = ar.store(ar.pow_sm(mbase, T));
    o.v[kPvBT1] = ar.store(ar.pow_sm(mbase, 2 * T));
    if (algo == 2) {  // a small prime of a residue-class job: the parity split, table span 2Q
      o.v[kPvG0] = ar.store(ar.pow_sm(binv, 2 * Q));
      o.v[kPvGT0] = ar.store(ar.pow_sm(binv, 2 * Q * T));
    } else {
      o.v[kPvG0] = ar.store(ar.pow_sm(binv, Q));
      o.v[kPvGT0] = ar.store(ar.pow_sm(binv, Q * T));
      o.v[kPvGT1] = ar.store(ar.pow_sm(binv, 2 * Q * T));
    }
  }
  *out = o;
}

// one lane per prime; every job of the batch on its own blockIdx.y
__global__ __launch_bounds__(256) void k_setup_pre(SegJob* __restrict__ jobs) {
  const SegJob& J = jobs[blockIdx.y];
  Counters* __restrict__ cnt = J.cnt;
  const uint32_t cap = J.cap;
  const uint32_t ngen = J.nprimes;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    atomicAdd((unsigned long long*)&cnt->primecount, (unsigned long long)ngen);
    atomicMax((unsigned long long*)&cnt->max_seg_primes, (unsigned long long)ngen);
    if (ngen > cap) atomicAdd((unsigned long long*)&cnt->seg_overflow, 1ull);
  }
  const uint32_t np = ngen < cap ? ngen : cap;
  const uint32_t pi = blockIdx.x * 256 + threadIdx.x;
  if (pi >= np) return;
  const uint64_t p = J.primes[pi];
  const uint64_t q = mont_inv(p);
  const uint64_t one = r_mod(p);
  PreRec* out = reinterpret_cast<PreRec*>(J.recs + pi);
  if ((p & 1) && p < kFpLimit) {
    FpArith ar;
    ar.p = (double)p;
    ar.pinv = 1.0 / ar.p;
    ar.R = (double)one;
    pre_prime(ar, J, p, q, one, out);
  } else {
    MontArith ar;
    ar.p = p;
    ar.q = q;
    ar.one_ = one;
    uint64_t r2 = madd(madd(one, one, p), madd(one, one, p), p);
#pragma unroll
    for (int i = 0; i < 5; i++) r2 = mmul(r2, r2, p, q);  // 4^(2^5) = 2^64: Montgomery form of R
    ar.r2 = r2;
    pre_prime(ar, J, p, q, one, out);
  }
}

// The setup of prime pre.p = primes[pi] by one wavefront (setup.cl:89-209 and
// sort.cl:11-104) from the values k_setup_pre staged; zv and zh are the
// wavefront's LDS for the z^i table.
template <class A>
__device__ __forceinline__ void setup_prime(const A& ar, const SegJob& J, const PreRec& pre, uint32_t pi,
                                            uint32_t lane, uint64_t* zv, uint32_t* zh) {
  using V = typename A::V;
  const JobDev* __restrict__ job = J.job;
  PrimeRec* __restrict__ R = J.recs + pi;
  Counters* __restrict__ cnt = J.cnt;
  const uint64_t p = pre.p, q = pre.q, one = pre.one;
  const uint32_t base = ld_const(&job->base);
  const int K = ld_const(&job->kcount);
  const uint32_t T = J.threads;
  const uint32_t algo = ld_const(&job->algo);
  const V vone = ar.one(), nmo = ar.neg_one();
  const V mbase = ar.base(base);
  // x^lane in every lane: an inclusive product scan of (lane ? x : 1), six
  // products instead of an exponentiation with a per-lane exponent
  auto lane_pow = [&](V x) {
    V v = lane ? x : vone;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const V a = shfl_up64(v, d);
      if ((int)lane >= d) v = ar.mul(v, a);
    }
    return v;
  };

  // k for k = lane and k = lane + 64; 'one' pads the products.
  const V mk0 = (int)lane < K ? ar.from_int(ld_global(&job->ak[lane])) : vone;
  const V mk1 = (int)lane + 64 < K ? ar.from_int(ld_global(&job->ak[lane + 64])) : vone;
  const bool act0 = (int)lane < K, act1 = (int)lane + 64 < K;
  const bool cp0 = act0 && ld_global(&job->cpos[lane]), cp1 = act1 && ld_global(&job->cpos[lane + 64]);

  // the values of k_setup_pre: the prime passed the invertibility test
  const V binv = ar.load(pre.v[kPvBinv]);   // b^-1
  const uint64_t dmask = (pre.info >> 1) & 0xFFFu;  // passing d tests, bit i = g_prd[i]
  const int resg = (int)((pre.info >> 13) & 3u) - 1;

  // residue classes: L = La * Lb * Lc * Ld (gcd(p - 1, 80), gcd(p - 1, 63),
  // gcd(p - 1, 143) and the fourth part)
  const uint64_t pm = p - 1;
  const bool split = algo == 2 && p > kClassMinP;
  uint32_t L = 2, La = 2, Lb = 1, Lc = 1, Ld = 1;
  uint64_t exL = 0;
  if (split) {
    La = rfl(pre.La);
    Lb = rfl(pre.Lb);
    Lc = rfl(pre.Lc);
    Ld = rfl(pre.Ld);
    L = La * Lb * Lc * Ld;
    ar.divides(pm, L, 1.0 / (double)L, &exL);
    exL = rfl64(exL);  // wavefront-uniform: the exponentiation loops branch on scalars
  }

  // The target h of a k (setup.cl:180-188): k^-1 for c = -1, -k^-1 for
  // c = +1.  A residue-class table runs on the inverse instead: its baby
  // steps walk b^-(nstart + S*e) and its giant steps multiply by b^(S*E), so
  // a solution b^n = h, n = nstart + d + S*(e + E*q), meets at
  // h^-1 * b^d * b^(S*E*q) = b^-(nstart + S*e), and h^-1 = -c*k needs no
  // inversion.  The characters of h and h^-1 agree (a power of h is 1 or -1
  // exactly when the same power of h^-1 is), and the class of h^-1 is the
  // negated class of h.  Below, h0 and h1 are h^-1 on that path and h on
  // the others.
  V h0, h1;
  if (split) {
    h0 = cp0 ? ar.neg(mk0) : mk0;
    h1 = cp1 ? ar.neg(mk1) : mk1;
  } else {
    // batch inversion by inclusive prefix (s) and suffix (u) products over the wavefront
    V s0 = mk0, s1 = mk1, u0 = mk0, u1 = mk1;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const V a0 = shfl_up64(s0, d), a1 = shfl_up64(s1, d);
      const V b0 = shfl_dn64(u0, d), b1 = shfl_dn64(u1, d);
      if ((int)lane >= d) {
        s0 = ar.mul(s0, a0);
        s1 = ar.mul(s1, a1);
      }
      if ((int)lane + d < 64) {
        u0 = ar.mul(u0, b0);
        u1 = ar.mul(u1, b1);
      }
    }
    const V A0 = shfl64(s0, 63), B0 = shfl64(s1, 63);
    const V ep0 = shfl_up64(s0, 1), ep1 = shfl_up64(s1, 1);
    const V es0 = shfl_dn64(u0, 1), es1 = shfl_dn64(u1, 1);
    // products of all other k: before and after, for both slots of this lane
    const V pre0 = lane ? ep0 : vone;
    const V pre1 = ar.mul(A0, lane ? ep1 : vone);
    const V suf0 = ar.mul(lane < 63 ? es0 : vone, B0);
    const V suf1 = lane < 63 ? es1 : vone;
    const V inv_k = ar.load(pre.v[kPvInvK]);  // (prod of all k)^-1
    const V kinv0 = ar.mul(inv_k, ar.mul(pre0, suf0));
    const V kinv1 = ar.mul(inv_k, ar.mul(pre1, suf1));
    h0 = cp0 ? ar.neg(kinv0) : kinv0;
    h1 = cp1 ? ar.neg(kinv1) : kinv1;
  }

  // prefilter (setup.cl:59-87); the order of the tests does not change the
  // result.  h^e for both slots, as +1, -1 or 0 (anything else).
  auto pow_sign2 = [&](uint64_t e, int* t0, int* t1) {
    V a0, a1;
    ar.pow_lg2(h0, h1, rfl64(e), &a0, &a1);  // a wavefront-uniform exponent: scalar loop control
    *t0 = (a0 == vone) - (a0 == nmo);
    *t1 = (a1 == vone) - (a1 == nmo);
  };
  // With the residue classes, y = h^((p-1)/L) is needed anyway, and its
  // parts ya = y^(L/La), yb = y^(L/Lb), yc = y^(L/Lc), yd = y^(L/Ld) give the
  // characters h^((p-1)/d) for d | L as short powers (the same values): the
  // reference's quadratic test and its power-residue tests with d | L take
  // them from there.
  V ya0 = vone, ya1 = vone, yb0 = vone, yb1 = vone, yc0 = vone, yc1 = vone, yd0 = vone, yd1 = vone;
  if (split) {
    // y = h^((p-1)/L); the exponent is 1 when L = p - 1 (pow_lg's masked
    // shift would take h^(2^64+1) there)
    V t0 = h0, t1 = h1;
    if (exL > 1) ar.pow_lg2(h0, h1, exL, &t0, &t1);
    if (Ld > 1) {
      yd0 = ar.pow_u64(t0, La * Lb * Lc);
      yd1 = ar.pow_u64(t1, La * Lb * Lc);
      t0 = ar.pow_u64(t0, Ld);
      t1 = ar.pow_u64(t1, Ld);
    }
    // on t = y^Ld: u = t^Lc, ya = u^Lb, yb = u^La, yc = t^(La*Lb); Lc = 1 for most p
    V u0 = t0, u1 = t1;
    if (Lc > 1) {
      u0 = ar.pow_u64(t0, Lc);
      u1 = ar.pow_u64(t1, Lc);
      yc0 = ar.pow_u64(t0, La * Lb);
      yc1 = ar.pow_u64(t1, La * Lb);
    }
    ya0 = ar.pow_u64(u0, Lb);
    ya1 = ar.pow_u64(u1, Lb);
    yb0 = ar.pow_u64(u0, La);
    yb1 = ar.pow_u64(u1, La);
  }
  auto sign_of = [&](V x0, V x1, uint32_t e, int* t0, int* t1) {
    const V a0 = ar.pow_u64(x0, rfl(e)), a1 = ar.pow_u64(x1, rfl(e));
    *t0 = (a0 == vone) - (a0 == nmo);
    *t1 = (a1 == vone) - (a1 == nmo);
  };
  int cls0 = act0 ? 1 : 0, cls1 = act1 ? 1 : 0;
  uint64_t dm = dmask;
  while (dm) {
    const int i = __ffsll((long long)dm) - 1;
    dm &= dm - 1;
    int t0, t1;
    const uint32_t d = g_prd[i];
    // h^((p-1)/d) = ya^(La/d), yb^(Lb/d), yc^(Lc/d) or yd for d of one part
    if (split && La % d == 0) {
      sign_of(ya0, ya1, La / d, &t0, &t1);
    } else if (split && Lb % d == 0) {
      sign_of(yb0, yb1, Lb / d, &t0, &t1);
    } else if (split && Lc % d == 0) {
      sign_of(yc0, yc1, Lc / d, &t0, &t1);
    } else if (split && d == Ld) {
      sign_of(yd0, yd1, 1, &t0, &t1);
    } else {
      uint64_t e;
      ar.divides(pm, d, g_prdinv[i], &e);
      pow_sign2(e, &t0, &t1);
    }
    if (t0 != 1) cls0 = 0;
    if (t1 != 1) cls1 = 0;
  }
  // the quadratic character h^((p-1)/2)
  {
    int t0, t1;
    if (split) {
      sign_of(ya0, ya1, La / 2, &t0, &t1);  // y^(L/2) = (y^(Lb*Lc))^(La/2), Lb*Lc odd
    } else {
      pow_sign2(pm >> 1, &t0, &t1);
    }
    if (cls0) {
      if (resg == -1) cls0 = (t0 == 1) ? 2 : (t0 == -1) ? 3 : 0;
      else if (resg == 1 && t0 == -1) cls0 = 0;
    }
    if (cls1) {
      if (resg == -1) cls1 = (t1 == 1) ? 2 : (t1 == -1) ? 3 : 0;
      else if (resg == 1 && t1 == -1) cls1 = 0;
    }
  }

  if (__ballot(cls0 != 0 || cls1 != 0) == 0) {  // every k impossible (setup.cl:205-208)
    if (lane == 0) {
      R->nk[0] = R->nk[1] = R->nk[2] = 0;
      atomicAdd((unsigned long long*)&cnt->skipped, 1ull);
    }
    return;
  }
  const bool merged = algo != 0;

  // Constants of the BSGS kernel: b^NMIN, and the baby strides b^T, b^(2T)
  // of the reference's tables and the merged table, from k_setup_pre
  const V bn = ar.load(pre.v[kPvBn]);
  V bT0 = vone, bT1 = vone;
  if (!split) {
    bT0 = ar.load(pre.v[kPvBT0]);
    bT1 = ar.load(pre.v[kPvBT1]);
  }

  uint32_t nk[3] = {0, 0, 0};
  uint32_t cS = 0, cE = 0, cM = 0;  // residue-class shape of this prime
  V g0v, g1v, gT0v, gT1v, bTa, bTb;
  KRec* kr = J.krecs + (size_t)pi * (merged ? 2 * K : K);
  if (split) {
    // the tables of za^i (i <= La, at zv[0..]), zb^i (i <= Lb, at zv[128..])
    // and zc^i (i <= Lc, at zv[192..]), and their orders sa, sb, sc (0: no
    // i > 0 with value 1 in range, possible only for a composite p)
    for (uint32_t h = lane; h < kZH / 2; h += 64) zh[h] = 0u;
    const V za = ar.load(pre.v[kPvZa]), zb = ar.load(pre.v[kPvZb]), zc = ar.load(pre.v[kPvZc]);
    const V zd = ar.load(pre.v[kPvZd]);
    // z^i for i <= Lz at zv[vb + i], 64 per pass; the order of z in range
    auto table = [&](V zz, uint32_t Lz, uint32_t vb) -> uint32_t {
      uint32_t sz = 0;
      V cur = lane_pow(zz);
      const V z64 = ar.mul(shfl64(cur, 63), zz);
      for (uint32_t j = 0; 64 * j <= Lz; j++) {
        const uint32_t idx = 64 * j + lane;
        const bool in = idx <= Lz;
        if (in) zv[vb + idx] = ar.bits(cur);
        const uint64_t m1 = __ballot(in && idx > 0 && cur == vone);
        if (!sz && m1) sz = 64 * j + (uint32_t)__ffsll((long long)m1) - 1;
        if (64 * (j + 1) <= Lz) cur = ar.mul(cur, z64);
      }
      return sz;
    };
    // (Lc = 1 for most p: then sc = 1 and the third table is not used; Ld = 1
    // likewise for the fourth)
    const uint32_t sa = table(za, La, 0u), sb = table(zb, Lb, 128u), sc = Lc > 1 ? table(zc, Lc, 192u) : 1u;
    // the fourth part: the baby steps zd^i, i < Bd = min(Ld, 128)
    const uint32_t Bd = Ld < 128 ? Ld : 128;
    if (Ld > 1) table(zd, Bd - 1, 336u);
    wave_sync();
    // sd = Ld when zd^Ld = 1 (zd != 1 and Ld prime: the order of zd), else 0;
    // gz = zd^-Bd = zd^(Ld - Bd), the giant step of its discrete logarithm
    uint32_t sd = 1;
    V gz = vone;
    if (Ld > 1) {
      const V zB = ar.mul(ar.from_bits(zv[336 + Bd - 1]), zd);  // zd^Bd
      const uint32_t qd = Ld / Bd, rd = Ld - qd * Bd;
      if (qd > 1 || rd) gz = ar.mul(ar.pow_u64(zB, qd - 1), rd ? ar.from_bits(zv[336 + rd]) : vone);
      sd = ar.mul(gz, zB) == vone ? Ld : 0u;
    }
    // The CRT needs pairwise coprime orders.  For a prime p they divide 80,
    // 63, 143 and the prime Ld and are coprime; a composite p may give other
    // orders, which are then tested by Euclid.
    auto gcd = [](uint32_t g, uint32_t h2) {
      while (h2) {
        const uint32_t t = g % h2;
        g = h2;
        h2 = t;
      }
      return g;
    };
    const uint32_t sab = sa * sb, sabc = sab * sc;
    const bool coprime = sa && sb && sc && sd &&
                         ((80u % sa == 0 && 63u % sb == 0 && 143u % sc == 0 && (sd == 1 || sd == Ld)) ||
                          (gcd(sa, sb) == 1 && gcd(sab, sc) == 1 && gcd(sabc, sd) == 1));
    const uint32_t s = coprime ? sabc * sd : 0u;
    // 16-bit slots, two per 32-bit word of zh (so that seven workgroups'
    // tables share a CU's LDS): a slot is taken by a compare-and-swap of its
    // word, retried while its own half stays empty
    auto zh_put = [&](uint32_t hb, uint32_t h, uint32_t hm, uint32_t val) {
      for (;; h = (h + 1) & hm) {
        const uint32_t sl = hb + h, sh = (sl & 1u) * 16u;
        uint32_t old = zh[sl >> 1];
        while (((old >> sh) & 0xFFFFu) == 0u) {
          const uint32_t got = atomicCAS(&zh[sl >> 1], old, old | (val << sh));
          if (got == old) return;
          old = got;
        }
      }
    };
    if (s) {
      for (uint32_t idx = lane; idx < sa; idx += 64) zh_put(0u, (uint32_t)zv[idx] & 255u, 255u, idx + 1);
      if (lane < sb) zh_put(256u, (uint32_t)zv[128 + lane] & 127u, 127u, lane + 1);
      for (uint32_t idx = lane; Lc > 1 && idx < sc; idx += 64) zh_put(384u, (uint32_t)zv[192 + idx] & 255u, 255u, idx + 1);
      for (uint32_t idx = lane; Ld > 1 && idx < Bd; idx += 64) zh_put(640u, (uint32_t)zv[336 + idx] & 255u, 255u, idx + 1);
    }
    wave_sync();
    // sa^-1 mod sb, (sa*sb)^-1 mod sc and (sa*sb*sc)^-1 mod sd for the CRT
    const SmallMod Mb(s ? sb : 1u), Mc(s ? sc : 1u), Md(s ? sd : 1u);
    const uint32_t sainv = s ? inv_mod_lanes(sa, Mb, lane) : 0u, sabinv = s ? inv_mod_lanes(sab, Mc, lane) : 0u;
    const uint32_t sabcinv = s && sd > 1 ? inv_mod_lanes(sabc, Md, lane) : 0u;
    // index of a key in one table (hash base hb, mask hm, values at vb), or -1
    auto lookup = [&](uint64_t key, uint32_t hb, uint32_t hm, uint32_t vb) -> int {
      for (uint32_t h = (uint32_t)key & hm;; h = (h + 1) & hm) {
        const uint32_t sl = (zh[(hb + h) >> 1] >> (((hb + h) & 1u) * 16u)) & 0xFFFFu;
        if (!sl) return -1;
        if (zv[vb + sl - 1] == key) return (int)sl - 1;
      }
    };
    // the x < s with z^x = y, from ya = y^(L/La), yb = y^(L/Lb), yc =
    // y^(L/Lc) and yd = y^(L/Ld), or -1
    auto dlog = [&](V yav, V ybv, V ycv, V ydv) -> int {
      const int xa = lookup(ar.bits(yav), 0u, 255u, 0u);
      if (xa < 0) return -1;
      const int xb = lookup(ar.bits(ybv), 256u, 127u, 128u);
      if (xb < 0) return -1;
      int xc = 0, xd = 0;
      if (Lc > 1) {
        xc = lookup(ar.bits(ycv), 384u, 255u, 192u);
        if (xc < 0) return -1;
      }
      if (Ld > 1) {  // the first j with yd * zd^(-Bd*j) = zd^i, i < Bd: xd = i + Bd*j < Ld
        V g = ydv;
        xd = -1;
        for (uint32_t j = 0; j * Bd < Ld; j++) {
          const int i = lookup(ar.bits(g), 640u, 255u, 336u);
          if (i >= 0) {
            xd = i + (int)(Bd * j);
            break;
          }
          g = ar.mul(g, gz);
        }
        if (xd < 0) return -1;
      }
      const uint32_t dx = Mb((uint32_t)xb + sb - Mb((uint32_t)xa));
      const uint32_t xab = (uint32_t)xa + sa * Mb(dx * sainv);
      const uint32_t dy = Mc((uint32_t)xc + sc - Mc(xab));
      const uint32_t xabc = xab + sab * Mc(dy * sabinv);
      const uint32_t dz = Md((uint32_t)xd + sd - Md(xabc));
      return (int)(xabc + sabc * Md(dz * sabcinv));
    };
    // Table stride S: s when s is even (it fixes the parity) or no k has a
    // parity class here, else 2s (CRT of the class mod s and the parity);
    // 2 without a usable s.  Each target gets one or two residues n mod S.
    const bool anyPar = __ballot(cls0 >= 2 || cls1 >= 2) != 0;
    const uint32_t S = s == 0 ? 2u : ((s & 1) && anyPar) ? 2 * s : s;
    auto classes = [&](int cls, V yav, V ybv, V ycv, V ydv, uint32_t (&rr)[2]) -> int {
      if (!cls) return 0;
      if (s == 0) {
        if (cls == 1) {
          rr[0] = 0;
          rr[1] = 1;
          return 2;
        }
        rr[0] = (uint32_t)(cls - 2);
        return 1;
      }
      const int xl = dlog(yav, ybv, ycv, ydv);  // the class of h^-1
      if (xl < 0) return 0;
      const uint32_t ux = xl ? s - (uint32_t)xl : 0u;  // the class of h
      if (S == s) {
        if (cls >= 2 && (ux & 1u) != (uint32_t)(cls - 2)) return 0;  // s even: the parity contradicts
        rr[0] = ux;
        return 1;
      }
      if (cls == 1) {
        rr[0] = ux;
        rr[1] = ux + s;
        return 2;
      }
      rr[0] = ((ux & 1u) == (uint32_t)(cls - 2)) ? ux : ux + s;
      return 1;
    };
    uint32_t ra[2] = {0, 0}, rb[2] = {0, 0};
    const int na = classes(cls0, ya0, yb0, yc0, yd0, ra), nb = classes(cls1, ya1, yb1, yc1, yd1, rb);
    const SmallMod MS(S);
    const uint32_t nmod = ld_const(&job->nmin) % S;
    // b^d for the shifts d < S < 2^24 (the target of h^-1 is h^-1 * b^d), in
    // base-64 digits: lane j holds b^j, b^(64j), b^(4096j) and b^(262144j)
    // (j < 64; the last two levels only when S needs them), so b^d is one
    // shuffle per digit and a product between (taken with every lane active,
    // outside the divergent stores)
    const V bl = lane_pow(mbase);
    const V b64 = ar.mul(shfl64(bl, 63), mbase);  // b^64
    const V bh = lane_pow(b64);
    V b3 = vone, b4 = vone;
    if (S > 4096) {
      b3 = lane_pow(ar.mul(shfl64(bh, 63), b64));  // (b^4096)^j
      if (S > 262144) b4 = lane_pow(ar.mul(shfl64(b3, 63), ar.mul(shfl64(bh, 63), b64)));
    }
    auto bdig_pow = [&](uint32_t d) {
      V r = ar.mul(shfl64(bl, (int)(d & 63u)), shfl64(bh, (int)((d >> 6) & 63u)));
      if (S > 4096) r = ar.mul(r, shfl64(b3, (int)((d >> 12) & 63u)));
      if (S > 262144) r = ar.mul(r, shfl64(b4, (int)(d >> 18)));
      return r;
    };
    uint32_t off = 0;
#pragma unroll
    for (int a = 0; a < 2; a++) {
      const bool va = na > a, vb = nb > a;
      if (a && __ballot(va || vb) == 0) break;  // no target with a second class on this prime
      const uint32_t da = MS(ra[a] + S - nmod), db = MS(rb[a] + S - nmod);
      const V ta = ar.mul(h0, bdig_pow(da)), tb = ar.mul(h1, bdig_pow(db));
      const uint64_t m0 = __ballot(va), m1 = __ballot(vb);
      const uint32_t c0 = __popcll(m0), c1 = __popcll(m1);
      if (va) {
        KRec e;
        e.hadj = ar.mont(ta);
        e.kidx = lane;
        e.cls = da;
        kr[off + lane_rank(m0)] = e;
      }
      if (vb) {
        KRec e;
        e.hadj = ar.mont(tb);
        e.kidx = lane + 64;
        e.cls = db;
        kr[off + c0 + lane_rank(m1)] = e;
      }
      off += c0 + c1;
    }
    nk[0] = off;
    // Table length: E* = sqrt(span * targets / (2S)) balances the baby
    // steps (set, then matched) against the giant-step probes.  The BSGS
    // kernel runs giant steps in tile rows of 64 (one per lane), or of 16
    // or 32 for items with at most 48, so the giant steps are rounded to a
    // multiple of 64, or of 16 below 56 (E = span / (S*M) rounded up, then
    // the M that E needs).  E is at most the job's Q entries, and a capped
    // table takes as many giant steps as it needs.
    const uint32_t span = ld_const(&job->nmax) - ld_const(&job->nmin) + 1;
    const double estar = sqrt((double)span * (double)(off + 1) / (2.0 * (double)S));
    const double mstar = (double)span / ((double)S * estar);
    // ceil(span / d) in binary64, exact for span < 2^53: a quotient just above
    // an integer n exceeds it by at least 1/d, more than its rounding error
    auto ceil_div = [&](double d) { return (uint64_t)ceil((double)span / d); };
    const uint64_t Mrow = mstar < 56.0 ? 16 * (uint64_t)fmax(1.0, rint(mstar / 16.0)) : 64 * (uint64_t)rint(mstar / 64.0);
    uint64_t E = ceil_div((double)S * (double)Mrow);
    if (E > ld_const(&job->Q)) E = ld_const(&job->Q);
    const uint64_t M = ceil_div((double)S * (double)E);
    const uint64_t SE = (uint64_t)S * E;
    cS = S;
    cE = (uint32_t)E;
    cM = (uint32_t)M;
    // the giant step b^(S*E) and the table stride b^-S (kept in bT[1]) by
    // lanes 0 and 1, then their T-th powers by log2(T) squarings: the giant
    // step of a workgroup round b^(S*E*T) and the baby stride b^-(S*T)
    uint64_t ye2 = 0;
    V yb2 = mbase;
    if (lane == 0) ye2 = SE;
    else if (lane == 1) {
      ye2 = S;
      yb2 = binv;
    }
    V y2 = vone;
    if (lane < 2) y2 = ar.pow_u64(yb2, ye2);
    V y2T = y2;
    for (uint32_t t = 1; t < T; t <<= 1) y2T = ar.mul(y2T, y2T);
    g0v = g1v = shfl64(y2, 0);
    gT0v = gT1v = shfl64(y2T, 0);
    bTa = shfl64(y2T, 1);
    bTb = shfl64(y2, 1);
  } else {
    // b^-Q, b^-QT, b^-2QT (giant step of the table span Q) from k_setup_pre
    const V gQ_inv = ar.load(pre.v[kPvG0]);
    g0v = gQ_inv;
    g1v = algo == 2 ? gQ_inv : ar.mul(gQ_inv, gQ_inv);
    gT0v = ar.load(pre.v[kPvGT0]);
    gT1v = algo == 2 ? gT0v : ar.load(pre.v[kPvGT1]);
    bTa = merged ? bT1 : bT0;
    bTb = bT1;
    if (!merged) {
      // compaction by class: full, even, odd (the reference's sort.cl:40-71)
      const uint64_t hm0 = ar.mont(h0), hm1 = ar.mont(h1);
      uint32_t off = 0;
#pragma unroll
      for (int c = 1; c <= 3; c++) {
        const uint64_t m0 = __ballot(cls0 == c), m1 = __ballot(cls1 == c);
        const uint32_t c0 = __popcll(m0), c1 = __popcll(m1);
        if (cls0 == c) {
          KRec e;
          e.hadj = hm0;
          e.kidx = lane;
          e.cls = c;
          kr[off + lane_rank(m0)] = e;
        }
        if (cls1 == c) {
          KRec e;
          e.hadj = hm1;
          e.kidx = lane + 64;
          e.cls = c;
          kr[off + c0 + lane_rank(m1)] = e;
        }
        nk[c - 1] = c0 + c1;
        off += c0 + c1;
      }
    } else {
      // One list against a table of even exponent offsets: a target with
      // shift d searches b^(n - d) = hadj * b^-d.  Full-range k get d = 0 and
      // d = 1, even-only k d = 0, odd-only k d = 1; cls holds d.
      uint32_t off = 0;
#pragma unroll
      for (int d = 0; d <= 1; d++) {
        const bool t0 = cls0 == 1 || cls0 == 2 + d, t1 = cls1 == 1 || cls1 == 2 + d;
        const uint64_t m0 = __ballot(t0), m1 = __ballot(t1);
        const uint32_t c0 = __popcll(m0), c1 = __popcll(m1);
        if (t0) {
          KRec e;
          e.hadj = ar.mont(d ? ar.mul(h0, binv) : h0);
          e.kidx = lane;
          e.cls = d;
          kr[off + lane_rank(m0)] = e;
        }
        if (t1) {
          KRec e;
          e.hadj = ar.mont(d ? ar.mul(h1, binv) : h1);
          e.kidx = lane + 64;
          e.cls = d;
          kr[off + c0 + lane_rank(m1)] = e;
        }
        off += c0 + c1;
      }
      nk[0] = off;
      if (algo == 2) {  // a small p of a residue-class job: the parity split (S = 2)
        cS = 2;
        cE = ld_const(&job->Q);
        cM = (uint32_t)(((uint64_t)ld_const(&job->nmax) - ld_const(&job->nmin) + 1 + 2ull * cE - 1) / (2ull * cE));
        // (g0v and gT0v are b^-2Q and b^-2QT here, staged for this case)
        bTb = ar.mul(mbase, mbase);  // the table stride b^S, S = 2
      }
    }
  }
  // the stored constants in Montgomery form, converted by lanes 0..7 at once
  V cv = vone;
  switch (lane) {
    case 0: cv = mbase; break;
    case 1: cv = bn; break;
    case 2: cv = g0v; break;
    case 3: cv = g1v; break;
    case 4: cv = gT0v; break;
    case 5: cv = gT1v; break;
    case 6: cv = bTa; break;
    case 7: cv = bTb; break;
    default: break;
  }
  const uint64_t cm = ar.mont(cv);
  PrimeRec rec;
  rec.p = p;
  rec.q = q;
  rec.one = one;
  rec.mbase = shfl64(cm, 0);
  rec.bn = shfl64(cm, 1);
  rec.g[0] = shfl64(cm, 2);
  rec.g[1] = shfl64(cm, 3);
  rec.gT[0] = shfl64(cm, 4);
  rec.gT[1] = shfl64(cm, 5);
  rec.bT[0] = shfl64(cm, 6);
  rec.bT[1] = shfl64(cm, 7);
  rec.nk[0] = nk[0];
  rec.nk[1] = nk[1];
  rec.nk[2] = nk[2];
  rec.flags = algo;
  rec.S = cS;
  rec.E = cE;
  rec.M = cM;
  rec.pad32 = 0;
  rec.pad = 0;
  if (lane == 0) *R = rec;
}

// seven wavefronts per SIMD (at most 73 VGPRs; 21.5 KiB of LDS per
// workgroup): the setup is issue-bound and needs the waves to cover its FP64
// and LDS latencies
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_setup(SegJob* __restrict__ jobs) {
  const SegJob& J = jobs[blockIdx.y];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t pi = rfl(blockIdx.x * 4 + (threadIdx.x >> 6));
  __shared__ uint64_t s_zv[4][kZV];  // per wavefront: z^i of the residue classes
  __shared__ uint32_t s_zh[4][kZH / 2];
  const uint32_t ngen = J.nprimes, cap = J.cap;
  const uint32_t np = ngen < cap ? ngen : cap;
  if (pi >= np) return;
  // the staged values, by scalar loads (one address per wavefront)
  const PreRec* __restrict__ src = reinterpret_cast<const PreRec*>(J.recs + pi);
  PreRec pre;
  pre.info = ld_const(&src->info);
  if (pre.info & kPreSkip) return;  // not invertible: k_setup_pre counted it and wrote nk = 0
  pre.p = ld_const(&src->p);
  pre.q = ld_const(&src->q);
  pre.one = ld_const(&src->one);
#pragma unroll
  for (int i = 0; i < 8; i++) pre.v[i] = ld_const(&src->v[i]);
  pre.La = ld_const(&src->La);
  pre.Lb = ld_const(&src->Lb);
  pre.Lc = ld_const(&src->Lc);
  pre.Ld = ld_const(&src->Ld);
  const uint64_t p = pre.p, q = pre.q, one = pre.one;
  uint64_t* zv = s_zv[threadIdx.x >> 6];
  uint32_t* zh = s_zh[threadIdx.x >> 6];
  if ((p & 1) && p < kFpLimit) {
    FpArith ar;
    ar.p = (double)p;
    ar.pinv = 1.0 / ar.p;
    ar.R = (double)one;
    setup_prime(ar, J, pre, pi, lane, zv, zh);
  } else {
    MontArith ar;
    ar.p = p;
    ar.q = q;
    ar.one_ = one;
    uint64_t r2 = madd(madd(one, one, p), madd(one, one, p), p);
#pragma unroll
    for (int i = 0; i < 5; i++) r2 = mmul(r2, r2, p, q);  // 4^(2^5) = 2^64: Montgomery form of R
    ar.r2 = r2;
    setup_prime(ar, J, pre, pi, lane, zv, zh);
  }
}

// ---------------------------------------------------------------------------
// bsgs (three lists, the reference's giantfull/giantparity shape)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void claim(const JobDev* __restrict__ job, uint32_t* __restrict__ bitmap,
                                      FactorRec* __restrict__ fac, Counters* __restrict__ cnt, uint64_t p,
                                      uint32_t kidx, uint32_t n, uint32_t* s_dups, uint32_t* s_claims) {
  // giant.cl:21-39: atomic test-and-clear of bit n - NMIN of sequence kidx
  if (n < job->nmin || n > job->nmax) {
    atomicAdd(s_dups, 1u);
    return;
  }
  const uint32_t off = n - job->nmin;
  const uint64_t w = (uint64_t)kidx * job->words_per_seq + (off >> 5);
  const uint32_t m = 1u << (off & 31u);
  const uint32_t old = atomicAnd(&bitmap[w], ~m);
  if (old & m) {
    atomicAdd(s_claims, 1u);
    if (job->record_factors) {
      const unsigned long long f = atomicAdd((unsigned long long*)&cnt->nfactors, 1ull);
      if (f < job->factor_cap) {
        FactorRec r;
        r.p = p;
        r.n = n;
        r.kidx = kidx;
        fac[f] = r;
      }
    }
  } else {
    atomicAdd(s_dups, 1u);
  }
}

// ---------------------------------------------------------------------------
// Baby-step table of the overflow recovery (global memory, one per
// workgroup): buckets of 8 words.  Word 0 counts the values sent to the
// bucket, words 1..7 hold up to seven of them; a bucket that received more
// passes the rest to the next bucket.  A slot holds an 18-bit fingerprint of
// the Montgomery residue above a 14-bit entry index.  The bucket comes from
// the residue's low bits, the fingerprint from the others; a fingerprint
// match is confirmed by recomputing the baby value, so the table answers
// exactly "which entry holds this residue".  At most half the slots fill.  A
// residue whose low 32 bits are 0 is never inserted: the reference marks
// empty slots with 0 and cannot find it (giant.cl:44-46, 61-63).  Slot words
// go through device-scope atomic loads and stores so no stale L1 line is read
// after the workgroup's inserts.
// ---------------------------------------------------------------------------
constexpr uint32_t kBucket = 8;

__device__ __forceinline__ uint32_t tab_fp(uint64_t v) {
  const uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  const uint32_t f = ((lo >> 14) ^ (hi * 0x9E3779B1u) ^ (hi >> 13)) & 0x3FFFFu;
  return f ? f : 1u;
}

__device__ __forceinline__ uint32_t gld(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gtab_insert(uint32_t* tab, uint32_t bmask, uint64_t v, uint32_t e) {
  const uint32_t lo = (uint32_t)v;
  if (lo == 0) return;
  const uint32_t word = (tab_fp(v) << 14) | e;
  uint32_t b = lo & bmask;
  while (true) {
    uint32_t* bk = tab + b * kBucket;
    const uint32_t c = atomicAdd(bk, 1u);
    if (c < kBucket - 1) {
      __hip_atomic_store(bk + 1 + c, word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    b = (b + 1) & bmask;
  }
}
// Entry e whose baby value b^(nstart + stride*e) equals g, or -1 (rare path,
// out of line).
__device__ __noinline__ int gtab_find_at(const uint32_t* tab, uint32_t bmask, uint64_t mbase, uint64_t p, uint64_t q,
                                         uint64_t one, uint32_t base, uint32_t nstart, uint32_t stride, uint64_t g) {
  const uint32_t fp = tab_fp(g);
  uint32_t b = (uint32_t)g & bmask;
  while (true) {
    const uint32_t* bk = tab + b * kBucket;
    const uint32_t n = gld(bk);
    for (uint32_t s = 0; s < 7 && s < n; s++) {
      const uint32_t w = gld(bk + 1 + s);
      if ((w >> 14) == fp) {
        const uint32_t e = w & 0x3FFFu;
        if (basepow_sm(mbase, nstart + stride * e, base, p, q, one) == g) return (int)e;
      }
    }
    if (n < kBucket) return -1;
    b = (b + 1) & bmask;
  }
}

// Blocked Bloom filter of the baby steps: hsize 32-bit words in LDS, four
// bits per baby value, one in each byte of the word its residue picks.  The
// word is residue bits 2 .. 1 + log2(hsize), used directly as the byte offset
// v & ((hsize - 1) << 2) (one AND); the bits are the low three bits of each
// byte of the 32 residue bits from bit s on (one v_alignbit and one AND),
// each made a one-hot byte by one v_perm_b32 from the table {1, 2, ..., 128}.
// With hsize/2 values the false-positive rate is about 5e-4, with hsize
// values about 2e-3 (simulated), so a giant step costs one ds_read_b32 and
// four integer instructions; the rare positives are queued and resolved
// exactly, and the fewer there are, the less often a wavefront takes the
// queueing branch (whose returning LDS atomic waits for every LDS read in
// flight).  s = 17 for p >= 2^45; below, the fields move down so that they
// stay inside the residue's bits, but never into the word index (for small p
// they then cover fewer residue bits, which only costs selectivity).
//
// The filter starts at a fixed LDS byte address, bsgs_filter_base(T), above
// the kernel's static LDS (the host checks that the static LDS fits below it
// and sizes the dynamic LDS so that the filter ends inside the workgroup's
// allocation).  A constant base goes into the LDS instructions' offset field,
// so a word's address is the one AND; the dynamic LDS would start at the
// static size, an address the compiler does not fold (one more add per
// baby step and per probe).
using lds_u32 = __attribute__((address_space(3))) uint32_t;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
using lds_u4 = __attribute__((address_space(3))) u32x4;
using lds_u8 = __attribute__((address_space(3))) uint8_t;
struct Filter {
  uint32_t bmask;  // (hsize - 1) << 2
  uint32_t s;      // first bit of the four 3-bit fields
};
__device__ __forceinline__ Filter filter_make(uint32_t hsize, uint64_t p) {
  const uint32_t lgh = 31u - (uint32_t)__clz((int)hsize);
  const uint32_t plen = 64u - (uint32_t)__clzll((long long)p);
  Filter f;
  f.bmask = (hsize - 1u) << 2;
  f.s = max(2u + lgh, min(17u, plen > 28u ? plen - 28u : 0u));
  return f;
}
__device__ __forceinline__ uint32_t flt_bits(const Filter& f, uint64_t v) {
  const uint32_t sel = __builtin_amdgcn_alignbit((uint32_t)(v >> 32), (uint32_t)v, f.s) & 0x07070707u;
  return __builtin_amdgcn_perm(0x80402010u, 0x08040201u, sel);
}
template <uint32_t B>
__device__ __forceinline__ lds_u32* flt_word(const Filter& f, uint64_t v) {
  return (lds_u32*)((lds_u8*)(uintptr_t)B + ((uint32_t)v & f.bmask));
}
template <uint32_t B>
__device__ __forceinline__ void flt_insert(const Filter& f, uint64_t v) {
  __hip_atomic_fetch_or(flt_word<B>(f, v), flt_bits(f, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The giant-step products of the tile loop in two arithmetics: binary64 plain
// residues for odd p < 2^51 (a probe's key is kFpBias + the residue, the
// integer residue is the key without the bias), integer Montgomery residues
// otherwise.  Targets are Montgomery residues and the giant powers plain
// residues (binary64) or Montgomery residues (integer), so both give the
// Montgomery residue the table holds.
struct GiantFp {
  double pd, pinv, pc;
  using Tgt = double;
  using Pow = double;
  __device__ Tgt target(uint64_t bits) const { return __longlong_as_double((long long)bits); }
  __device__ Pow power(uint64_t bits) const { return __longlong_as_double((long long)bits); }
  __device__ Pow unit() const { return 1.0; }
  __device__ Pow mul(Pow a, Pow b) const { return fp_mulmod(a, b, pd, pinv); }
  __device__ uint64_t key(Tgt t, Pow g) const { return fp_mulmod_bits(t, g, pd, pinv, pc); }
  __device__ uint64_t value(uint64_t key) const { return key ^ kFpBias; }
};
struct GiantInt {
  uint64_t p, q, one;
  using Tgt = uint64_t;
  using Pow = uint64_t;
  __device__ Tgt target(uint64_t bits) const { return bits; }
  __device__ Pow power(uint64_t bits) const { return bits; }
  __device__ Pow unit() const { return one; }
  __device__ Pow mul(Pow a, Pow b) const { return mmul(a, b, p, q); }
  __device__ uint64_t key(Tgt t, Pow g) const { return mmul(t, g, p, q); }
  __device__ uint64_t value(uint64_t key) const { return key; }
};

// Hash of the match pass's filter over the queued positives.  It must not
// reuse the main filter's fields: a false positive passed the main filter
// because some baby value shares its word and bit fields, and that baby value
// would then pass here too.  A multiplicative mix of all 64 bits, read from
// its upper bits (which depend on every input bit): word index bits 23..31,
// bit positions from bits 7..21.
__device__ __forceinline__ uint32_t qmix(uint64_t v) {
  return ((uint32_t)v ^ ((uint32_t)(v >> 32) * 0x85EBCA77u)) * 0x9E3779B1u;
}
__device__ __forceinline__ uint32_t qmix_bits(uint32_t h) {
  return (1u << ((h >> 7) & 31u)) | (1u << ((h >> 12) & 31u)) | (1u << ((h >> 17) & 31u));
}

// Trapezoid self-scheduling of the persistent kernel's work: ticket k (one
// atomic increment of a shared counter) is the item range [start(k),
// start(k+1)), chunk sizes falling linearly from f = N/(2B) items (B
// workgroups) to 1 over C = ceil(2N/(f+1)) tickets, so
//   start(k) = floor(k f - d k (k-1) / 2),  d = (f-1)/(C-1),
// clipped to N.  About 4B atomics per launch instead of one per item (one
// device-wide counter serialises at tens of nanoseconds per atomic), and small
// chunks at the end balance the finish.  Every block computes the same
// boundaries from k; consecutive boundaries differ by at least 1 - 1e-9, so
// the chunks tile [0, N) exactly.
struct Guide {
  double f, d;
  uint32_t N, C;
};
__device__ __forceinline__ Guide guide_make(uint32_t N, uint32_t nblocks) {
  Guide g;
  g.N = N;
  g.f = fmax(1.0, floor((double)N / (2.0 * (double)nblocks)));
  g.C = (uint32_t)ceil(2.0 * (double)N / (g.f + 1.0));
  g.d = g.C > 1 ? (g.f - 1.0) / (double)(g.C - 1) : 0.0;
  return g;
}
__device__ __forceinline__ uint32_t guide_start(const Guide& g, uint32_t k) {
  if (k >= g.C) return g.N;
  const double kd = (double)k;
  const uint32_t s = (uint32_t)floor(kd * g.f - g.d * kd * (kd - 1.0) * 0.5);
  return s < g.N ? s : g.N;
}

// Filter positives are queued in LDS during the giant steps and resolved by
// all threads together at round boundaries (once the queue is half full, and
// at the end of the item), instead of one lane at a time inside the loop
// (bsgs_queue_cap entries: 128 at 128 and 256 threads, 384 above).
constexpr uint32_t kHitCap = 64;

// One persistent kernel serves every job of a batch.  An item is (job, prime,
// class) for a job with the reference's three tables (giant.cl giantfull /
// giantparity) and (job, prime) for a job with the merged table of
// b^(NMIN + 2j), whose targets carry their shift d.  The LDS holds the largest
// table of the batch; each item uses its own job's table size.
// Register budgets for the occupancy the LDS allows at the production table
// (8192 filter words, 40 KiB per 256-thread workgroup): four 256-thread or
// two 512-thread workgroups or one 1024-thread workgroup per CU, four waves
// per SIMD, at most 128 VGPRs.
template <int T>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(T >= 256 ? 4 : 3))) void k_bsgs(
    SegJob* __restrict__ jobs, BatchState* __restrict__ bs) {
  // Four wavefronts per SIMD from 256 threads on (128 VGPRs, a few spilled):
  // at 256 threads four 40 KiB workgroups then share a CU's LDS with filters
  // of 8192 words, which hides more of the items' latencies than three
  // workgroups at 139 VGPRs without spills (44.5 against 47.2 ms per launch
  // at the production shape).  Targets per giant-step tile: four (43.3 ms;
  // five the same, three 44.0, six 44.5), six at 128 threads.
  constexpr uint32_t kU = T <= 128 ? 6 : 4;
  constexpr uint32_t kQCap = bsgs_queue_cap(T);
  constexpr uint32_t kHQ = bsgs_hash_words(T);
  constexpr uint32_t kHQShift = kHQ == 1024 ? 22 : 23;  // qmix's word index bits
  constexpr uint32_t kFB = bsgs_filter_base(T);
  // the discrete-log hash holds j + 1 <= ceil(16383 / 64) + 1 in 9 bits and at
  // most 257 entries in its kHQ slots (fewer than kHQ: some bucket keeps an
  // empty last slot, which ends every search)
  static_assert(16383 / 64 + 2 < 512 && 257 < kHQ && (1u << (32 - kHQShift)) == kHQ, "discrete-log hash");
  constexpr uint32_t kW = T / 64;  // wavefronts per workgroup
  constexpr int kLgT = T >= 1024 ? 10 : T >= 512 ? 9 : T >= 256 ? 8 : 7;  // log2(T)
  static_assert((1 << kLgT) == T, "T is a power of two from 128 to 1024");
  __shared__ uint64_t s_pw[2][kLgT];  // per-item power tables (see vstart)
  __shared__ uint64_t s_h[2 * RS_MAXK + 24];  // targets of the item (+ the last tile's pad, rg*kU <= 24)
  static_assert(RS_MAXK <= 256, "a sequence index fits the low 8 bits of s_k");
  __shared__ uint32_t s_k[2 * RS_MAXK];   // sequence index | shift (< 2^21) << 8
  __shared__ uint32_t s_item, s_order, s_qn, s_wkdone, s_qovf;
  // duplicates and claims of the current item, by item parity (thread 0 sums
  // one item's while the next item's claims count into the other pair)
  __shared__ uint32_t s_dups[2], s_claims[2];
  __shared__ uint32_t s_next;  // the block's next item when it is known (same chunk and job), else ~0
  // The fields of the block's current job that every item reads, copied by
  // thread 0 when the job changes (its items are read through these, not
  // through the job's descriptors, which would cost a chain of dependent
  // global loads per item).
  struct JobCache {
    const JobDev* job;
    const PrimeRec* recs;
    const KRec* krecs;
    uint32_t* bitmap;
    FactorRec* fac;
    Counters* cnt;
    uint32_t item_base, algo, hsize, base, kcount, nmin, nmax, Q, QQ, M, MM, qcap;
  };
  __shared__ JobCache s_jc;
  __shared__ uint4 s_q[kQCap];  // filter positives of the item: residue, target, giant step
  // Bloom filter of the queued residues (order path) or hash of the W_j (discrete logs)
  __shared__ __align__(16) uint32_t s_hq[kHQ];
  // G^f, f < 64 (the giant steps' first tiles), or, on the order path, the
  // baby-step hits of a match pass: (entry, queue slot).  One array for both
  // keeps the static LDS within 16 KiB, so two workgroups with 64 KiB filters
  // share a CU.
  static_assert(kHitCap * sizeof(uint2) == 64 * sizeof(uint64_t), "s_gl and s_hit share one array");
  __shared__ uint64_t s_gl[64];
  uint2* const s_hit = (uint2*)s_gl;
  __shared__ uint64_t s_pf[64];     // b^(stride*f), f < 64 (the positives' discrete logs)
  __shared__ uint32_t s_nhit;
  const uint32_t tid = threadIdx.x;
  const uint32_t njobs = bs->njobs;
  const uint32_t nitems = bs->item_base[njobs];

  // Thread 0 takes chunks of items by guided self-scheduling (one atomic per
  // chunk, see Guide) with the next ticket in flight while a chunk runs.  A
  // block's items only grow, so its job index only moves forward.
  const Guide gd = guide_make(nitems, gridDim.x);
  uint32_t tk = 0, ci = 0, ce = 0;       // ticket in flight, current chunk [ci, ce)
  uint32_t jcur = 0, jend = 0;           // jend: first item past job jcur
  uint32_t jcached = 0xFFFFFFFFu;        // the job in s_jc
  // Thread 0 sums the counters of its items in LDS and adds them to the
  // job's once (when the block moves to the next job and at exit): every
  // block adding per item would serialise on the few counter addresses.
  __shared__ unsigned long long s_acc[8];  // dups, claims, work, probes, lists[3], order
  __shared__ Counters* s_acc_cnt;
  auto acc_flush = [&]() {
    Counters* c = s_acc_cnt;
    if (!c) return;
    unsigned long long* dst[8] = {(unsigned long long*)&c->dups,       (unsigned long long*)&c->claims,
                                  (unsigned long long*)&c->work_mmul,  (unsigned long long*)&c->work_probe,
                                  (unsigned long long*)&c->lists[0],   (unsigned long long*)&c->lists[1],
                                  (unsigned long long*)&c->lists[2],   (unsigned long long*)&c->order_lists};
#pragma unroll
    for (int i = 0; i < 8; i++) {
      if (s_acc[i]) atomicAdd(dst[i], s_acc[i]);
      s_acc[i] = 0;
    }
  };
  if (tid == 0) {
    s_acc_cnt = nullptr;
#pragma unroll
    for (int i = 0; i < 8; i++) s_acc[i] = 0;
  }
  if (tid == 0) {
    tk = atomicAdd(&bs->work_next, 1u);
    jend = bs->item_base[1];
  }
  // Thread 0 takes the next item: s_item, s_next and, on a job change, the
  // job cache.  It runs once before the loop and then during each item, after
  // the walk's barrier (every wavefront has read the current item's values by
  // then), so the item's last barrier also publishes the next item and no
  // barrier of its own is needed.
  auto fetch = [&]() {
    while (ci >= ce && ci < nitems) {  // chunk used up: take the ticket, fetch the next
      ci = guide_start(gd, tk);
      ce = guide_start(gd, tk + 1);
      tk = atomicAdd(&bs->work_next, 1u);
    }
    const uint32_t it = ci < ce ? ci++ : nitems;
    if (it < nitems) {
      while (it >= jend) jend = bs->item_base[++jcur + 1];
      if (jcur != jcached) {
        const SegJob& J = jobs[jcur];
        const JobDev* jd = J.job;
        s_jc.job = jd;
        s_jc.recs = J.recs;
        s_jc.krecs = J.krecs;
        s_jc.bitmap = J.bitmap;
        s_jc.fac = J.fac;
        s_jc.cnt = J.cnt;
        s_jc.item_base = bs->item_base[jcur];
        s_jc.algo = jd->algo;
        s_jc.hsize = jd->hsize;
        s_jc.base = jd->base;
        s_jc.kcount = (uint32_t)jd->kcount;
        s_jc.nmin = jd->nmin;
        s_jc.nmax = jd->nmax;
        s_jc.Q = jd->Q;
        s_jc.QQ = jd->QQ;
        s_jc.M = jd->M;
        s_jc.MM = jd->MM;
        s_jc.qcap = jd->qcap;
        jcached = jcur;
      }
    }
    s_item = it;
    s_next = (ci < ce && ci < jend) ? ci : 0xFFFFFFFFu;
  };
  if (tid == 0) {
    fetch();
    s_dups[0] = s_claims[0] = 0;
    s_qn = s_wkdone = s_qovf = 0;
  }
  __syncthreads();
  for (uint32_t par = 0;; par ^= 1u) {
    // workgroup-uniform values read from LDS go through readfirstlane, so
    // the compiler keeps the item's values (and the loads that follow from
    // them) in scalar registers
    // (thread 0 writes the next item's only after the walk's barrier; an
    // item without targets takes it between two barriers of its own)
    const uint32_t item = rfl(s_item);
    if (item >= nitems) break;
    const JobDev* __restrict__ job = rflp(s_jc.job);
    const uint32_t jalgo = rfl(s_jc.algo);
    const bool merged = jalgo != 0;  // one list per prime (merged or residue-class table)
    const uint32_t local = item - rfl(s_jc.item_base);
    const uint32_t pi = merged ? local : local / 3, c = merged ? 0 : local % 3;
    const PrimeRec* __restrict__ R = rflp(s_jc.recs) + pi;
    // the record's fields in one batch of scalar loads, before the test of nk
    const uint32_t nk = ld_const(&R->nk[c]), nk0 = ld_const(&R->nk[0]), nk1 = ld_const(&R->nk[1]);
    const uint64_t p = ld_const(&R->p), q = ld_const(&R->q), one = ld_const(&R->one), mbase = ld_const(&R->mbase);
    const uint64_t rbn = ld_const(&R->bn), rbT0 = ld_const(&R->bT[0]), rbT1 = ld_const(&R->bT[1]);
    const uint64_t rg0 = ld_const(&R->g[0]), rg1 = ld_const(&R->g[1]);
    const uint32_t rS = ld_const(&R->S), rE = ld_const(&R->E), rM = ld_const(&R->M);
    // (an empty asm that reads them all keeps the loads together, ahead of the branch)
    asm volatile("" ::"s"(nk), "s"(nk0), "s"(nk1), "s"(p), "s"(q), "s"(one), "s"(mbase), "s"(rbn), "s"(rbT0),
                 "s"(rbT1), "s"(rg0), "s"(rg1), "s"(rS), "s"(rE), "s"(rM));
    if (nk == 0) {
      __syncthreads();
      if (tid == 0) {
        fetch();
        s_dups[par ^ 1u] = s_claims[par ^ 1u] = 0;
      }
      __syncthreads();
      continue;
    }
    const uint32_t hsize = rfl(s_jc.hsize);
    const uint32_t base = rfl(s_jc.base);
    const int K = (int)rfl(s_jc.kcount);
    const uint32_t nmin = rfl(s_jc.nmin), nmax = rfl(s_jc.nmax);
    const uint32_t cp = (c && !merged) ? 1 : 0;
    const Filter flt = filter_make(hsize, p);
    // table stride, baby-step entries, exponent span and giant steps of the
    // item: the reference's (three lists), fixed (merged), or per prime
    // (residue classes)
    const bool rcls = jalgo == 2;
    const uint32_t jQ = rfl(s_jc.Q), jQQ = rfl(s_jc.QQ), jM = rfl(s_jc.M), jMM = rfl(s_jc.MM);
    const uint32_t stride = rcls ? rS : (cp || merged) ? 2 : 1;
    const uint32_t nent = rcls ? rE : (cp ? jQQ : jQ) / stride;
    const uint32_t Ql = rcls ? stride * nent : cp ? jQQ : jQ;
    const uint32_t Ml = rcls ? rM : cp ? jMM : jM;
    const uint32_t nstart = nmin + (c == 2 ? 1 : 0);
    const uint64_t v0 = (c == 2) ? mul_base(rbn, mbase, base, p, q) : rbn;
    const uint64_t bstep = cp ? rbT1 : rbT0;
    const uint32_t koff = (c >= 1 ? nk0 : 0) + (c >= 2 ? nk1 : 0);
    const KRec* __restrict__ jkrecs = rflp(s_jc.krecs);
    const uint32_t kstride = merged ? 2 * K : K;  // target records per prime
    const KRec* __restrict__ kr = jkrecs + (size_t)pi * kstride + koff;
    uint32_t* __restrict__ bitmap = rflp(s_jc.bitmap);
    FactorRec* __restrict__ fac = rflp(s_jc.fac);
    Counters* __restrict__ cnt = rflp(s_jc.cnt);
    const uint32_t jq = rfl(s_jc.qcap), qcap = jq ? min(jq, kQCap) : kQCap;  // (the host never passes more)
    // Queue a filter positive.  Past qcap the entry is dropped and the
    // item's remaining rounds are redone exactly (settle); the hot loop
    // itself holds no table lookups, calls or global accesses.
    auto defer = [&](uint64_t g, uint32_t i, uint32_t qi) {
      const uint32_t slot = atomicAdd(&s_qn, 1u);
      if (slot < qcap) {
        // the queue's top slots hold the staged W_j until every wavefront has
        // hashed its share: a positive that would land there earlier makes
        // the settle redo the steps exactly
        if (slot < kQCap - 65 || __hip_atomic_load(&s_wkdone, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == kW)
          s_q[slot] = make_uint4((uint32_t)g, (uint32_t)(g >> 32), i, qi);
        else
          s_qovf = 1u;
      }
    };
    // Giant steps in tiles: a tile is a row of rw consecutive giant steps
    // (rw = 64, or 16 or 32 for an item with few giant steps) times rg =
    // 64 / rw groups of kU consecutive targets: lane l takes giant step
    // q = rw*qb + (l mod rw) and the targets of group g*rg + l / rw.  Tile
    // t = qb*ng + g, and wavefront w of the workgroup takes the tiles w,
    // w + kW, w + 2 kW, ... (step j holds the tiles j*kW .. j*kW + kW - 1).
    // Every probe of the item, whatever M and the target count, then runs
    // through the one loop below, and the wavefronts get equal shares (lanes
    // with q >= M and targets past nk are computed and ignored).
    // (the wavefront index through readfirstlane: the tile walk is scalar)
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(tid >> 6)), lane = tid & 63;
    // row width: the fewest idle lanes, rows of 64 from 49 giant steps on
    const uint32_t lw = Ml <= 16 ? 4u : Ml <= 32 ? 5u : Ml <= 48 ? 4u : 6u;
    const uint32_t rw = 1u << lw, rg = 64u >> lw, lsub = lane & (rw - 1), lgrp = lane >> lw;
    const uint32_t ng = (nk + rg * kU - 1) / (rg * kU), nblk = (Ml + rw - 1) / rw, ntiles = nblk * ng;
    const uint32_t nsteps = (ntiles + kW - 1) / kW;
    // Steps between queue checks: a step queues about T * kU * 5e-4
    // positives (the filter's rate at half load); checking when a quarter of
    // the queue could be filled at 2e-3 (one value per word) keeps overflows
    // rare without a barrier per step.
    const uint32_t every = max(1u, 125u * qcap / (T * kU));

    // Baby steps: entry e = tid + T*i is b^(nstart + stride*e).  Below 2^51
    // the steps multiply in binary64 by the plain residue of the stride,
    // which keeps the Montgomery representation of the integer path.
    const bool fp = (p & 1) && p < kFpLimit;
    const double pd = (double)p, pinv = fp ? 1.0 / pd : 0.0;
    const double bd = fp ? (double)mmul(mbase, 1, p, q) : 0.0;  // plain b
    // Per-item power tables, one entry per bit of tid: s_pw[0][i] =
    // b^(stride*2^i) and s_pw[1][i] = G^(2^i) (plain residues as double bits
    // on the binary64 path, Montgomery residues otherwise), written by
    // 2*log2(T) threads before the barrier below.  A thread's first baby
    // value b^(nstart + stride*tid) and giant power G^tid then cost
    // popcount(tid) products instead of an exponentiation each.
    const uint64_t Gm = cp ? rg1 : rg0;
    if (tid < 2 * kLgT) {
      const uint32_t i = tid % kLgT, side = tid / kLgT;
      if (fp) {
        double x = side ? (double)mmul(Gm, 1, p, q)
                        : rcls ? (double)mmul(rbT1, 1, p, q) : fp_powmod_sm(bd, stride, pd, pinv);
        for (uint32_t k = 0; k < i; k++) x = fp_mulmod(x, x, pd, pinv);
        s_pw[side][i] = (uint64_t)__double_as_longlong(x);
      } else {
        uint64_t x = side ? Gm : rcls ? rbT1 : basepow_sm(mbase, stride, base, p, q, one);
        for (uint32_t k = 0; k < i; k++) x = mmul(x, x, p, q);
        s_pw[side][i] = x;
      }
    }
    uint64_t vstart = 0;  // b^(nstart + stride*tid), Montgomery; set after the barrier
    const double bsd = fp ? (double)mmul(bstep, 1, p, q) : 0.0;
    // Every baby value of this thread, e = tid + T*i < emax: fn(e, value),
    // in four interleaved chains (i mod 4) stepping by b^(4*stride*T), so
    // four independent products are in flight instead of one dependent chain.
    auto baby_walk = [&](uint32_t emax, auto&& fn) {
      if (fp) {
        double v[4];
        v[0] = (double)vstart;
#pragma unroll
        for (int c = 1; c < 4; c++) v[c] = fp_mulmod(v[c - 1], bsd, pd, pinv);
        const double b2 = fp_mulmod(bsd, bsd, pd, pinv), b4 = fp_mulmod(b2, b2, pd, pinv);
        uint32_t e = tid;
        for (; e + 3 * T < emax; e += 4 * T) {
#pragma unroll
          for (int c = 0; c < 4; c++) fn(e + c * T, fp_to_u64(v[c]));
#pragma unroll
          for (int c = 0; c < 4; c++) v[c] = fp_mulmod(v[c], b4, pd, pinv);
        }
#pragma unroll
        for (int c = 0; c < 3; c++)
          if (e + c * T < emax) fn(e + c * T, fp_to_u64(v[c]));
      } else {
        uint64_t v = vstart;
        for (uint32_t e = tid; e < emax; e += T) {
          fn(e, v);
          v = mmul(v, bstep, p, q);
        }
      }
    };

    // Exact resolution of the queued values s_q[0, nq): a small Bloom filter
    // of them in LDS (s_hq, kHQ words hashed by qmix), then
    // every baby value e < emax is tested against it with one LDS read; the
    // rare value that passes is compared with every queued value.  A hit is
    // a solution n = nstart + (giant step) * Ql + stride * e + d, or, with
    // ord != 0 (the order path), n = nstart + stride * e + d + t * ord.  Hits
    // are listed in LDS during the walk and claimed afterwards, so the walk
    // stays small; a full list sends the whole walk again with claims on the
    // spot (each hit still claimed once).  Baby values whose low 32 bits are
    // 0 never match: the reference's table cannot hold them (giant.cl:44-46,
    // 61-63).
    auto claim_hit = [&](uint32_t e, uint32_t slot, uint32_t ord) {
      const uint4 qe = s_q[slot];
      const uint32_t kk = s_k[qe.z], kidx = kk & 0xFFu, d = kk >> 8;
      if (ord) {
        for (uint64_t n = (uint64_t)nstart + stride * e + d; n <= nmax; n += ord)
          claim(job, bitmap, fac, cnt, p, kidx, (uint32_t)n, &s_dups[par], &s_claims[par]);
      } else {
        claim(job, bitmap, fac, cnt, p, kidx, nstart + qe.w * Ql + stride * e + d, &s_dups[par], &s_claims[par]);
      }
    };
    auto match = [&](uint32_t nq, uint32_t emax, uint32_t ord) __attribute__((always_inline)) {
      for (uint32_t s = tid; s < kHQ; s += T) s_hq[s] = 0u;
      if (tid == 0) s_nhit = 0;
      __syncthreads();
      for (uint32_t s = tid; s < nq; s += T) {
        const uint4 qe = s_q[s];
        const uint32_t h = qmix(((uint64_t)qe.y << 32) | qe.x);
        atomicOr(&s_hq[h >> kHQShift], qmix_bits(h));
      }
      __syncthreads();
      // queued values equal to baby value v: slots s with fn(e, s)
      auto scan = [&](uint32_t e, uint64_t v, auto&& fn) {
        const uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
        const uint32_t h = qmix(v), m = qmix_bits(h);
        if (!lo || (s_hq[h >> kHQShift] & m) != m) return;
        for (uint32_t s = 0; s < nq; s++) {
          const uint4 qe = s_q[s];
          if (qe.x == lo && qe.y == hi) fn(e, s);
        }
      };
      baby_walk(emax, [&](uint32_t e, uint64_t v) {
        scan(e, v, [&](uint32_t ee, uint32_t s) {
          const uint32_t at = atomicAdd(&s_nhit, 1u);
          if (at < kHitCap) s_hit[at] = make_uint2(ee, s);
        });
      });
      __syncthreads();
      const uint32_t nh = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_nhit);
      if (nh <= kHitCap) {
        for (uint32_t s = tid; s < nh; s += T) claim_hit(s_hit[s].x, s_hit[s].y, ord);
      } else {
        uint64_t v = vstart;  // rare: walk again, claiming on the spot
        for (uint32_t e = tid; e < emax; e += T) {
          scan(e, v, [&](uint32_t ee, uint32_t s) { claim_hit(ee, s, ord); });
          v = mmul(v, bstep, p, q);
        }
      }
    };

    // Exact resolution of the queued positives of a table without the order
    // shortcut, one discrete logarithm each.  With f = lane (0..63), a queued
    // value g is the baby value of entry e = 64j - f exactly when
    // g * b^(stride*f) = W_j, the baby value of entry 64j, for the one j in
    // [0, J], J = ceil((E - 1) / 64).  The hash s_hq holds a 23-bit
    // fingerprint of each W_j and j + 1 in buckets of four slots (kHB buckets
    // picked by residue bits 2..8, linear probing over buckets; a bucket
    // fills its slots in order, so a search ends at the first bucket with an
    // empty last slot).  Lane 0 of the baby walk writes each W_j's bucket
    // and fingerprint (32 bits; 64j < E) as it passes entry 64j, and W_J's
    // from W_(J-1) * b^(64*stride) when 64J >= E, to s_wk, the top 258 words
    // of the positive queue's LDS (idle until the giant steps); after the
    // walk's barrier each wavefront inserts its share (whash_build, at most
    // two per lane, one returning LDS atomic each) and counts itself done in
    // s_wkdone, with no barrier: the giant steps queue positives from slot 0,
    // and one that would land in the top slots before all wavefronts are done
    // (more than 62 positives by then) makes the settle redo the steps
    // exactly.  A fingerprint
    // hit, rare unless it is a solution, is compared in full with W_j = v0 *
    // (b^(64*stride))^j, computed then, so a hit is exact and the values
    // need no LDS.  Every wavefront takes every kW-th queued
    // value, two at a time: per lane and value one product (by b^(stride*f)
    // from s_pf) and, nearly always, one 16-byte LDS read.  A value whose low
    // 32 bits are 0 is never a hit (the reference's table cannot hold it,
    // giant.cl:44-46, 61-63).
    constexpr uint32_t kHB = kHQ / 4;
    const uint32_t nJ = nent > 1 ? (nent - 1 + 63) / 64 : 0;  // J
    auto wfp = [](uint64_t w) { return (uint32_t)(w >> 20) & 0x7FFFFFu; };
    constexpr uint32_t kHBBits = kHB == 256 ? 8 : 7;
    static_assert((1u << kHBBits) == kHB && kQCap * 4 >= 258, "s_wk: the bucket and 258 entries in the queue's LDS");
    uint32_t* const s_wk = (uint32_t*)(s_q + kQCap) - 258;
    auto winfo = [&](uint64_t w) { return (wfp(w) << kHBBits) | (((uint32_t)w >> 2) & (kHB - 1)); };
    // W_j from the walk (lane 0, entry 64j), and W_J after W_(J-1) when 64J >= E
    auto w_put = [&](uint32_t j, uint64_t w) {
      s_wk[j] = winfo(w);
      if (j + 1 == nJ && 64 * nJ >= nent)
        s_wk[nJ] = winfo(fp ? fp_mulmod_bits(__longlong_as_double((long long)w) - kFpTwo52,
                                             __longlong_as_double(s_pw[0][6]), pd, pinv, pd + kFpTwo52)
                            : mmul(w, s_pw[0][6], p, q));
    };
    auto whash_build = [&]() {
      for (uint32_t j = wave + kW * lane; j <= nJ; j += T) {
        const uint32_t wi = s_wk[j], entry = ((wi >> kHBBits) << 9) | (j + 1);
        for (uint32_t bk = wi & (kHB - 1);; bk = (bk + 1) & (kHB - 1)) {
          uint32_t k = 0;
          while (k < 4 && atomicCAS(&s_hq[4 * bk + k], 0u, entry) != 0u) k++;
          if (k < 4) break;
        }
      }
      if (lane == 0) atomicAdd(&s_wkdone, 1u);  // this wavefront's reads of s_wk are done
    };
    // W_j in the form the walk gives it (binary64 keys kFpBias + residue, or
    // Montgomery residues)
    auto w_value = [&](uint32_t j) -> uint64_t {
      if (fp)
        return fp_mulmod_bits((double)v0, fp_powmod_sm(__longlong_as_double(s_pw[0][6]), j, pd, pinv), pd, pinv,
                              pd + kFpTwo52);
      return mmul(v0, powmod_sm(s_pw[0][6], j, p, q, one), p, q);
    };
    auto dlog = [&](uint32_t nq) {
      const uint64_t plane = s_pf[lane];
      // the hits among one bucket's slots; true when the bucket is full (the search goes on)
      auto scan = [&](const uint4& en, uint32_t fg, uint64_t g, uint64_t gam, uint32_t sq) {
        const uint32_t sl[4] = {en.x, en.y, en.z, en.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
          if (sl[k] && (sl[k] >> 9) == fg) {
            const uint32_t j = (sl[k] & 511u) - 1;
            const uint32_t e = 64 * j - lane;
            if (64 * j >= lane && e < nent && (uint32_t)g && w_value(j) == gam) claim_hit(e, sq, 0u);
          }
        }
        return en.w != 0u;
      };
      auto look = [&](uint32_t sq, uint64_t g, uint64_t gam, uint4 en, uint32_t bk) {
        const uint32_t fg = wfp(gam);
        while (scan(en, fg, g, gam, sq)) {
          bk = (bk + 1) & (kHB - 1);
          en = ((const uint4*)s_hq)[bk];
        }
      };
      auto lpow = [&](uint64_t g) {  // in the form of W_j
        return fp ? fp_mulmod_bits((double)g, __longlong_as_double(plane), pd, pinv, pd + kFpTwo52)
                  : mmul(g, plane, p, q);
      };
      for (uint32_t sq = wave; sq < nq; sq += 2 * kW) {
        const bool two = sq + kW < nq;  // wavefront-uniform
        const uint4 qa = s_q[sq], qb = s_q[two ? sq + kW : sq];
        const uint64_t ga = ((uint64_t)qa.y << 32) | qa.x, gb = ((uint64_t)qb.y << 32) | qb.x;
        const uint64_t gama = lpow(ga), gamb = lpow(gb);
        const uint32_t ba = ((uint32_t)gama >> 2) & (kHB - 1), bb = ((uint32_t)gamb >> 2) & (kHB - 1);
        const uint4 ea = ((const uint4*)s_hq)[ba], eb = ((const uint4*)s_hq)[bb];
        look(sq, ga, gama, ea, ba);
        if (two) look(sq + kW, gb, gamb, eb, bb);
      }
    };

    // The overflow recovery: the bucketed table of the baby values is built
    // in this workgroup's global scratch (at most half full) and every probe
    // of the tiles of steps ja on is looked up directly with integer
    // Montgomery products h * G^q.
    auto redo = [&](uint32_t ja) {
      const uint32_t tw = recovery_table_words(nent), gbmask = tw / kBucket - 1;
      uint32_t* gtab = bs->scratch + (size_t)blockIdx.x * bs->scratch_words;
      __syncthreads();
      for (uint32_t s = tid; s < tw; s += T) __hip_atomic_store(gtab + s, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __threadfence();
      __syncthreads();
      {  // one chain per thread: this path is rare
        uint64_t v = vstart;
        for (uint32_t e = tid; e < nent; e += T) {
          gtab_insert(gtab, gbmask, v, e);
          v = mmul(v, bstep, p, q);
        }
      }
      __threadfence();
      __syncthreads();
      auto resolve = [&](uint64_t g, uint32_t i, uint32_t qi) {
        const int e = gtab_find_at(gtab, gbmask, mbase, p, q, one, base, nstart, stride, g);
        if (e >= 0) {
          const uint32_t kk = s_k[i];
          claim(job, bitmap, fac, cnt, p, kk & 0xFFu, nstart + qi * Ql + stride * (uint32_t)e + (kk >> 8),
                &s_dups[par], &s_claims[par]);
        }
      };
      const uint64_t G = Gm;
      for (uint32_t t = ja * kW + wave; t < ntiles; t += kW) {
        const uint32_t qi = (t / ng) * rw + lsub, g0 = ((t % ng) * rg + lgrp) * kU;
        if (qi >= Ml) continue;
        const uint64_t Gq = powmod_sm(G, qi, p, q, one);
        for (uint32_t i = g0; i < g0 + kU && i < nk; i++) resolve(mmul(kr[i].hadj, Gq, p, q), i, qi);
      }
    };
    // After the steps before r1 (block-uniform): resolve the queue when it
    // is half full or at the end.  If it overflowed, the positives of steps
    // [r0, r1) are incomplete: redo every step from r0 on directly and
    // report the item finished (true).
    auto settle = [&](uint32_t& r0, uint32_t r1, bool final) -> bool {
      __syncthreads();
      const uint32_t nq = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_qn);
      if (nq > qcap || rfl(s_qovf)) {
        redo(r0);
        return true;
      }
      if (nq && (final || 2 * nq >= qcap)) {  // no positive, nothing to resolve
        dlog(nq);
        if (!final) {  // after the final one the item's last barrier follows, and the next item's fetch clears s_qn
          __syncthreads();
          if (tid == 0) s_qn = 0;
          __syncthreads();
        }
        r0 = r1;
      }
      return false;
    };

    {
      const uint4 z = make_uint4(0u, 0u, 0u, 0u);
      for (uint32_t s = tid; s < hsize / 4; s += T) ((lds_u4*)(uintptr_t)kFB)[s] = u32x4{0u, 0u, 0u, 0u};  // the filter
      for (uint32_t s = tid; s < kHQ / 4; s += T) ((uint4*)s_hq)[s] = z;  // the discrete-log hash
    }
    // targets: Montgomery residues, as binary64 values on the binary64 path
    // (the giant steps' operands; exact below 2^51)
    for (uint32_t i = tid; i < nk; i += T) {
      const uint64_t h = ld_global(&kr[i].hadj);
      const uint64_t kc = ld_global((const uint64_t*)&kr[i].kidx);  // kidx, cls
      s_h[i] = fp ? (uint64_t)__double_as_longlong((double)h) : h;
      s_k[i] = (uint32_t)kc | ((merged ? (uint32_t)(kc >> 32) : 0u) << 8);
    }
    if (tid == 0) s_order = 0xFFFFFFFFu;
    __syncthreads();
    // vstart = v0 * b^(stride*lane) * b^(stride*64*wave): the lane's factor
    // as a product tree over its six bits (wavefront 0 keeps it in s_pf for
    // the discrete logs), the wavefront's bits by uniform branches; the last
    // wavefront (not wavefront 0, which computes the power tables and runs
    // the item fetch) also writes G^lane to s_gl (the giant steps' first tiles)
    if (fp) {
      double f[6];
#pragma unroll
      for (int i = 0; i < 6; i++) f[i] = ((lane >> i) & 1u) ? __longlong_as_double(s_pw[0][i]) : 1.0;
      const double f01 = fp_mulmod_signed(f[0], f[1], pd, pinv), f23 = fp_mulmod_signed(f[2], f[3], pd, pinv),
                   f45 = fp_mulmod_signed(f[4], f[5], pd, pinv);
      const double fl = fp_mulmod_signed(fp_mulmod_signed(f01, f23, pd, pinv), f45, pd, pinv);
      if (tid < 64) s_pf[tid] = (uint64_t)__double_as_longlong(fl);
      if (wave == kW - 1) {
        double g[6];
#pragma unroll
        for (int i = 0; i < 6; i++) g[i] = ((lane >> i) & 1u) ? __longlong_as_double(s_pw[1][i]) : 1.0;
        const double g01 = fp_mulmod_signed(g[0], g[1], pd, pinv), g23 = fp_mulmod_signed(g[2], g[3], pd, pinv),
                     g45 = fp_mulmod_signed(g[4], g[5], pd, pinv);
        const double gl = fp_mulmod_signed(fp_mulmod_signed(g01, g23, pd, pinv), g45, pd, pinv);
        s_gl[lane] = (uint64_t)__double_as_longlong(fp_canon(gl, pd));
      }
      double v = fp_mulmod_signed((double)v0, fl, pd, pinv);
#pragma unroll
      for (int i = 6; i < kLgT; i++)
        if ((wave >> (i - 6)) & 1u) v = fp_mulmod_signed(v, __longlong_as_double(s_pw[0][i]), pd, pinv);
      vstart = fp_to_u64(fp_canon(v, pd));
    } else {
      uint64_t f[6];
#pragma unroll
      for (int i = 0; i < 6; i++) f[i] = ((lane >> i) & 1u) ? s_pw[0][i] : one;
      const uint64_t fl = mmul(mmul(mmul(f[0], f[1], p, q), mmul(f[2], f[3], p, q), p, q), mmul(f[4], f[5], p, q), p, q);
      if (tid < 64) s_pf[tid] = fl;
      if (wave == kW - 1) {
        uint64_t g[6];
#pragma unroll
        for (int i = 0; i < 6; i++) g[i] = ((lane >> i) & 1u) ? s_pw[1][i] : one;
        s_gl[lane] = mmul(mmul(mmul(g[0], g[1], p, q), mmul(g[2], g[3], p, q), p, q), mmul(g[4], g[5], p, q), p, q);
      }
      uint64_t v = mmul(v0, fl, p, q);
#pragma unroll
      for (int i = 6; i < kLgT; i++)
        if ((wave >> (i - 6)) & 1u) v = mmul(v, s_pw[0][i], p, q);
      vstart = v;
    }
    // Touch the next item's record and its first K target records' lines
    // (one 128-byte line per thread) so that they are in this XCD's L2 when
    // it starts; the value is only consumed by an empty asm at the end of the
    // item, so nothing waits for these loads on the way.
    uint32_t pf = 0;
    {
      const uint32_t nx = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_next);
      // the targets of a prime rarely exceed K (room is 2K): K records' lines
      const uint32_t kbytes = (uint32_t)K * (uint32_t)sizeof(KRec);
      if (nx != 0xFFFFFFFFu && (tid == 0 || 128u * (tid - 1) < kbytes)) {
        const uint32_t lx = nx - rfl(s_jc.item_base);
        const uint32_t px = merged ? lx : lx / 3;
        const uint32_t* a = tid == 0 ? (const uint32_t*)(R - pi + px)
                                     : (const uint32_t*)((const char*)(jkrecs + (size_t)px * kstride) + 128u * (tid - 1));
        pf = ld_global(a);
      }
    }
    // Baby steps: set the filter bits, hash W_j, the value of entry 64j
    // (lane 0's values: entry e = 64*wave + T*i is W_(wave + kW*i)), and find
    // ord_p(b) below the span (a repeat of v0).  A value whose low 32 bits
    // are 0 is inserted like any other: its probes pass the filter and are
    // rejected by the exact resolution, as the reference's table never holds
    // such a value (giant.cl:44-46, 61-63).
    if (fp) {
      // binary64 keys (kFpBias + residue: the filter reads residue bits below
      // 49 only), four chains of signed products stepping by b^(4*stride*T)
      const uint64_t k0 = v0 ^ kFpBias;
      const double pc = pd + kFpTwo52;
      auto visit = [&](uint32_t e, uint64_t key) {
        if (key == k0) atomicMin(&s_order, e * stride - 1u);  // entry 0 gives ~0: no repeat
        flt_insert<kFB>(flt, key);
      };
      double v[4];
      v[0] = (double)vstart;
#pragma unroll
      for (int c = 1; c < 4; c++) v[c] = fp_mulmod_signed(v[c - 1], bsd, pd, pinv);
      const double b2 = fp_mulmod_signed(bsd, bsd, pd, pinv), b4 = fp_mulmod_signed(b2, b2, pd, pinv);
      uint32_t e = tid, jw = wave;
      for (; e + 3 * T < nent; e += 4 * T, jw += 4 * kW) {
        uint64_t key[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          key[c] = fp_signed_bits(v[c], pc);
          visit(e + c * T, key[c]);
        }
        if (lane == 0) {
#pragma unroll
          for (int c = 0; c < 4; c++) w_put(jw + c * kW, key[c]);
        }
#pragma unroll
        for (int c = 0; c < 4; c++) v[c] = fp_mulmod_signed(v[c], b4, pd, pinv);
      }
#pragma unroll
      for (int c = 0; c < 3; c++) {
        if (e + c * T < nent) {
          const uint64_t key = fp_signed_bits(v[c], pc);
          visit(e + c * T, key);
          if (lane == 0) w_put(jw + c * kW, key);
        }
      }
    } else {
      baby_walk(nent, [&](uint32_t e, uint64_t v) {
        if (v == v0) atomicMin(&s_order, e * stride - 1u);
        flt_insert<kFB>(flt, v);
        if (lane == 0) w_put(e >> 6, v);  // e = 64 * (e >> 6) on lane 0 (T is a multiple of 64)
      });
    }
    __syncthreads();
    if (tid == 0) {  // the next item (everyone has read this item's values)
      fetch();
      s_dups[par ^ 1u] = s_claims[par ^ 1u] = 0;
    }
    // the first repeat's exponent offset (s_order holds it minus one), or ~0
    const uint32_t omin = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_order);
    const uint32_t order = omin == 0xFFFFFFFFu ? omin : omin + 1u;
    unsigned long long work, probes;
    if (order != 0xFFFFFFFFu) {
      // ord_p(b) (or its even multiple) is below the table span: the baby
      // values b^(nstart + stride*e) with stride*e < order are distinct and
      // hold every reachable residue (giant.cl:121-134); the targets are
      // matched against them directly, at most kQCap targets (the queue's
      // size) per pass over the baby values.
      for (uint32_t c0 = 0; c0 < nk; c0 += kQCap) {
        const uint32_t nc = min(nk - c0, kQCap);
        for (uint32_t i = tid; i < nc; i += T) {
          const uint64_t h = fp ? (uint64_t)__longlong_as_double((long long)s_h[c0 + i]) : s_h[c0 + i];
          s_q[i] = make_uint4((uint32_t)h, (uint32_t)(h >> 32), c0 + i, 0u);
        }
        __syncthreads();
        match(nc, (order + stride - 1) / stride, order);
        __syncthreads();  // the claims have read the queue before the next pass refills it
      }
      work = order / stride;  // the baby steps up to the repeat (giant.cl:121-134); no giant products
      probes = nk;
    } else {
      // The wavefront's tiles, one at a time: products, filter words, test.
      // The three workgroups a CU holds cover each other's LDS waits; a
      // software pipeline over two tiles (products of tile j while tile j-1's
      // words arrive) measured 5 % slower, the compiler copying its two
      // register sets at every step.
      auto giant = [&](const auto& ar) __attribute__((always_inline)) {
        using A = typename std::decay<decltype(ar)>::type;
        using Tg = typename A::Tgt;
        using Pw = typename A::Pow;
        // this wavefront's first tile and its giant power G^(rw*qb + lsub):
        // G^lsub from s_gl times the power table's entries for the bits of
        // the wavefront-uniform rw*qb (below T)
        uint32_t qb = wave / ng, g = wave % ng;
        Pw gq = ar.power(s_gl[lsub]);
#pragma unroll
        for (int i = 4; i < kLgT; i++)
          if (i >= (int)lw && ((qb >> (i - lw)) & 1u)) gq = ar.mul(gq, ar.power(s_pw[1][i]));
        const Pw grw = ar.power(s_pw[1][lw]);  // G^rw: one tile row further
        Tg tt[kU];
        auto load_t = [&]() {
#pragma unroll
          for (uint32_t k = 0; k < kU; k++) tt[k] = ar.target(s_h[__umul24(g * rg + lgrp, kU) + k]);
        };
        // products of the current tile (targets in tt), then advance to the
        // wavefront's next tile
        auto mul_t = [&](uint64_t(&uu)[kU], uint32_t& i0, uint32_t& qbt) {
#pragma unroll
          for (uint32_t k = 0; k < kU; k++) uu[k] = ar.key(tt[k], gq);
          i0 = __umul24(g * rg + lgrp, kU);  // (24-bit products: a full-rate multiply)
          qbt = qb;
          g = (uint32_t)__builtin_amdgcn_readfirstlane((int)(g + kW));
          while (g >= ng) {
            g -= ng;
            qb++;
            gq = ar.mul(gq, grw);
          }
        };
        auto read_w = [&](const uint64_t(&uu)[kU], uint32_t(&ww)[kU]) {
#pragma unroll
          for (uint32_t k = 0; k < kU; k++) ww[k] = *flt_word<kFB>(flt, uu[k]);
        };
        auto consume = [&](const uint64_t(&uu)[kU], const uint32_t(&ww)[kU], uint32_t i0, uint32_t qbt) {
          uint32_t x[kU], mn = 0xFFFFFFFFu;
#pragma unroll
          for (uint32_t k = 0; k < kU; k++) {
            x[k] = ~ww[k] & flt_bits(flt, uu[k]);
            mn = min(mn, x[k]);
          }
          const uint32_t qq = qbt * rw + lsub;
          if (mn == 0u && qq < Ml) {
#pragma unroll
            for (uint32_t k = 0; k < kU; k++)
              if (!x[k] && i0 + k < nk) defer(ar.value(uu[k]), i0 + k, qq);
          }
        };
        load_t();
        uint32_t r0 = 0;  // first step whose positives are not yet resolved
        for (uint32_t j0 = 0; j0 < nsteps; j0 += every) {
          const uint32_t j1 = min(nsteps, j0 + every);
          // this wavefront's tiles in steps [j0, j1)
          const uint32_t tf = j0 * kW + wave;
          const uint32_t nt = tf < ntiles ? min(j1 - j0, (ntiles - tf + kW - 1) / kW) : 0u;
          if (nt) {
            uint64_t uA[kU];
            uint32_t wA[kU], iA, qA;
            for (uint32_t n = 0; n < nt; n++) {
              mul_t(uA, iA, qA);
              load_t();
              read_w(uA, wA);
              consume(uA, wA, iA, qA);
            }
          }
          if (settle(r0, j1, j1 == nsteps)) break;  // the overflow recovery finished the item
        }
      };
      whash_build();
      if (fp) {
        GiantFp ar;
        ar.pd = pd;
        ar.pinv = pinv;
        ar.pc = pd + kFpTwo52;
        giant(ar);
      } else {
        GiantInt ar;
        ar.p = p;
        ar.q = q;
        ar.one = one;
        giant(ar);
      }
      work = (unsigned long long)nent + (unsigned long long)Ml * (nk + 1);
      probes = (unsigned long long)Ml * nk;
    }
    asm volatile("" ::"v"(pf));  // the prefetch's only use
    __syncthreads();
    if (tid == 0) {
      if (cnt != s_acc_cnt) {
        acc_flush();
        s_acc_cnt = cnt;
      }
      s_acc[0] += s_dups[par];
      s_acc[1] += s_claims[par];
      s_acc[2] += work;
      s_acc[3] += probes;
      s_acc[4 + c] += 1;
      if (order != 0xFFFFFFFFu) s_acc[7] += 1;
      s_qn = s_wkdone = s_qovf = 0;  // the next item's queue (it reaches its own only after its first barrier)
    }
  }
  if (tid == 0) acc_flush();
}

// Work items of every job in prefix order; one thread (a handful of jobs).
__global__ void k_round_items(SegJob* __restrict__ jobs, BatchState* __restrict__ bs, uint32_t njobs) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t base = 0;
  for (uint32_t j = 0; j < njobs; j++) {
    bs->item_base[j] = base;
    const uint32_t np = jobs[j].nprimes < jobs[j].cap ? jobs[j].nprimes : jobs[j].cap;
    base += np * (jobs[j].job->algo != 0 ? 1u : 3u);
  }
  bs->item_base[njobs] = base;
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
// static LDS of the BSGS kernel (target lists, positive queue, counters),
// queried once per workgroup size
hipError_t bsgs_static_lds(uint32_t threads, size_t* out) {
  static std::atomic<size_t> cache[4];  // 0: not queried yet (the static LDS is never empty)
  const int slot = threads == 128 ? 0 : threads == 256 ? 1 : threads == 512 ? 2 : threads == 1024 ? 3 : -1;
  if (slot < 0) return hipErrorInvalidValue;
  size_t v = cache[slot].load(std::memory_order_relaxed);
  if (!v) {
    hipFuncAttributes a;
    hipError_t e;
    switch (threads) {
      case 128: e = hipFuncGetAttributes(&a, (const void*)k_bsgs<128>); break;
      case 256: e = hipFuncGetAttributes(&a, (const void*)k_bsgs<256>); break;
      case 512: e = hipFuncGetAttributes(&a, (const void*)k_bsgs<512>); break;
      default: e = hipFuncGetAttributes(&a, (const void*)k_bsgs<1024>); break;
    }
    if (e != hipSuccess) return e;
    v = a.sharedSizeBytes;
    cache[slot].store(v, std::memory_order_relaxed);
  }
  *out = v;
  return hipSuccess;
}

// LDS of one BSGS workgroup in all: the filter ends at its base + 4*hsize
size_t bsgs_block_lds_bytes(uint32_t hsize, uint32_t threads) {
  return (size_t)bsgs_filter_base(threads) + (size_t)hsize * 4;
}

// dynamic LDS of a BSGS workgroup: from the end of the static LDS to the end
// of the filter (the padding below the filter's base included)
size_t bsgs_lds_bytes(uint32_t hsize, uint32_t threads) {
  size_t st = 0;
  if (bsgs_static_lds(threads, &st) != hipSuccess || st > bsgs_filter_base(threads)) st = 0;  // refused at rs_open
  return bsgs_block_lds_bytes(hsize, threads) - st;
}

// every BSGS kernel's static LDS must end below its filter's fixed base; the
// largest excess (or 0) in *worst
hipError_t bsgs_check_static_lds(size_t* worst) {
  size_t w = 0;
  hipError_t r = hipSuccess;
  for (uint32_t t : {128u, 256u, 512u, 1024u}) {
    size_t st = 0;
    const hipError_t e = bsgs_static_lds(t, &st);
    if (e != hipSuccess) return e;
    if (st > bsgs_filter_base(t)) {
      r = hipErrorInvalidKernelFile;
      w = st - bsgs_filter_base(t) > w ? st - bsgs_filter_base(t) : w;
    }
  }
  *worst = w;
  return r;
}

hipError_t bsgs_blocks_per_cu(uint32_t threads, uint32_t hsize, int* out) {
  const size_t lds = bsgs_lds_bytes(hsize, threads);
  switch (threads) {
    case 128: return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, k_bsgs<128>, 128, lds);
    case 256: return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, k_bsgs<256>, 256, lds);
    case 512: return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, k_bsgs<512>, 512, lds);
    case 1024: return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, k_bsgs<1024>, 1024, lds);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_set_cursor(SegJob* sj, uint64_t pmin, uint64_t pmax, uint32_t threads, uint64_t range,
                             uint32_t gen_blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_set_cursor, dim3(1), dim3(64), 0, s, sj, pmin, pmax, threads, range, gen_blocks);
  return hipGetLastError();
}

// One round of a batch in two parts: the preparation (round_begin, primegen,
// setup, round_items) and the BSGS launch, so a caller can time the BSGS
// launch with events between them.
hipError_t launch_round_prep(SegJob* jobs, BatchState* bs, uint32_t njobs, uint32_t gen_blocks, uint32_t setup_blocks,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_round_begin, dim3((njobs + 63) / 64), dim3(64), 0, s, jobs, bs, njobs);
  if (gen_blocks) hipLaunchKernelGGL(k_primegen, dim3(gen_blocks, njobs), dim3(256), 0, s, jobs);
  // k_setup takes four primes per workgroup, k_setup_pre 256
  hipLaunchKernelGGL(k_setup_pre, dim3((4 * setup_blocks + 255) / 256, njobs), dim3(256), 0, s, jobs);
  hipLaunchKernelGGL(k_setup, dim3(setup_blocks, njobs), dim3(256), 0, s, jobs);
  hipLaunchKernelGGL(k_round_items, dim3(1), dim3(64), 0, s, jobs, bs, njobs);
  return hipGetLastError();
}

hipError_t launch_round_bsgs(SegJob* jobs, BatchState* bs, uint32_t threads, uint32_t hsize, uint32_t bsgs_blocks,
                             hipStream_t s) {
  const size_t lds = bsgs_lds_bytes(hsize, threads);
  switch (threads) {
    case 128: hipLaunchKernelGGL(k_bsgs<128>, dim3(bsgs_blocks), dim3(128), lds, s, jobs, bs); break;
    case 256: hipLaunchKernelGGL(k_bsgs<256>, dim3(bsgs_blocks), dim3(256), lds, s, jobs, bs); break;
    case 512: hipLaunchKernelGGL(k_bsgs<512>, dim3(bsgs_blocks), dim3(512), lds, s, jobs, bs); break;
    case 1024: hipLaunchKernelGGL(k_bsgs<1024>, dim3(bsgs_blocks), dim3(1024), lds, s, jobs, bs); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Static scratch of the allocation-free sieve (rs_sieve_noalloc): the BSGS
// overflow-recovery tables of every workgroup live in this file-scope device
// array instead of context memory.  Every GPU the library runs on holds its
// own copy from the module load until the process ends; host code uses the
// address hipGetSymbolAddress returns for the current device (the array's
// name on the host is only a stand-in), resolved once per device by
// static_scratch_address at context set-up.  The recovery kernel clears the
// part it uses before every use, so nothing in it needs to start at zero.
// All calls on one GPU share it.
// ---------------------------------------------------------------------------
constexpr size_t kStaticScratchBytes = RS_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) uint32_t scratch[kStaticScratchBytes / 4];

size_t static_scratch_bytes() { return kStaticScratchBytes; }

hipError_t static_scratch_address(int device, uint32_t** out) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static uint32_t* addr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {  // first context on this device: the caller made it current
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    addr[device] = (uint32_t*)p;
  }
  *out = addr[device];
  return hipSuccess;
}

// Points a launch's work queue at a recovery scratch of 'words' 32-bit words
// per workgroup (a kernel rather than a copy, so the call records no host
// buffer and can be captured).
__global__ void k_set_scratch(BatchState* __restrict__ bs, uint32_t* ptr, uint64_t words) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    bs->scratch = ptr;
    bs->scratch_words = words;
  }
}

hipError_t launch_set_scratch(BatchState* bs, uint32_t* ptr, uint64_t words, hipStream_t s) {
  hipLaunchKernelGGL(k_set_scratch, dim3(1), dim3(64), 0, s, bs, ptr, words);
  return hipGetLastError();
}

// One round with events around its BSGS launch (either may be null).
hipError_t launch_round(SegJob* jobs, BatchState* bs, uint32_t njobs, uint32_t gen_blocks, uint32_t setup_blocks,
                        uint32_t threads, uint32_t hsize, uint32_t bsgs_blocks, hipEvent_t ev0, hipEvent_t ev1,
                        hipStream_t s) {
  hipError_t e = launch_round_prep(jobs, bs, njobs, gen_blocks, setup_blocks, s);
  if (e == hipSuccess && ev0) e = hipEventRecord(ev0, s);
  if (e == hipSuccess) e = launch_round_bsgs(jobs, bs, threads, hsize, bsgs_blocks, s);
  if (e == hipSuccess && ev1) e = hipEventRecord(ev1, s);
  return e;
}

// The generator alone (rs_primegen): round_begin + primegen for one job.
hipError_t launch_primegen_round(SegJob* jobs, BatchState* bs, uint32_t gen_blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_round_begin, dim3(1), dim3(64), 0, s, jobs, bs, 1u);
  if (gen_blocks) hipLaunchKernelGGL(k_primegen, dim3(gen_blocks, 1), dim3(256), 0, s, jobs);
  return hipGetLastError();
}

}  // namespace rs
