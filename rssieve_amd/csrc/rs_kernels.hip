// rs_kernels.hip -- HIP kernels for the k*b^n+-1 BSGS sieve on gfx950.
//
// One p-segment runs as: seg_begin -> primegen -> setup -> bsgs, all enqueued
// on one stream with no host round trip (sizes are read on the device).
//
//  primegen  all base-2 strong probable primes in [low, high) with no prime
//            factor <= 113 (the reference's getsegprimes.cl:128-182 output set)
//  setup     one wavefront per prime: Montgomery constants, batch inversion of
//            the k by wavefront product scans, power-residue / quadratic-
//            character prefilter (setup.cl:11-209) and per-class compaction
//            (the reference's sort.cl, folded in)
//  bsgs      persistent workgroups pull (prime, class) items from a device
//            work queue; the baby-step table lives entirely in LDS (32-bit
//            fingerprint + 16-bit index per slot, no global spill), giant
//            steps probe it (giant.cl:77-323)
#include "rs_internal.h"
#include "rs_mont.h"

namespace rs {

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

// Base-2 strong probable-prime test (getsegprimes.cl:37-66), same shift
// semantics: N = 2^t + 1 squares 64 times before the final doubling.
__device__ bool sprp2(uint64_t N) {
  const uint64_t q = mont_inv(N);
  const uint64_t one = r_mod(N);
  const uint64_t nmo = N - one;
  const uint64_t two = madd(one, one, N);
  const int t = __ffsll((long long)(N - 1)) - 1;
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

__global__ void k_seg_begin(SegState* seg) {
  if (threadIdx.x == 0) {
    seg->nprimes = 0;
    seg->work_next = 0;
  }
}

// ---------------------------------------------------------------------------
// primegen: work-item gid covers [low + 60 gid, low + 60 gid + 60) with the
// mod-30 wheel; survivors of the small-prime bit sieve are packed in LDS, then
// every lane of the workgroup tests them; passing primes are appended with one
// global atomic per wavefront.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_primegen(uint64_t low, uint64_t high, int32_t widx,
                                                  uint64_t* __restrict__ out, uint32_t cap,
                                                  SegState* __restrict__ seg) {
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
      ok = sprp2(N);
    }
    const uint64_t m = __ballot(ok);
    if (m) {
      const uint32_t leader = __ffsll((long long)m) - 1;
      uint32_t wbase = 0;
      if (lane == leader) wbase = atomicAdd(&seg->nprimes, (uint32_t)__popcll(m));
      wbase = __shfl(wbase, leader);
      if (ok) {
        const uint32_t w = wbase + lane_rank(m);
        if (w < cap) out[w] = N;
      }
    }
  }
}

// primes <= 113 in [low, high) (addsmallprimes.cl:8-27)
__global__ void k_smallprimes(uint64_t low, uint64_t high, uint64_t* __restrict__ out, uint32_t cap,
                              SegState* __restrict__ seg) {
  const uint32_t small[30] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47,
                              53, 59, 61, 67, 71, 73, 79, 83, 89, 97, 101, 103, 107, 109, 113};
  const uint32_t i = threadIdx.x;
  if (i >= 30) return;
  const uint64_t p = small[i];
  if (p < low || p >= high) return;
  const uint32_t w = atomicAdd(&seg->nprimes, 1u);
  if (w < cap) out[w] = p;
}

// ---------------------------------------------------------------------------
// setup: one wavefront per prime.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t shfl_up64(uint64_t v, int d) { return __shfl_up(v, d, 64); }
__device__ __forceinline__ uint64_t shfl_dn64(uint64_t v, int d) { return __shfl_down(v, d, 64); }
__device__ __forceinline__ uint64_t shfl64(uint64_t v, int l) { return __shfl(v, l, 64); }

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

// d of the power-residue tests (setup.cl:22-48)
__constant__ uint32_t g_prd[12] = {3, 4, 5, 7, 8, 9, 11, 13, 17, 19, 23, 29};

__global__ __launch_bounds__(256) void k_setup(const JobDev* __restrict__ job, const uint64_t* __restrict__ primes,
                                               const SegState* __restrict__ seg, PrimeRec* __restrict__ recs,
                                               KRec* __restrict__ krecs, Counters* __restrict__ cnt, uint32_t cap) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t pi = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t ngen = seg->nprimes;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    atomicAdd((unsigned long long*)&cnt->primecount, (unsigned long long)ngen);
    atomicMax((unsigned long long*)&cnt->max_seg_primes, (unsigned long long)ngen);
    if (ngen > cap) atomicAdd((unsigned long long*)&cnt->seg_overflow, 1ull);
  }
  const uint32_t np = ngen < cap ? ngen : cap;
  if (pi >= np) return;

  const uint32_t base = job->base;
  const int K = job->kcount;
  const uint32_t T = job->threads;
  const uint64_t p = primes[pi];
  const uint64_t q = mont_inv(p);
  const uint64_t one = r_mod(p);
  const uint64_t nmo = p - one;
  const uint64_t two = madd(one, one, p);
  const uint64_t four = madd(two, two, p);
  uint64_t r2 = four;
#pragma unroll
  for (int i = 0; i < 5; i++) r2 = mmul(r2, r2, p, q);  // 4^(2^5) = 2^64: Montgomery form of R
  uint64_t mbase;
  if (base == 2) mbase = two;
  else if (base == 3) mbase = madd(one, two, p);
  else if (base == 5) mbase = madd(four, one, p);
  else mbase = mmul(base, r2, p, q);

  // Montgomery k for k = lane and k = lane + 64; 'one' pads the products.
  const uint64_t mk0 = (int)lane < K ? mmul(job->ak[lane], r2, p, q) : one;
  const uint64_t mk1 = (int)lane + 64 < K ? mmul(job->ak[lane + 64], r2, p, q) : one;
  const uint64_t gQ = basepow_sm(mbase, job->Q, base, p, q, one);

  // inclusive prefix (s) and suffix (u) products over the wavefront
  uint64_t s0 = mk0, s1 = mk1, u0 = mk0, u1 = mk1;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t a0 = shfl_up64(s0, d), a1 = shfl_up64(s1, d);
    const uint64_t b0 = shfl_dn64(u0, d), b1 = shfl_dn64(u1, d);
    if ((int)lane >= d) {
      s0 = mmul(s0, a0, p, q);
      s1 = mmul(s1, a1, p, q);
    }
    if ((int)lane + d < 64) {
      u0 = mmul(u0, b0, p, q);
      u1 = mmul(u1, b1, p, q);
    }
  }
  const uint64_t A = shfl64(s0, 63), B = shfl64(s1, 63);
  const uint64_t ep0 = shfl_up64(s0, 1), ep1 = shfl_up64(s1, 1);
  const uint64_t es0 = shfl_dn64(u0, 1), es1 = shfl_dn64(u1, 1);
  // products of all other k: before and after, for both slots of this lane
  const uint64_t pre0 = lane ? ep0 : one;
  const uint64_t pre1 = mmul(A, lane ? ep1 : one, p, q);
  const uint64_t suf0 = mmul(lane < 63 ? es0 : one, B, p, q);
  const uint64_t suf1 = lane < 63 ? es1 : one;
  const uint64_t AB = mmul(A, B, p, q);
  const uint64_t total = mmul(AB, gQ, p, q);  // prefix[KCOUNT] of setup.cl:147-151

  // Lane-parallel exponentiations, one code path (no serialized branches):
  //  lane 0: Fermat inverse total^(p-2)          (setup.cl:154)
  //  lanes 1..12: base^((p-1)/d) for d | p-1       (setup.cl:22-48)
  //  lane 13: base^((p-1)/2)                       (setup.cl:51-53)
  const uint64_t pm = p - 1;
  uint64_t ex = 0, x = mbase;
  bool dvalid = false;
  if (lane == 0) {
    ex = p - 2;
    x = total;
  } else if (lane <= 12) {
    const uint32_t d = g_prd[lane - 1];
    ex = pm / d;
    dvalid = (ex * d == pm);
    if (!dvalid) ex = 0;
  } else if (lane == 13) {
    ex = pm >> 1;
  }
  uint64_t r = one;
  if (ex) r = powmod_lg(x, ex, p, q);
  const uint64_t inv_total = shfl64(r, 0);
  const bool invertible = mmul(total, inv_total, p, q) == one;  // setup.cl:156-162
  const uint64_t dmask = __ballot(dvalid && r == one) >> 1;      // passing d tests, bit i = g_prd[i]
  const uint64_t rg = shfl64(r, 13);
  const int resg = (rg == one) - (rg == nmo);

  PrimeRec* R = recs + pi;
  if (!invertible) {
    if (lane == 0) {
      R->nk[0] = R->nk[1] = R->nk[2] = 0;
      atomicAdd((unsigned long long*)&cnt->skipped, 1ull);
    }
    return;
  }
  const uint64_t gQ_inv = mmul(inv_total, AB, p, q);  // setup.cl:164-165
  const uint64_t inv_k = mmul(inv_total, gQ, p, q);   // (prod of all k)^-1

  // k^-1 and the target hadj for both slots (setup.cl:180-188)
  const uint64_t kinv0 = mmul(inv_k, mmul(pre0, suf0, p, q), p, q);
  const uint64_t kinv1 = mmul(inv_k, mmul(pre1, suf1, p, q), p, q);
  const bool act0 = (int)lane < K, act1 = (int)lane + 64 < K;
  const uint64_t h0 = (act0 && job->cpos[lane]) ? mmul(nmo, kinv0, p, q) : kinv0;
  const uint64_t h1 = (act1 && job->cpos[lane + 64]) ? mmul(nmo, kinv1, p, q) : kinv1;

  // prefilter (setup.cl:59-87); the order of the tests does not change the result
  int cls0 = act0 ? 1 : 0, cls1 = act1 ? 1 : 0;
  uint64_t dm = dmask;
  while (dm) {
    const int i = __ffsll((long long)dm) - 1;
    dm &= dm - 1;
    const uint64_t e = shfl64(ex, i + 1);
    uint64_t ra0, ra1;
    powmod_lg2(h0, h1, e, p, q, &ra0, &ra1);
    if (ra0 != one) cls0 = 0;
    if (ra1 != one) cls1 = 0;
  }
  {
    uint64_t ra0, ra1;
    powmod_lg2(h0, h1, pm >> 1, p, q, &ra0, &ra1);
    if (cls0) {
      if (resg == -1) cls0 = (ra0 == one) ? 2 : (ra0 == nmo) ? 3 : 0;
      else if (resg == 1 && ra0 == nmo) cls0 = 0;
    }
    if (cls1) {
      if (resg == -1) cls1 = (ra1 == one) ? 2 : (ra1 == nmo) ? 3 : 0;
      else if (resg == 1 && ra1 == nmo) cls1 = 0;
    }
  }

  // compaction by class: full, even, odd (the reference's sort.cl:40-71)
  uint32_t nk[3];
  uint32_t off = 0;
  KRec* kr = krecs + (size_t)pi * K;
#pragma unroll
  for (int c = 1; c <= 3; c++) {
    const uint64_t m0 = __ballot(cls0 == c), m1 = __ballot(cls1 == c);
    const uint32_t c0 = __popcll(m0), c1 = __popcll(m1);
    if (cls0 == c) {
      KRec e;
      e.hadj = h0;
      e.kidx = lane;
      e.cls = c;
      kr[off + lane_rank(m0)] = e;
    }
    if (cls1 == c) {
      KRec e;
      e.hadj = h1;
      e.kidx = lane + 64;
      e.cls = c;
      kr[off + c0 + lane_rank(m1)] = e;
    }
    nk[c - 1] = c0 + c1;
    off += c0 + c1;
  }
  if (off == 0) {  // every k impossible (setup.cl:205-208)
    if (lane == 0) {
      R->nk[0] = R->nk[1] = R->nk[2] = 0;
      atomicAdd((unsigned long long*)&cnt->skipped, 1ull);
    }
    return;
  }

  // Constants of the BSGS kernel, lane-parallel:
  //  lane 0: b^NMIN; 1: (b^-Q)^T; 2: (b^-QQ)^T; 3: b^T; 4: b^(2T)
  const uint64_t g1 = mmul(gQ_inv, gQ_inv, p, q);
  uint64_t y = one, yb = one;
  uint32_t ye = 0;
  if (lane == 0) { yb = mbase; ye = job->nmin; }
  else if (lane == 1) { yb = gQ_inv; ye = T; }
  else if (lane == 2) { yb = g1; ye = T; }
  else if (lane == 3) { yb = mbase; ye = T; }
  else if (lane == 4) { yb = mbase; ye = 2 * T; }
  if (lane < 5) y = powmod_sm(yb, ye, p, q, one);
  const uint64_t bn = shfl64(y, 0), gT0 = shfl64(y, 1), gT1 = shfl64(y, 2);
  const uint64_t bT0 = shfl64(y, 3), bT1 = shfl64(y, 4);
  if (lane == 0) {
    PrimeRec rec;
    rec.p = p;
    rec.q = q;
    rec.one = one;
    rec.mbase = mbase;
    rec.bn = bn;
    rec.g[0] = gQ_inv;
    rec.g[1] = g1;
    rec.gT[0] = gT0;
    rec.gT[1] = gT1;
    rec.bT[0] = bT0;
    rec.bT[1] = bT1;
    rec.nk[0] = nk[0];
    rec.nk[1] = nk[1];
    rec.nk[2] = nk[2];
    rec.flags = 0;
    rec.pad[0] = rec.pad[1] = rec.pad[2] = 0;
    *R = rec;
  }
}

// ---------------------------------------------------------------------------
// bsgs (three lists, the reference's giantfull/giantparity shape)
// ---------------------------------------------------------------------------
struct Tally {
  uint32_t dups, claims;
  uint64_t work, probes;
};

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

// Insert baby value v with exponent offset j.  A value whose low 32 bits are 0
// is left out: the reference marks empty slots with 0 and never finds it
// (giant.cl:44-46, 61-63).
__device__ __forceinline__ void tab_insert(uint32_t* fp, uint16_t* ix, uint32_t hmask, uint64_t v, uint32_t j) {
  const uint32_t lo = (uint32_t)v;
  if (lo == 0) return;
  uint32_t s = lo & hmask;
  while (true) {
    const uint32_t old = atomicCAS(&fp[s], 0u, lo);
    if (old == 0u) {
      ix[s] = (uint16_t)j;
      return;
    }
    s = (s + 1) & hmask;
  }
}

// Find exponent offset j with baby value == g, or -1.  Slots hold the low 32
// bits; when p >= 2^32 a match is confirmed by recomputing b^(nstart + j).
__device__ __forceinline__ int tab_lookup(const uint32_t* fp, const uint16_t* ix, uint32_t hmask, uint64_t g,
                                          uint64_t mbase, uint32_t base, uint32_t nstart, uint64_t p, uint64_t q,
                                          uint64_t one) {
  const uint32_t lo = (uint32_t)g;
  uint32_t s = lo & hmask;
  while (true) {
    const uint32_t f = fp[s];
    if (f == 0u) return -1;
    if (f == lo) {
      const uint32_t j = ix[s];
      if ((p >> 32) == 0) return (int)j;
      if (basepow_sm(mbase, nstart + j, base, p, q, one) == g) return (int)j;
    }
    s = (s + 1) & hmask;
  }
}

template <int T>
__global__ __launch_bounds__(T) void k_bsgs3(const JobDev* __restrict__ job, const PrimeRec* __restrict__ recs,
                                             const KRec* __restrict__ krecs, SegState* __restrict__ seg,
                                             uint32_t* __restrict__ bitmap, FactorRec* __restrict__ fac,
                                             Counters* __restrict__ cnt, uint32_t cap) {
  extern __shared__ uint32_t lds[];
  const uint32_t hsize = job->hsize, hmask = job->hmask;
  uint32_t* fp = lds;
  uint16_t* ix = (uint16_t*)(lds + hsize);
  __shared__ uint32_t s_item, s_order, s_dups, s_claims;
  __shared__ unsigned long long s_work, s_probe;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) {
    s_dups = 0;
    s_claims = 0;
    s_work = 0;
    s_probe = 0;
  }
  const uint32_t ngen = seg->nprimes;
  const uint32_t np = ngen < cap ? ngen : cap;
  const uint32_t nitems = np * 3;
  const uint32_t base = job->base;
  const int K = job->kcount;
  const uint32_t nmin = job->nmin, nmax = job->nmax;

  for (;;) {
    if (tid == 0) s_item = atomicAdd(&seg->work_next, 1u);
    __syncthreads();
    const uint32_t item = s_item;
    __syncthreads();
    if (item >= nitems) break;
    const uint32_t pi = item / 3, c = item % 3;
    const PrimeRec* __restrict__ R = recs + pi;
    const uint32_t nk = R->nk[c];
    if (nk == 0) continue;
    const uint32_t cp = c ? 1 : 0;
    const uint64_t p = R->p, q = R->q, one = R->one, mbase = R->mbase;
    const uint32_t Ql = cp ? job->QQ : job->Q;
    const uint32_t Ml = cp ? job->MM : job->M;
    const uint32_t stride = cp ? 2 : 1;
    const uint32_t nstart = nmin + (c == 2 ? 1 : 0);
    const uint32_t nent = Ql / stride;
    const uint64_t v0 = (c == 2) ? mul_base(R->bn, mbase, base, p, q) : R->bn;
    const uint64_t bstep = R->bT[cp];
    const uint32_t koff = (c >= 1 ? R->nk[0] : 0) + (c >= 2 ? R->nk[1] : 0);
    const KRec* __restrict__ kr = krecs + (size_t)pi * K + koff;

    for (uint32_t s = tid; s < hsize; s += T) fp[s] = 0u;
    if (tid == 0) s_order = 0xFFFFFFFFu;
    __syncthreads();
    // baby steps: entry e = tid + T*i holds b^(nstart + stride*e)
    const uint64_t vstart = mmul(v0, basepow_sm(mbase, stride * tid, base, p, q, one), p, q);
    {
      uint64_t v = vstart;
      for (uint32_t e = tid; e < nent; e += T) {
        const uint32_t j = e * stride;
        if (j && v == v0) atomicMin(&s_order, j);
        tab_insert(fp, ix, hmask, v, j);
        v = mmul(v, bstep, p, q);
      }
    }
    __syncthreads();
    const uint32_t order = s_order;
    if (order != 0xFFFFFFFFu) {
      // ord_p(b) (or its even multiple) is below the table size: keep only
      // j < order (giant.cl:121-134), then solve without giant steps.
      __syncthreads();
      for (uint32_t s = tid; s < hsize; s += T) fp[s] = 0u;
      __syncthreads();
      uint64_t v = vstart;
      for (uint32_t e = tid; e < nent; e += T) {
        const uint32_t j = e * stride;
        if (j < order) tab_insert(fp, ix, hmask, v, j);
        v = mmul(v, bstep, p, q);
      }
      __syncthreads();
      if (tid < nk) {
        const int j = tab_lookup(fp, ix, hmask, kr[tid].hadj, mbase, base, nstart, p, q, one);
        if (j >= 0) {
          const uint32_t kidx = kr[tid].kidx;
          for (uint64_t n = (uint64_t)nstart + (uint32_t)j; n <= nmax; n += order)
            claim(job, bitmap, fac, cnt, p, kidx, (uint32_t)n, &s_dups, &s_claims);
        }
      }
      if (tid == 0) {
        atomicAdd(&s_work, (unsigned long long)(order / stride + nk));
        atomicAdd(&s_probe, (unsigned long long)nk);
        atomicAdd((unsigned long long*)&cnt->order_lists, 1ull);
      }
    } else {
      // giant steps: thread tid handles q = tid, tid + T, ...
      const uint64_t g = R->g[cp], gT = R->gT[cp];
      uint64_t Gq = powmod_sm(g, tid, p, q, one);
      for (uint32_t qq = tid; qq < Ml; qq += T) {
        const uint32_t nq = nstart + qq * Ql;
        for (uint32_t i = 0; i < nk; i++) {
          const uint64_t gamma = mmul(kr[i].hadj, Gq, p, q);
          const int j = tab_lookup(fp, ix, hmask, gamma, mbase, base, nstart, p, q, one);
          if (j >= 0) claim(job, bitmap, fac, cnt, p, kr[i].kidx, nq + (uint32_t)j, &s_dups, &s_claims);
        }
        Gq = mmul(Gq, gT, p, q);
      }
      if (tid == 0) {
        atomicAdd(&s_work, (unsigned long long)nent + (unsigned long long)Ml * (nk + 1));
        atomicAdd(&s_probe, (unsigned long long)Ml * nk);
      }
    }
    if (tid == 0) atomicAdd((unsigned long long*)&cnt->lists[c], 1ull);
    __syncthreads();
  }
  __syncthreads();
  if (tid == 0) {
    if (s_dups) atomicAdd((unsigned long long*)&cnt->dups, (unsigned long long)s_dups);
    if (s_claims) atomicAdd((unsigned long long*)&cnt->claims, (unsigned long long)s_claims);
    atomicAdd((unsigned long long*)&cnt->work_mmul, s_work);
    atomicAdd((unsigned long long*)&cnt->work_probe, s_probe);
  }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
hipError_t launch_seg_begin(SegState* seg, hipStream_t s) {
  hipLaunchKernelGGL(k_seg_begin, dim3(1), dim3(64), 0, s, seg);
  return hipGetLastError();
}

hipError_t launch_smallprimes(uint64_t low, uint64_t high, uint64_t* out, uint32_t cap, SegState* seg,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_smallprimes, dim3(1), dim3(64), 0, s, low, high, out, cap, seg);
  return hipGetLastError();
}

hipError_t launch_primegen(uint64_t low, uint64_t high, int32_t widx, uint64_t* out, uint32_t cap, SegState* seg,
                           hipStream_t s) {
  if (high <= low) return hipSuccess;
  const uint64_t nthreads = (high - low) / 60 + 1;
  const uint64_t nblocks = (nthreads + 255) / 256;
  if (nblocks > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_primegen, dim3((uint32_t)nblocks), dim3(256), 0, s, low, high, widx, out, cap, seg);
  return hipGetLastError();
}

hipError_t launch_setup(const JobDev* job, const uint64_t* primes, const SegState* seg, PrimeRec* recs, KRec* krecs,
                        Counters* cnt, uint32_t cap, hipStream_t s) {
  const uint32_t nblocks = (cap + 3) / 4;
  hipLaunchKernelGGL(k_setup, dim3(nblocks), dim3(256), 0, s, job, primes, seg, recs, krecs, cnt, cap);
  return hipGetLastError();
}

size_t bsgs_lds_bytes(uint32_t hsize) { return (size_t)hsize * 6; }

hipError_t bsgs_blocks_per_cu(uint32_t threads, uint32_t hsize, int* out) {
  const size_t lds = bsgs_lds_bytes(hsize);
  switch (threads) {
    case 128: return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, k_bsgs3<128>, (int)threads, lds);
    case 256: return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, k_bsgs3<256>, (int)threads, lds);
    case 512: return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, k_bsgs3<512>, (int)threads, lds);
    case 1024: return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, k_bsgs3<1024>, (int)threads, lds);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_bsgs(uint32_t threads, uint32_t hsize, uint32_t nblocks, const JobDev* job, const PrimeRec* recs,
                       const KRec* krecs, SegState* seg, uint32_t* bitmap, FactorRec* fac, Counters* cnt,
                       uint32_t cap, hipStream_t s) {
  const size_t lds = bsgs_lds_bytes(hsize);
  switch (threads) {
    case 128:
      hipLaunchKernelGGL(k_bsgs3<128>, dim3(nblocks), dim3(128), lds, s, job, recs, krecs, seg, bitmap, fac, cnt, cap);
      break;
    case 256:
      hipLaunchKernelGGL(k_bsgs3<256>, dim3(nblocks), dim3(256), lds, s, job, recs, krecs, seg, bitmap, fac, cnt, cap);
      break;
    case 512:
      hipLaunchKernelGGL(k_bsgs3<512>, dim3(nblocks), dim3(512), lds, s, job, recs, krecs, seg, bitmap, fac, cnt, cap);
      break;
    case 1024:
      hipLaunchKernelGGL(k_bsgs3<1024>, dim3(nblocks), dim3(1024), lds, s, job, recs, krecs, seg, bitmap, fac, cnt,
                         cap);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace rs
