/*
 * rs_oracle.c -- CPU oracle for the k*b^n+-1 BSGS sieve path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (rssieve_amd/, the CLI,
 * the C-ABI library) links, loads or calls this file.  It is used by tests/,
 * by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and only as
 * the checker or the timed CPU baseline.
 *
 * It restates, in plain C with unsigned __int128, the algorithm of the
 * reference (mfl0p/RSSieve v1.3).  Every function cites the reference
 * file:line it follows.  Semantics reproduced on purpose, including the
 * reference's quirks:
 *   - OpenCL masks shift counts to 6 bits, so a left-to-right powmod whose
 *     exponent is 1 squares 64 times (common.cl:52-53, getsegprimes.cl:40-41);
 *   - a baby-step value whose low 32 bits are 0 is never found by a lookup
 *     (giant.cl:44-46, 61-63: 0 marks an empty LDS slot);
 *   - the signed 'int klist[]' encoding of k and c (inputfile.cpp:366), which
 *     wraps for k >= 2^31, when the caller asks for it (k31_compat).
 *
 * Parity status of each function is listed in the header comment of
 * tests/test_oracle.py, which holds the checks that pin it.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <pthread.h>

typedef unsigned __int128 u128;

#define OR_MAX_SEQ 100 /* cl_sieve.h:2 */

/* ------------------------------------------------------------------------ */
/* Montgomery arithmetic, R = 2^64 (common.cl:24-176, verify_factor.c:14-50) */
/* ------------------------------------------------------------------------ */

/* common.cl:24-28: Newton iteration for p^-1 mod 2^64 */
uint64_t or_invert(uint64_t p) {
  uint64_t p_inv = 1, prev = 0;
  while (p_inv != prev) {
    prev = p_inv;
    p_inv *= 2 - p * p_inv;
  }
  return p_inv;
}

/* common.cl:33-40: a*b*2^-64 mod p with q = p^-1 mod 2^64 (not -p^-1) */
uint64_t or_mmul(uint64_t a, uint64_t b, uint64_t p, uint64_t q) {
  u128 ab = (u128)a * b;
  uint64_t lo = (uint64_t)ab, hi = (uint64_t)(ab >> 64);
  uint64_t m = lo * q;
  uint64_t mp = (uint64_t)(((u128)m * p) >> 64);
  uint64_t r = hi - mp;
  return (hi < mp) ? r + p : r;
}

/* common.cl:42-47 */
uint64_t or_add(uint64_t a, uint64_t b, uint64_t p) {
  uint64_t c = (a >= p - b) ? p : 0;
  return a + b - c;
}

/* count leading zeros with the OpenCL convention clz(0) = width */
static int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }
static int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }

/* OpenCL 'x >> s' for 64-bit x: the shift count is taken modulo 64 */
static uint64_t shr64_cl(uint64_t x, int s) { return x >> (s & 63); }
static uint32_t shr32_cl(uint32_t x, int s) { return x >> (s & 31); }

/* Multiply a Montgomery residue by the sequence base b
 * (common.cl:73-85: adds for b = 2, 3, 5, a Montgomery product otherwise). */
static uint64_t mul_base(uint64_t a, uint64_t mbase, uint32_t base, uint64_t p, uint64_t q) {
  if (base == 2) return or_add(a, a, p);
  if (base == 3) { uint64_t b2 = or_add(a, a, p); return or_add(a, b2, p); }
  if (base == 5) { uint64_t b2 = or_add(a, a, p); b2 = or_add(b2, b2, p); return or_add(a, b2, p); }
  return or_mmul(a, mbase, p, q);
}

/* common.cl:51-63 (exp > 1 assumed by the reference; exp == 1 squares 64x) */
uint64_t or_powmodlg(uint64_t mbase, uint64_t exp, uint64_t p, uint64_t q) {
  uint64_t curBit = shr64_cl(0x8000000000000000ull, clz64(exp) + 1);
  uint64_t a = mbase;
  /* With a masked shift count of 0 (exp == 1) the loop starts at bit 63. */
  while (curBit) {
    a = or_mmul(a, a, p, q);
    if (exp & curBit) a = or_mmul(a, mbase, p, q);
    curBit >>= 1;
  }
  return a;
}

/* common.cl:67-90 */
uint64_t or_basepowmodlg(uint64_t mbase, uint64_t exp, uint32_t base, uint64_t p, uint64_t q) {
  uint64_t curBit = shr64_cl(0x8000000000000000ull, clz64(exp) + 1);
  uint64_t a = mbase;
  while (curBit) {
    a = or_mmul(a, a, p, q);
    if (exp & curBit) a = mul_base(a, mbase, base, p, q);
    curBit >>= 1;
  }
  return a;
}

/* common.cl:93-107 */
uint64_t or_powmodsm(uint64_t mbase, uint32_t exp, uint64_t p, uint64_t q, uint64_t one) {
  if (!exp) return one;
  if (exp == 1) return mbase;
  uint32_t curBit = shr32_cl(0x80000000u, clz32(exp) + 1);
  uint64_t a = mbase;
  while (curBit) {
    a = or_mmul(a, a, p, q);
    if (exp & curBit) a = or_mmul(a, mbase, p, q);
    curBit >>= 1;
  }
  return a;
}

/* common.cl:110-135 */
uint64_t or_basepowmodsm(uint64_t mbase, uint32_t exp, uint32_t base, uint64_t p, uint64_t q, uint64_t one) {
  if (!exp) return one;
  if (exp == 1) return mbase;
  uint32_t curBit = shr32_cl(0x80000000u, clz32(exp) + 1);
  uint64_t a = mbase;
  while (curBit) {
    a = or_mmul(a, a, p, q);
    if (exp & curBit) a = mul_base(a, mbase, base, p, q);
    curBit >>= 1;
  }
  return a;
}

/* ------------------------------------------------------------------------ */
/* Segmented 2-PRP generation (getsegprimes.cl, addsmallprimes.cl,          */
/* cl_sieve.cpp:453-512 findWheelOffset)                                     */
/* ------------------------------------------------------------------------ */

/* getsegprimes.cl:37-66: base-2 strong probable prime test in Montgomery
 * form.  exp = N >> t (equal to (N-1) >> t for odd N). */
int or_strong_prp_two(uint64_t N) {
  uint64_t q = or_invert(N);
  uint64_t one = (0 - N) % N;
  uint64_t nmo = N - one;
  uint64_t two = or_add(one, one, N);
  int t = __builtin_ctzll(N - 1);
  uint64_t exp = N >> t;
  uint64_t curBit = shr64_cl(0x8000000000000000ull, clz64(exp) + 1);
  uint64_t a = two;
  while (curBit) {
    a = or_mmul(a, a, N, q);
    if (exp & curBit) a = or_add(a, a, N);
    curBit >>= 1;
  }
  if (a == one || a == nmo) return 1;
  for (int s = 1; s < t; ++s) {
    a = or_mmul(a, a, N, q);
    if (a == nmo) return 1;
  }
  return 0;
}

/* cl_sieve.cpp:453-512: first N >= start coprime to 30 and its index in the
 * mod-30 wheel {4,2,4,2,4,6,2,6}. */
void or_find_wheel_offset(uint64_t *start, int32_t *index) {
  static const int32_t wheel[8] = {4, 2, 4, 2, 4, 6, 2, 6};
  int32_t idx = -1;
  uint64_t k = *start / 6;
  int32_t i = 1;
  uint64_t N = (k * 6) - 1;
  while (N < *start || N % 5 == 0) {
    if (i) { i = 0; N += 2; }
    else { i = 1; N += 4; }
  }
  *start = N;
  while (idx < 0) {
    if (i) {
      N += 2; i = 0;
      if (N % 5 == 0) { N -= 2; idx = 5; }
    } else {
      N += 4; i = 1;
      if (N % 5 == 0) { N -= 4; idx = 7; }
    }
  }
  while (N != *start) {
    --idx;
    if (idx < 0) idx = 7;
    N -= wheel[idx];
  }
  *index = idx;
}

/* The 27 sieving primes of getsegprimes.cl:74-126 (7..113).  Their bit
 * tables are the closed form "bit i of table_p[r] is set iff p divides
 * r + 2i" (i < 32), computed here rather than copied. */
static const uint32_t kSmallPrimes[27] = {7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59,
                                          61, 67, 71, 73, 79, 83, 89, 97, 101, 103, 107, 109, 113};
static uint32_t g_bits[27][113];
static pthread_once_t g_bits_once = PTHREAD_ONCE_INIT;
static void init_bits(void) {
  for (int s = 0; s < 27; s++) {
    uint32_t p = kSmallPrimes[s];
    for (uint32_t r = 0; r < p; r++) {
      uint32_t w = 0;
      for (uint32_t i = 0; i < 32; i++)
        if ((r + 2 * i) % p == 0) w |= 1u << i;
      g_bits[s][r] = w;
    }
  }
}
uint32_t or_bitsieve_word(uint32_t prime_index, uint32_t residue) {
  pthread_once(&g_bits_once, init_bits);
  return g_bits[prime_index][residue];
}

/* getsegprimes.cl:128-182: emits every 2-PRP in [low, high) found by the
 * per-work-item walk (work-item gid covers [low + 60 gid, low + 60 gid + 60)),
 * with CKOVERFLOW wrap guards.  'low' must come from or_find_wheel_offset.
 * Returns the number of primes found; writes at most cap of them. */
size_t or_getsegprimes(uint64_t low, uint64_t high, int32_t wheelidx, uint64_t *out, size_t cap) {
  static const int wheel3[24] = {2, 1, 2, 1, 2, 3, 1, 3, 2, 1, 2, 1, 2, 3, 1, 3, 2, 1, 2, 1, 2, 3, 1, 3};
  pthread_once(&g_bits_once, init_bits);
  size_t count = 0;
  if (low >= high) return 0;
  uint64_t nthreads = (high - low) / 60 + 1;
  for (uint64_t gid = 0; gid < nthreads; gid++) {
    uint64_t P = low + gid * 60;
    if (P < low) P = high;
    uint64_t end = P + 60;
    if (end > high) end = high;
    if (end < P) end = high;
    if (P >= end) continue;
    uint32_t bitsieve = 0;
    for (int s = 0; s < 27; s++) bitsieve |= g_bits[s][P % kSmallPrimes[s]];
    int idx = wheelidx;
    while (P < end) {
      if ((bitsieve & 1) == 0) {
        if (or_strong_prp_two(P)) {
          if (count < cap) out[count] = P;
          count++;
        }
      }
      int inc = wheel3[idx++];
      uint64_t nP = P + (uint64_t)inc * 2;
      bitsieve >>= inc;
      if (nP < low) nP = end;
      P = nP;
    }
  }
  return count;
}

/* addsmallprimes.cl:8-27: primes <= 113 in [low, high) */
size_t or_addsmallprimes(uint64_t low, uint64_t high, uint64_t *out, size_t cap) {
  static const uint32_t primes[30] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47,
                                      53, 59, 61, 67, 71, 73, 79, 83, 89, 97, 101, 103, 107, 109, 113};
  size_t count = 0;
  for (int i = 0; i < 30; i++) {
    uint64_t p = primes[i];
    if (p < low || p >= high) continue;
    if (count < cap) out[count] = p;
    count++;
  }
  return count;
}

/* cl_sieve.cpp:1291-1320: the primes one sieve pass over [pmin, pmax)
 * generates, segment by segment (segment width 'range').  Returns the count. */
size_t or_primes_in_range(uint64_t pmin, uint64_t pmax, uint64_t range, uint64_t *out, size_t cap) {
  size_t count = 0;
  uint64_t p = pmin;
  if (range == 0) range = 1;
  while (p < pmax) {
    uint64_t stop = p + range;
    if (stop > pmax || stop < p) stop = pmax;
    if (p < 114) {
      uint64_t stop_sm = (stop > 114) ? 114 : stop;
      size_t c = or_addsmallprimes(p, stop_sm, out ? out + (count < cap ? count : cap) : NULL,
                                   count < cap ? cap - count : 0);
      count += c;
      p = stop_sm;
    }
    uint64_t ks = p;
    int32_t widx;
    or_find_wheel_offset(&ks, &widx);
    size_t c = or_getsegprimes(ks, stop, widx, out ? out + (count < cap ? count : cap) : NULL,
                               count < cap ? cap - count : 0);
    count += c;
    p = stop;
  }
  return count;
}

/* ------------------------------------------------------------------------ */
/* Host sizing (cl_sieve.cpp:693-735 profileGPU, 885-938 build_giant_kernels) */
/* ------------------------------------------------------------------------ */

typedef struct {
  uint32_t Q, hsize, M, QQ, MM, cache_k;
} or_bsgs_params;

/* cl_sieve.cpp:893-938 (non-NVIDIA device: no 4-byte per-array reserve).
 * Returns 0, or -1 when the device LDS is too small. */
int or_bsgs_params_for(uint32_t nmin, uint32_t nmax, int32_t kcount, uint32_t lmemsize, or_bsgs_params *out) {
  uint32_t L = nmax - nmin + 1;
  uint32_t targetQ = (uint32_t)ceil(sqrt((double)L));
  uint32_t Q = 1024, hsize = 2048;
  while (Q < 4096 && Q < targetQ) { Q <<= 1; hsize <<= 1; }
  uint32_t lmem = hsize * 4 + 16;
  while (lmem > lmemsize) {
    Q >>= 1; hsize >>= 1;
    if (Q < 512) return -1;
    lmem = hsize * 4 + 16;
  }
  out->cache_k = (lmem + (uint32_t)kcount * 8u <= lmemsize) ? 1 : 0;
  out->Q = Q;
  out->hsize = hsize;
  out->M = (uint32_t)ceil((double)L / Q);
  out->QQ = Q << 1;
  out->MM = (uint32_t)ceil((double)L / out->QQ);
  return 0;
}

/* cl_sieve.cpp:693-735: p-range per segment so that about 256*CU primes
 * fall in it; psize = 1.5 * 256 * CU. */
void or_profile_range(uint64_t start, uint32_t computeunits, uint64_t *range, uint32_t *psize) {
  uint64_t range_primes = (uint64_t)computeunits * 256;
  double C = range_primes + start / log(start);
  double x = C * (log(C) + log(log(C)));
  for (int i = 0; i < 10; i++) {
    double fx = x / log(x) - C;
    double dfx = (log(x) - 1) / (log(x) * log(x));
    x = x - fx / dfx;
  }
  uint64_t stop = (uint64_t)x;
  uint64_t calc = stop - start;
  if (calc > 4294900000ull) calc = 4294900000ull;
  *range = calc;
  *psize = (uint32_t)(1.5 * (double)range_primes);
}

/* ------------------------------------------------------------------------ */
/* Jobs, setup (setup.cl), BSGS (sort.cl + giant.cl), claims                 */
/* ------------------------------------------------------------------------ */

typedef struct {
  uint32_t base;
  int32_t kcount;
  uint32_t k[OR_MAX_SEQ]; /* |k| as read from the ABCD header */
  int32_t c[OR_MAX_SEQ];  /* +1 or -1 */
  uint32_t nmin, nmax;    /* bitmap NMIN (even) and NMAX (inputfile.cpp:405-411) */
  uint64_t words_per_seq;
  int32_t k31_compat;     /* 1: reproduce the signed-int klist encoding */
  uint32_t lmemsize;      /* device local memory for build_giant_kernels */
} or_job;

typedef struct {
  uint64_t p;
  uint32_t n;
  uint32_t kidx;
} or_factor;

typedef struct {
  uint64_t primecount;  /* 2-PRPs generated (setup.cl:99-101) */
  uint64_t claims;      /* successful bitmap claims (factors found) */
  uint64_t dups;        /* failed claims (giant.cl:163, 187, 195-196) */
  uint64_t skipped;     /* primes dropped by setup (setup.cl:156-162, 206-208) */
  uint64_t n_full, n_even, n_odd; /* prime lists with work (sort.cl:40-71) */
  uint64_t order_lists; /* lists solved by the order shortcut (giant.cl:143-168) */
  uint64_t mmul_work;   /* Montgomery products of the BSGS work model */
} or_counts;

/* Signed klist entry as the reference stores it (inputfile.cpp:366): the
 * conversions wrap modulo 2^32 like the reference's int arithmetic. */
static int32_t klist_entry(const or_job *job, int i) {
  uint32_t k = job->k[i];
  if (job->c[i] > 0) return (int32_t)k;
  return (int32_t)(0u - k);
}

/* |k| and c the kernels use for sequence i.  k31_compat follows setup.cl:143
 * and 188 on the signed entry (sign extension included); otherwise k and c
 * are taken as read. */
static void job_kc(const or_job *job, int i, uint64_t *ak, int *cpos) {
  if (job->k31_compat) {
    int32_t e = klist_entry(job, i);
    /* OpenCL: ulong aki = (klist[i] < 0) ? -klist[i] : klist[i];  the negation
     * is int arithmetic (INT_MIN stays INT_MIN), then sign-extended. */
    int32_t a = (e < 0) ? (int32_t)(0u - (uint32_t)e) : e;
    *ak = (uint64_t)(int64_t)a;
    *cpos = e > 0;
  } else {
    *ak = job->k[i];
    *cpos = job->c[i] > 0;
  }
}

typedef struct {
  uint64_t hadj;
  int32_t parity; /* 1 full, 2 even, 3 odd */
  int32_t kidx;
} or_kdata;

typedef struct {
  uint64_t p, q, one, mbase, nmo, gQ_inv;
} or_prime;

/* setup.cl:11-17 */
static int good_pr(const or_prime *pr, uint32_t base, uint64_t exponent) {
  return or_basepowmodlg(pr->mbase, exponent, base, pr->p, pr->q) == pr->one;
}

/* setup.cl:22-56 */
static int setup_pr(const or_prime *pr, uint32_t base, uint64_t *expo, int *resg) {
  const uint64_t pm = pr->p - 1;
  static const uint32_t ds[12] = {3, 4, 5, 7, 8, 9, 11, 13, 17, 19, 23, 29};
  int r = 0;
  for (int t = 0; t < 12; t++) {
    uint32_t d = ds[t];
    expo[r] = pm / d;
    if (pm == expo[r] * d && good_pr(pr, base, expo[r])) r++;
  }
  expo[12] = pm >> 1;
  uint64_t rg = or_basepowmodlg(pr->mbase, expo[12], base, pr->p, pr->q);
  *resg = (rg == pr->one) - (rg == pr->nmo);
  return r;
}

/* setup.cl:59-87 */
static int prefilter(uint64_t hk_inv, const or_prime *pr, const uint64_t *power, int powcnt, int rg) {
  uint64_t ra;
  for (int i = 0; i < powcnt; ++i) {
    ra = or_powmodlg(hk_inv, power[i], pr->p, pr->q);
    if (ra != pr->one) return 0;
  }
  ra = or_powmodlg(hk_inv, power[12], pr->p, pr->q);
  if (rg == -1) {
    if (ra == pr->one) return 2;
    else if (ra == pr->nmo) return 3;
    else return 0;
  }
  if (rg == 1 && ra == pr->nmo) return 0;
  return 1;
}

/* setup.cl:89-209 for one prime.  Returns the number of kdata written
 * (in the reference's order, kidx descending), or -1 if the prime is
 * skipped (inverse does not exist, or every k is impossible). */
int or_setup_prime(const or_job *job, uint64_t p, uint32_t Q, or_prime *pr_out, or_kdata *kd) {
  const int K = job->kcount;
  const uint32_t base = job->base;
  or_prime pr;
  pr.p = p;
  pr.q = or_invert(p);
  pr.one = (0 - p) % p;
  pr.nmo = p - pr.one;
  uint64_t two = or_add(pr.one, pr.one, p);
  uint64_t four = or_add(two, two, p);
  uint64_t r2 = or_mmul(four, four, p, pr.q);
  for (int i = 0; i < 4; i++) r2 = or_mmul(r2, r2, p, pr.q); /* 4^(2^5) = 2^64 */
  if (base == 2) pr.mbase = two;
  else if (base == 3) pr.mbase = or_add(pr.one, two, p);
  else if (base == 5) pr.mbase = or_add(four, pr.one, p);
  else pr.mbase = or_mmul(base, r2, p, pr.q);

  uint64_t mk[OR_MAX_SEQ + 1], prefix[OR_MAX_SEQ + 1];
  uint64_t gQ = or_basepowmodsm(pr.mbase, Q, base, p, pr.q, pr.one);
  mk[K] = gQ;
  for (int i = 0; i < K; ++i) {
    uint64_t ak; int cpos;
    job_kc(job, i, &ak, &cpos);
    mk[i] = or_mmul(ak, r2, p, pr.q);
  }
  prefix[0] = mk[0];
  for (int i = 1; i < K + 1; ++i) prefix[i] = or_mmul(prefix[i - 1], mk[i], p, pr.q);
  uint64_t inv_total = or_powmodlg(prefix[K], p - 2, p, pr.q);
  uint64_t one = or_mmul(prefix[K], inv_total, p, pr.q);
  if (one != pr.one) return -1;
  uint64_t prev = prefix[K - 1];
  pr.gQ_inv = or_mmul(inv_total, prev, p, pr.q);
  inv_total = or_mmul(inv_total, mk[K], p, pr.q);

  int resg;
  uint64_t expo[13];
  int tests = setup_pr(&pr, base, expo, &resg);
  int count0 = 0, pos = 0;
  for (int i = K - 1; i >= 0; --i) {
    prev = (i == 0) ? pr.one : prefix[i - 1];
    uint64_t k_inv = or_mmul(inv_total, prev, p, pr.q);
    inv_total = or_mmul(inv_total, mk[i], p, pr.q);
    uint64_t ak; int cpos;
    job_kc(job, i, &ak, &cpos);
    uint64_t hk_inv = cpos ? or_mmul(pr.nmo, k_inv, p, pr.q) : k_inv;
    int parity = prefilter(hk_inv, &pr, expo, tests, resg);
    if (!parity) { count0++; continue; }
    kd[pos].hadj = hk_inv;
    kd[pos].parity = parity;
    kd[pos].kidx = i;
    pos++;
  }
  if (pr_out) *pr_out = pr;
  if (count0 == K) return -1;
  return pos;
}

/* ---- the LDS hash of giant.cl:41-75 as a map from value to baby index ---- */
typedef struct {
  uint64_t *key;
  int32_t *idx;
  uint8_t *used;
  uint32_t mask;
} vmap;

static void vmap_init(vmap *m, uint32_t entries) {
  uint32_t sz = 1;
  while (sz < 2 * entries + 2) sz <<= 1;
  m->key = (uint64_t *)malloc(sizeof(uint64_t) * sz);
  m->idx = (int32_t *)malloc(sizeof(int32_t) * sz);
  m->used = (uint8_t *)calloc(sz, 1);
  m->mask = sz - 1;
}
static void vmap_clear(vmap *m) { memset(m->used, 0, (size_t)m->mask + 1); }
static void vmap_free(vmap *m) { free(m->key); free(m->idx); free(m->used); }
static uint32_t vmix(uint64_t v) { return (uint32_t)((v * 0x9E3779B97F4A7C15ull) >> 40); }
static void vmap_put(vmap *m, uint64_t v, int32_t j) {
  /* giant.cl:44-46: a value with low 32 bits 0 leaves its slot marked empty. */
  if ((uint32_t)v == 0) return;
  uint32_t pos = vmix(v) & m->mask;
  while (m->used[pos]) pos = (pos + 1) & m->mask;
  m->used[pos] = 1; m->key[pos] = v; m->idx[pos] = j;
}
static int vmap_get(const vmap *m, uint64_t v, int32_t *j) {
  uint32_t pos = vmix(v) & m->mask;
  while (m->used[pos]) {
    if (m->key[pos] == v) { *j = m->idx[pos]; return 1; }
    pos = (pos + 1) & m->mask;
  }
  return 0;
}

typedef struct {
  const or_job *job;
  or_bsgs_params prm;
  uint32_t *bitmap;  /* shared, claimed with atomic test-and-clear */
  or_factor *fac;
  uint64_t fac_cap;
  uint64_t fac_n;    /* atomic */
  or_counts cnt;     /* per-thread partial */
} sieve_ctx;

/* giant.cl:21-39 */
static int claim(sieve_ctx *sc, uint32_t kidx, uint32_t n) {
  const or_job *job = sc->job;
  if (n < job->nmin || n > job->nmax) return 0;
  uint32_t off = n - job->nmin;
  uint64_t word = (uint64_t)kidx * job->words_per_seq + (off >> 5);
  uint32_t mask = 1u << (off & 31u);
  uint32_t old = __atomic_fetch_and(&sc->bitmap[word], ~mask, __ATOMIC_RELAXED);
  return (old & mask) != 0;
}

static void record(sieve_ctx *sc, or_counts *c, uint64_t p, uint32_t kidx, uint32_t n) {
  if (claim(sc, kidx, n)) {
    uint64_t f = __atomic_fetch_add(&sc->fac_n, 1, __ATOMIC_RELAXED);
    if (f < sc->fac_cap) {
      sc->fac[f].p = p; sc->fac[f].n = n; sc->fac[f].kidx = kidx;
    }
    c->claims++;
  } else {
    c->dups++;
  }
}

/* sort.cl + giant.cl:77-323 for one prime list (parity 1 = giantfull,
 * 2/3 = giantparity even/odd). */
static void bsgs_list(sieve_ctx *sc, or_counts *c, vmap *vm, const or_prime *pr, const or_kdata *kd, int nk, int parity) {
  const or_job *job = sc->job;
  const uint32_t base = job->base;
  const uint64_t p = pr->p, q = pr->q;
  const uint32_t Ql = (parity == 1) ? sc->prm.Q : sc->prm.QQ;
  const uint32_t Ml = (parity == 1) ? sc->prm.M : sc->prm.MM;
  const uint32_t stride = (parity == 1) ? 1 : 2;
  const uint32_t nstart = (parity == 3) ? job->nmin + 1 : job->nmin;
  /* sort.cl:284 and giant.cl:111-112, 234-236 */
  uint64_t gj_start = or_basepowmodsm(pr->mbase, job->nmin, base, p, q, pr->one);
  uint64_t gj = gj_start;
  if (parity == 3) gj = mul_base(gj, pr->mbase, base, p, q);
  uint64_t gstep = (stride == 1) ? pr->mbase : mul_base(pr->mbase, pr->mbase, base, p, q);
  const uint64_t v0 = gj;
  uint32_t order = 0;
  vmap_clear(vm);
  uint32_t j;
  for (j = 0; j < Ql; j += stride) {
    if (j && gj == v0) { order = j; break; }
    vmap_put(vm, gj, (int32_t)j);
    gj = or_mmul(gj, gstep, p, q);
  }
  c->mmul_work += j / stride;
  if (order) {
    c->order_lists++;
    for (int i = 0; i < nk; i++) {
      int32_t r;
      if (vmap_get(vm, kd[i].hadj, &r)) {
        for (uint64_t n = (uint64_t)nstart + (uint32_t)r; n <= job->nmax; n += order)
          record(sc, c, p, (uint32_t)kd[i].kidx, (uint32_t)n);
      }
    }
    return;
  }
  /* giant step G = gQ_inv (full) or gQ_inv^2 (parity), sort.cl:292-294 */
  uint64_t G = (parity == 1) ? pr->gQ_inv : or_mmul(pr->gQ_inv, pr->gQ_inv, p, q);
  uint64_t Gq = pr->one;
  for (uint32_t qq = 0; qq < Ml; qq++) {
    for (int i = 0; i < nk; i++) {
      uint64_t gamma = or_mmul(kd[i].hadj, Gq, p, q);
      int32_t r;
      if (vmap_get(vm, gamma, &r)) {
        int32_t n = (int32_t)nstart + (int32_t)(qq * Ql) + r;
        record(sc, c, p, (uint32_t)kd[i].kidx, (uint32_t)n);
      }
    }
    Gq = or_mmul(Gq, G, p, q);
  }
  c->mmul_work += (uint64_t)Ml * (uint64_t)(nk + 1);
}

static void sieve_prime(sieve_ctx *sc, or_counts *c, vmap *vm, uint64_t p) {
  or_kdata kd[OR_MAX_SEQ], lst[OR_MAX_SEQ];
  or_prime pr;
  int nk = or_setup_prime(sc->job, p, sc->prm.Q, &pr, kd);
  if (nk < 0) { c->skipped++; return; }
  for (int parity = 1; parity <= 3; parity++) {
    int m = 0;
    for (int i = 0; i < nk; i++)
      if (kd[i].parity == parity) lst[m++] = kd[i];
    if (!m) continue;
    if (parity == 1) c->n_full++;
    else if (parity == 2) c->n_even++;
    else c->n_odd++;
    bsgs_list(sc, c, vm, &pr, lst, m, parity);
  }
}

typedef struct {
  sieve_ctx *sc;
  const uint64_t *primes;
  size_t n;
  size_t next; /* atomic work index */
  pthread_mutex_t mu;
} pool_t;

static void *worker(void *arg) {
  pool_t *pool = (pool_t *)arg;
  sieve_ctx *sc = pool->sc;
  or_counts c;
  memset(&c, 0, sizeof(c));
  vmap vm;
  vmap_init(&vm, sc->prm.QQ);
  for (;;) {
    size_t i0 = __atomic_fetch_add(&pool->next, 64, __ATOMIC_RELAXED);
    if (i0 >= pool->n) break;
    size_t i1 = i0 + 64 < pool->n ? i0 + 64 : pool->n;
    for (size_t i = i0; i < i1; i++) sieve_prime(sc, &c, &vm, pool->primes[i]);
  }
  vmap_free(&vm);
  pthread_mutex_lock(&pool->mu);
  sc->cnt.claims += c.claims; sc->cnt.dups += c.dups; sc->cnt.skipped += c.skipped;
  sc->cnt.n_full += c.n_full; sc->cnt.n_even += c.n_even; sc->cnt.n_odd += c.n_odd;
  sc->cnt.order_lists += c.order_lists; sc->cnt.mmul_work += c.mmul_work;
  pthread_mutex_unlock(&pool->mu);
  return NULL;
}

/* The sieve over [pmin, pmax): prime generation, setup, BSGS and claims on
 * 'bitmap' (in/out, the layout of inputfile.cpp:15-21).  With nthreads == 1
 * primes are processed in ascending order, so the smallest p claims a term.
 * Factors beyond fac_cap are counted but not stored.  Returns the number of
 * claims, or -1 on a parameter error. */
int64_t or_sieve(const or_job *job, uint64_t pmin, uint64_t pmax, uint32_t *bitmap, or_factor *fac, uint64_t fac_cap,
                 or_counts *counts, int nthreads, uint64_t *nfac_out) {
  sieve_ctx sc;
  memset(&sc, 0, sizeof(sc));
  sc.job = job;
  if (job->kcount < 1 || job->kcount > OR_MAX_SEQ) return -1;
  if (or_bsgs_params_for(job->nmin, job->nmax, job->kcount, job->lmemsize ? job->lmemsize : 65536, &sc.prm)) return -1;
  sc.bitmap = bitmap; sc.fac = fac; sc.fac_cap = fac_cap;
  if (nthreads < 1) nthreads = 1;
  /* generate primes in bounded chunks so memory stays small */
  const uint64_t chunk = 1ull << 24;
  uint64_t *buf = NULL;
  size_t bufcap = 0;
  uint64_t p = pmin;
  while (p < pmax) {
    uint64_t stop = p + chunk;
    if (stop > pmax || stop < p) stop = pmax;
    size_t np = or_primes_in_range(p, stop, stop - p, NULL, 0);
    if (np > bufcap) { free(buf); bufcap = np + 1024; buf = (uint64_t *)malloc(bufcap * 8); }
    or_primes_in_range(p, stop, stop - p, buf, bufcap);
    sc.cnt.primecount += np;
    pool_t pool;
    pool.sc = &sc; pool.primes = buf; pool.n = np; pool.next = 0;
    pthread_mutex_init(&pool.mu, NULL);
    if (nthreads == 1) {
      or_counts c;
      memset(&c, 0, sizeof(c));
      vmap vm;
      vmap_init(&vm, sc.prm.QQ);
      for (size_t i = 0; i < np; i++) sieve_prime(&sc, &c, &vm, buf[i]);
      vmap_free(&vm);
      sc.cnt.claims += c.claims; sc.cnt.dups += c.dups; sc.cnt.skipped += c.skipped;
      sc.cnt.n_full += c.n_full; sc.cnt.n_even += c.n_even; sc.cnt.n_odd += c.n_odd;
      sc.cnt.order_lists += c.order_lists; sc.cnt.mmul_work += c.mmul_work;
    } else {
      pthread_t th[256];
      int nt = nthreads > 256 ? 256 : nthreads;
      for (int t = 0; t < nt; t++) pthread_create(&th[t], NULL, worker, &pool);
      for (int t = 0; t < nt; t++) pthread_join(th[t], NULL);
    }
    pthread_mutex_destroy(&pool.mu);
    p = stop;
  }
  free(buf);
  if (counts) *counts = sc.cnt;
  if (nfac_out) *nfac_out = sc.fac_n;
  return (int64_t)sc.cnt.claims;
}

/* ------------------------------------------------------------------------ */
/* Host factor verification (verify_factor.c:100-181)                        */
/* ------------------------------------------------------------------------ */

/* verify_factor.c:57-95 */
static int strong_prp_base(uint32_t base, uint64_t p, uint64_t q, uint64_t one, uint64_t pmo, uint64_t r2, int t,
                           uint64_t exp, uint64_t curBit) {
  uint64_t a = or_mmul(base, r2, p, q);
  const uint64_t mbase = a;
  while (curBit) {
    a = or_mmul(a, a, p, q);
    if (exp & curBit) a = or_mmul(a, mbase, p, q);
    curBit >>= 1;
  }
  if (a == one || a == pmo) return 1;
  for (int s = 1; s < t; ++s) {
    a = or_mmul(a, a, p, q);
    if (a == pmo) return 1;
  }
  return 0;
}

/* verify_factor.c:100-128: deterministic Miller-Rabin for 64-bit p with the
 * 7-base set {2, 325, 9375, 28178, 450775, 9780504, 1795265022}. */
int or_is_prime(uint64_t p) {
  static const uint32_t bases[7] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
  if (p % 2 == 0) return 0;
  uint64_t q = or_invert(p);
  uint64_t one = (0 - p) % p;
  uint64_t pmo = p - one;
  uint64_t two = or_add(one, one, p);
  uint64_t t4 = or_add(two, two, p);
  for (int i = 0; i < 5; ++i) t4 = or_mmul(t4, t4, p, q);
  uint64_t r2 = t4;
  int t = __builtin_ctzll(p - 1);
  uint64_t exp = p >> t;
  /* x86 masks the shift count like OpenCL does: exp == 1 starts at bit 63 */
  uint64_t curBit = shr64_cl(0x8000000000000000ull, __builtin_clzll(exp) + 1);
  for (int i = 0; i < 7; ++i) {
    uint32_t base = bases[i];
    if (base >= p) {
      base %= p;
      if (base == 0) continue;
    }
    if (!strong_prp_base(base, p, q, one, pmo, r2, t, exp, curBit)) return 0;
  }
  return 1;
}

/* verify_factor.c:131-181: 1 = p divides k*b^n+c, 0 = it does not,
 * -1 = p is composite. */
int or_verify_factor(uint64_t p, uint64_t k, uint32_t n, int32_t c, uint32_t base) {
  uint64_t q = or_invert(p);
  uint64_t one = (0 - p) % p;
  uint64_t two = or_add(one, one, p);
  uint64_t t4 = or_add(two, two, p);
  for (int i = 0; i < 5; ++i) t4 = or_mmul(t4, t4, p, q);
  uint64_t r2 = t4;
  uint64_t mbase = (base == 2) ? two : or_mmul(base, r2, p, q);
  if (!or_is_prime(p)) return -1;
  uint32_t exp = n;
  uint32_t curBit = shr32_cl(0x80000000u, clz32(exp) + 1);
  uint64_t a = mbase;
  while (curBit) {
    a = or_mmul(a, a, p, q);
    if (exp & curBit) a = (base == 2) ? or_add(a, a, p) : or_mmul(a, mbase, p, q);
    curBit >>= 1;
  }
  a = or_mmul(a, k, p, q);
  if (a == 1 && c == -1) return 1;
  if (a == p - 1 && c == 1) return 1;
  return 0;
}

/* ABI probe for the Python side */
uint32_t or_sizeof_counts(void) { return (uint32_t)sizeof(or_counts); }
uint32_t or_sizeof_job(void) { return (uint32_t)sizeof(or_job); }
uint32_t or_sizeof_factor(void) { return (uint32_t)sizeof(or_factor); }
