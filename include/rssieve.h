/*
 * rssieve.h -- C ABI of the MI355X (gfx950) k*b^n+-1 BSGS sieve library
 * (librssieve.so).  Plain C types only: pointers, sizes, integers.
 *
 * The reference (mfl0p/RSSieve v1.3) has no library boundary: its host
 * (cl_sieve.cpp) drives eight OpenCL kernels through simpleCL
 * (simpleCL.h:67-141).  Each entry point below names the reference code it
 * replaces.  Conventions:
 *   - every function returns 0 on success or a negative RS_E* code, and never
 *     exits the process; rs_last_error() gives the message of the last failure
 *     on the calling thread (the reference prints and calls exit(), e.g.
 *     simpleCL.c:347, cl_sieve.cpp:550-554);
 *   - the caller owns every host buffer, the library owns device memory;
 *   - one job handle is used by one host thread at a time.
 */
#ifndef RSSIEVE_H
#define RSSIEVE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RS_MAX_SEQUENCES 100 /* cl_sieve.h:2 */

enum {
  RS_OK = 0,
  RS_E_ARG = -1,      /* invalid argument */
  RS_E_HIP = -2,      /* HIP runtime error */
  RS_E_NOMEM = -3,    /* device or host allocation failed */
  RS_E_OVERFLOW = -4, /* a device buffer overflowed (cl_sieve.cpp:550-554, 578-581) */
  RS_E_IO = -5,       /* file I/O or format error */
  RS_E_VERIFY = -6    /* a reported factor failed the exact CPU check (cl_sieve.cpp:607-611) */
};

/* BSGS algorithm of a job. */
enum {
  RS_ALGO_THREE_LISTS = 0, /* reference shape: full / even / odd k lists, one table each (giant.cl) */
  RS_ALGO_MERGED = 1,      /* one table of even exponents shared by all classes */
  RS_ALGO_CLASSES = 2      /* one table per prime of stride S: the class of n mod S each target's
                              power-residue character fixes (S | 2 * gcd(p - 1, 720720)); Q is the
                              largest table, in entries; the eliminated terms are the same */
};

typedef struct rs_ctx rs_ctx;
typedef struct rs_job rs_job;

/* Factor record (cl_sieve.h:4-8 'factor', with the sequence index in place of
 * the signed k; rs_job_sequence() maps it back to k and c). */
typedef struct {
  uint64_t p;
  uint32_t n;
  uint32_t kidx;
} rs_factor;

/* Job counters (the reference's primecount[] slots, clearn.cl:13-18 and
 * clearresult.cl:13-19, plus g_sum, setup.cl:99-101). */
typedef struct {
  uint64_t primecount;     /* 2-PRPs generated in the sieved p-range ('prime count') */
  uint64_t claims;         /* successful bitmap claims (factors found) */
  uint64_t dups;           /* failed claims ('duplicate factors') */
  uint64_t nfactors;       /* factor records written since the last drain (may exceed capacity) */
  uint64_t skipped;        /* primes dropped by setup: not invertible, or every k impossible */
  uint64_t lists[3];       /* (prime, class) work items: full, even, odd */
  uint64_t order_lists;    /* items solved by the order shortcut (ord_p(b) below the table) */
  uint64_t work_mmul;      /* Montgomery products of the BSGS work model */
  uint64_t work_probe;     /* hash-table lookups */
  uint64_t max_seg_primes; /* largest prime count of one segment */
  uint64_t seg_overflow;   /* segments whose primes exceeded the segment buffer */
} rs_counters;

/* Tunable shape of a job.  rs_params_reference() gives the reference's own
 * choice (cl_sieve.cpp:693-735 and 885-938). */
typedef struct {
  int32_t algo;            /* RS_ALGO_* */
  uint32_t Q;              /* baby steps of the full list; parity lists use QQ = 2Q (merged: the
                              exponent span 2 * entries; residue classes: the largest table's entries) */
  uint32_t hsize;          /* LDS filter words, power of 2, >= 64 and >= the baby-step entries */
  uint32_t threads;        /* BSGS workgroup size: 128, 256, 512 or 1024 */
  uint32_t seg_primes;     /* target primes per segment (reference: 256 * compute units) */
  uint32_t blocks_per_cu;  /* persistent BSGS workgroups per CU (0 = occupancy maximum) */
  int32_t k31_compat;      /* 1: encode k and c as the reference's signed int (k >= 2^31 wraps, inputfile.cpp:366) */
  int32_t record_factors;  /* 0: claims only clear bitmap bits (pre-sieving) */
  uint64_t factor_cap;     /* factor records held on the device between drains */
  uint32_t queue_cap;      /* filter positives a BSGS workgroup queues before resolving them (0: the
                              maximum: 128 entries at 128 and 256 threads, 384 at 512 and 1024; more is
                              refused); a queue that overflows is recovered exactly by redoing its
                              giant steps, so small values only cost time (they exercise that path) */
  uint32_t reserved;       /* 0 */
} rs_params;

/* --- context ---------------------------------------------------------- */

/* Replaces main.cpp:200-325 (platform/device/context/queue) for HIP device
 * 'device'.  One stream per context. */
int rs_open(int device, rs_ctx **out);
void rs_close(rs_ctx *ctx);
/* Message of the last failure on this thread ("" if none). */
const char *rs_last_error(void);
/* The HIP stream (hipStream_t) the context enqueues on, for callers that
 * order their own work with the sieve's. */
int rs_context_stream(rs_ctx *ctx, void **stream);
/* Device facts: compute units, LDS bytes per workgroup, device name. */
int rs_device_info(rs_ctx *ctx, uint32_t *compute_units, uint32_t *lds_per_block, char *name, size_t name_len);

/* --- parameters ------------------------------------------------------- */

/* The reference's BSGS split and segment sizing: Q = smallest power of two
 * >= ceil(sqrt(L)) in [1024, 4096] that fits lmemsize, HSIZE = 2Q, M =
 * ceil(L/Q), QQ = 2Q, MM = ceil(L/QQ) (cl_sieve.cpp:893-938), segment target
 * 256 * compute_units primes (cl_sieve.cpp:698), 256-thread workgroups
 * (CL_DEVICE_MAX_WORK_GROUP_SIZE of an AMD GPU under OpenCL). */
int rs_params_reference(uint32_t nmin, uint32_t nmax, int32_t kcount, uint32_t lmemsize, uint32_t compute_units,
                        rs_params *out);
/* Derived giant-step counts M and MM for a job shape. */
int rs_params_derive(const rs_params *prm, uint32_t nmin, uint32_t nmax, uint32_t *M, uint32_t *QQ, uint32_t *MM);

/* --- jobs ------------------------------------------------------------- */

/* One sieve job: kcount sequences k[i]*base^n+c[i] (c = +1 or -1) sharing
 * the candidate bitmap 'bitmap' of kcount * words_per_seq uint32 words, bit
 * (n - nmin) of sequence i at word i*words_per_seq + (n-nmin)/32
 * (inputfile.cpp:15-21).  Replaces build_giant_kernels + buffer allocation +
 * bitmap upload (cl_sieve.cpp:885-1266). */
int rs_job_create(rs_ctx *ctx, uint32_t base, int32_t kcount, const uint32_t *k, const int32_t *c, uint32_t nmin,
                  uint32_t nmax, const uint32_t *bitmap, uint64_t words_per_seq, const rs_params *prm, rs_job **out);
void rs_job_destroy(rs_job *job);
/* A second handle on the same job: it shares the job's sequences,
 * configuration, bitmap, counters and factor list, and has its own segment
 * buffers, so rs_sieve_batch can sieve several p-ranges of one job in one
 * launch ("batch split").  Destroy views before their job; a view cannot be
 * reconfigured.  No reference counterpart. */
int rs_job_view(rs_job *job, rs_job **view);
/* k, c of sequence kidx as reported in factor lines (the signed-int encoding
 * when k31_compat is set, cl_sieve.cpp:604-605). */
int rs_job_sequence(const rs_job *job, uint32_t kidx, uint32_t *k, int32_t *c);

/* Enqueue the sieve of [pmin, pmax) on the context stream: the segment loop
 * of cl_sieve.cpp:1291-1403 (prime generation, setup, BSGS) with no host
 * round trip per segment.  Returns once enqueued. */
int rs_sieve(rs_job *job, uint64_t pmin, uint64_t pmax);
/* Allocation-free variant of rs_sieve, with the same rounds and results.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  The BSGS overflow-recovery tables go to a static device array
 * of RS_NOALLOC_SCRATCH_BYTES per GPU instead of context memory, and no
 * timing events are recorded, so the call makes no hipMalloc, hipFree,
 * hipEventCreate or other allocating call (the job's own buffers come from
 * rs_job_create).  RS_E_ARG if the job's launch needs more than that array:
 * its persistent BSGS workgroups times a table of the least power of two >=
 * twice the largest baby-step list (>= 64) 32-bit words.  Every call on one
 * GPU, from any context, shares the array and leaves its contents behind:
 * calls on one context are stream-ordered, calls that overlap on one GPU
 * (several streams or contexts at once) race. */
#define RS_NOALLOC_SCRATCH_BYTES (48u << 20)
int rs_sieve_noalloc(rs_job *job, uint64_t pmin, uint64_t pmax);
/* Enqueue several jobs into one persistent BSGS launch per segment round
 * (no reference counterpart: the reference runs one job per process).  The
 * launch has one shape: every job needs the same BSGS workgroup size
 * (rs_params.threads, else RS_E_ARG); the LDS holds the largest filter of the
 * batch, and the workgroups per CU are the occupancy maximum, capped by the
 * smallest nonzero blocks_per_cu of the jobs.  A job handle may appear once
 * (RS_E_ARG otherwise): several p-ranges of one job go in as views
 * (rs_job_view). */
int rs_sieve_batch(rs_job **jobs, int njobs, const uint64_t *pmin, const uint64_t *pmax);
/* Wait for everything enqueued on the context. */
int rs_sync(rs_ctx *ctx);

/* Counters (getResults, cl_sieve.cpp:530-554); synchronises. */
int rs_counters_read(rs_job *job, rs_counters *out);
/* Copy out and clear the device factor list and its counter (cl_sieve.cpp:
 * 583-590 + clearresult.cl); synchronises.  *n receives the number of records
 * written to out (<= cap); RS_E_OVERFLOW if the device list overflowed. */
int rs_factors_drain(rs_job *job, rs_factor *out, uint64_t cap, uint64_t *n);
/* Read / replace the device bitmap (cl_sieve.cpp:1266). */
int rs_bitmap_read(rs_job *job, uint32_t *out, uint64_t nwords);
int rs_bitmap_write(rs_job *job, const uint32_t *in, uint64_t nwords);
/* Seconds the BSGS kernel ran since the last rs_job_timer_reset (HIP events
 * around every BSGS launch on the context stream), the number of launches,
 * and the stream time from the first to the last sieve call; synchronises. */
int rs_job_kernel_time(rs_job *job, double *bsgs_seconds, uint64_t *bsgs_launches, double *total_seconds);
int rs_job_timer_reset(rs_job *job);

/* --- stage entry points (parity tests) -------------------------------- */

/* Every prime the generator emits for [pmin, pmax) (getsegprimes.cl +
 * addsmallprimes.cl), unordered.  *count receives the total (<= cap copied). */
int rs_primegen(rs_ctx *ctx, uint64_t pmin, uint64_t pmax, uint64_t *out, uint64_t cap, uint64_t *count);

/* --- measurement --------------------------------------------------------- */

/* Issue rate of v_mad_u64_u32 on the context's device, in lane operations per
 * second (independent multiply-add chains on every lane of 16 workgroups per
 * compute unit, best of three timed launches): the ceiling of the integer
 * Montgomery path's roofline (a product issues 8 of them, 4 v_mul_lo_u32 and
 * 2 v_mul_hi_u32).  No counterpart in the reference. */
int rs_peak_mad_u64_u32(rs_ctx *ctx, double *lane_ops_per_s);

/* --- host-side exact arithmetic (verify_factor.c) ---------------------- */

/* 1 if p divides k*b^n+c, 0 if not, -1 if p is composite
 * (verify_factor.c:131-181).  Exact 128-bit arithmetic. */
int rs_verify_factor(uint64_t p, uint64_t k, uint32_t n, int32_t c, uint32_t b);
/* Deterministic primality for 64-bit p (verify_factor.c:100-128). */
int rs_is_prime(uint64_t p);
/* First N >= start coprime to 30 and its wheel index (cl_sieve.cpp:453-512). */
void rs_wheel_offset(uint64_t *start, int32_t *index);

#ifdef __cplusplus
}
#endif
#endif /* RSSIEVE_H */
