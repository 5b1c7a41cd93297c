// rs_internal.h -- device data layouts shared by the HIP kernels and the
// host side of the C-ABI library.  Not part of the public interface.
#pragma once
#include <stdint.h>

#define RS_MAXK 100          // sequences per job (reference cl_sieve.h:2)
#define RS_NCLASS 3          // k classes with work: full n range, even n, odd n

namespace rs {

// Per-job constants, resident in device memory and read with uniform loads.
struct JobDev {
  uint32_t base;
  int32_t kcount;
  uint32_t nmin, nmax;        // bitmap NMIN (even) and NMAX
  uint64_t words_per_seq;
  uint64_t ak[RS_MAXK];       // |k| fed to the Montgomery setup (setup.cl:143)
  int32_t cpos[RS_MAXK];      // 1 if the target is -k^-1 (c = +1), setup.cl:188
  // BSGS shape
  uint32_t algo;              // 0: three lists (reference shape), 1: one merged table, 2: residue-class tables
  uint32_t Q, M, QQ, MM;      // baby steps / giant steps, full and parity lists
  uint32_t Qref;              // the reference's Q for this n-range (the setup invertibility test)
  uint32_t hsize, hmask;      // LDS hash slots (power of 2)
  uint32_t threads;           // BSGS workgroup size
  uint32_t record_factors;    // 0: claims clear bitmap bits only
  uint64_t factor_cap;
  uint32_t qcap;              // filter positives queued per BSGS workgroup (0: the kernel's queue size)
  uint32_t pad;
};

// Per-prime record written by the setup kernel (128 bytes).
struct PrimeRec {
  uint64_t p, q, one, mbase;
  uint64_t bn;                // b^NMIN (Montgomery); residue-class tables: b^-NMIN
  uint64_t g[2];              // giant step: b^-Q (full) and b^-QQ (parity); residue classes: b^(S*E)
  uint64_t gT[2];             // g[c]^threads
  uint64_t bT[2];             // baby stride per round b^(stride*threads) (bT[0]; three lists: b^threads,
                              // b^(2*threads)); residue classes: bT[0] = b^-(S*threads), bT[1] = b^-S
  uint32_t nk[RS_NCLASS];     // k count per class: full, even, odd (0 = skipped)
  uint32_t flags;             // the job's algo when written
  // residue-class tables (algo 2): baby values b^-(NMIN + S*e), e < E; M giant
  // steps of b^(S*E); every target carries its n mod S as a shift
  uint32_t S, E, M, pad32;
  uint64_t pad;
};
static_assert(sizeof(PrimeRec) == 128, "PrimeRec layout");

// Per (prime, k) entry, grouped by class: full, then even, then odd.
struct KRec {
  uint64_t hadj;              // target b^n: k^-1 (c=-1) or -k^-1 (c=+1), Montgomery (merged tables: times
                              // b^-d); residue-class tables: the inverse times b^d, -c*k*b^d
  uint32_t kidx;              // sequence index
  uint32_t cls;               // three lists: 1 full, 2 even, 3 odd (the reference's parity code);
                              // merged and residue-class tables: the shift d of the target
};

// Factor record handed to the host (p, n, sequence index).
struct FactorRec {
  uint64_t p;
  uint32_t n;
  uint32_t kidx;
};

// Device counters of one job (read by the host at checkpoints).
struct Counters {
  uint64_t primecount;        // 2-PRPs generated
  uint64_t claims;            // successful bitmap claims
  uint64_t dups;              // failed claims
  uint64_t nfactors;          // factor records appended (may exceed cap)
  uint64_t skipped;           // primes dropped by setup
  uint64_t lists[RS_NCLASS];  // (prime, class) items with work
  uint64_t order_lists;       // items solved by the order shortcut
  uint64_t work_mmul;         // Montgomery products of the work model
  uint64_t work_probe;        // hash lookups
  uint64_t max_seg_primes;    // largest per-segment prime count (overflow check)
  uint64_t seg_overflow;      // segments whose primes exceeded the buffer
};

// One job of a batch, device-resident.  The segment cursor advances on the
// device (k_round_begin), so a round of kernels needs no host input and can be
// replayed from a HIP graph.
struct SegJob {
  const JobDev* job;
  PrimeRec* recs;
  KRec* krecs;
  uint64_t* primes;
  uint32_t* bitmap;
  FactorRec* fac;
  Counters* cnt;
  uint64_t p_cur, p_end;      // next segment start, end of the requested range
  uint64_t range;             // integers per segment
  uint64_t low, high;         // this round's generator range (after small primes)
  uint32_t cap;               // prime capacity of one segment
  uint32_t nprimes;           // primes written this round
  int32_t widx;               // wheel index of 'low'
  uint32_t threads;           // BSGS workgroup size of the launch (sets the strides)
  uint32_t gen_blocks;        // primegen blocks this job needs per round
  uint32_t pad;
};

#define RS_MAX_BATCH 256
// LDS shape of the BSGS kernel by workgroup size: the positive queue
// (entries), the queue filter / discrete-log hash (32-bit words) and the
// fixed LDS byte address of the Bloom filter, above the static LDS.  At 128
// and 256 threads the static LDS stays within 8 KiB, so four workgroups with
// 32 KiB filters share a CU's 160 KiB; at 512 and 1024 threads the register
// budget allows no more workgroups than a 16 KiB base does, and the larger
// queue is checked less often.
#if defined(__HIPCC__)
__host__ __device__
#endif
constexpr uint32_t bsgs_queue_cap(uint32_t threads) { return threads >= 512 ? 384u : 128u; }
#if defined(__HIPCC__)
__host__ __device__
#endif
constexpr uint32_t bsgs_hash_words(uint32_t threads) { return threads >= 512 ? 1024u : 512u; }
#if defined(__HIPCC__)
__host__ __device__
#endif
constexpr uint32_t bsgs_filter_base(uint32_t threads) { return threads >= 512 ? 16384u : 8192u; }

// Work queue of one round over all jobs of a batch.
struct BatchState {
  uint32_t* scratch;       // per-workgroup tables of the BSGS overflow recovery (global memory)
  uint64_t scratch_words;  // 32-bit words per workgroup
  uint32_t njobs;
  uint32_t work_next;
  uint32_t item_base[RS_MAX_BATCH + 1];  // prefix sums of work items per job
};

// Words of the overflow-recovery table for a baby-step list of n entries:
// the least power of two >= 2n (and >= 64), so at most half the slots fill.
#if defined(__HIPCC__)
__host__ __device__
#endif
inline uint32_t recovery_table_words(uint32_t n) {
  uint32_t w = 64;
  while (w < 2 * n) w <<= 1;
  return w;
}

}  // namespace rs
