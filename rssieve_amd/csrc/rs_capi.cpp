// rs_capi.cpp -- host side of the C-ABI library: device context, jobs, the
// segment loop and counter / factor read-back.  Replaces the reference host's
// OpenCL plumbing (simpleCL.c) and the device half of cl_sieve.cpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rssieve.h"
#include "../../include/rssieve_mlp.h"
#include "rs_internal.h"

namespace rs {
hipError_t bsgs_blocks_per_cu(uint32_t threads, uint32_t hsize, int* out);
size_t bsgs_block_lds_bytes(uint32_t hsize, uint32_t threads);
hipError_t bsgs_check_static_lds(size_t* worst);
// dynamic + static LDS of one BSGS workgroup (the filter ends at a fixed
// address above the static LDS, checked at rs_open)
static size_t bsgs_block_lds(uint32_t hsize, uint32_t threads) { return bsgs_block_lds_bytes(hsize, threads); }
hipError_t launch_set_cursor(SegJob* sj, uint64_t pmin, uint64_t pmax, uint32_t threads, uint64_t range,
                             uint32_t gen_blocks, hipStream_t s);
hipError_t launch_round(SegJob* jobs, BatchState* bs, uint32_t njobs, uint32_t gen_blocks, uint32_t setup_blocks,
                        uint32_t threads, uint32_t hsize, uint32_t bsgs_blocks, hipEvent_t ev0, hipEvent_t ev1,
                        hipStream_t s);
hipError_t launch_primegen_round(SegJob* jobs, BatchState* bs, uint32_t gen_blocks, hipStream_t s);
hipError_t launch_round_prep(SegJob* jobs, BatchState* bs, uint32_t njobs, uint32_t gen_blocks, uint32_t setup_blocks,
                             hipStream_t s);
hipError_t launch_round_bsgs(SegJob* jobs, BatchState* bs, uint32_t threads, uint32_t hsize, uint32_t bsgs_blocks,
                             hipStream_t s);
size_t static_scratch_bytes();
hipError_t static_scratch_address(int device, uint32_t** out);
hipError_t launch_set_scratch(BatchState* bs, uint32_t* ptr, uint64_t words, hipStream_t s);
}  // namespace rs

using namespace rs;

// The message of the last failure on this thread, in a fixed buffer (setting
// it allocates nothing, so error paths of allocation-free calls stay so).
static thread_local char g_err[512];

namespace rs {
void set_error(const std::string& msg) { snprintf(g_err, sizeof(g_err), "%s", msg.c_str()); }
}  // namespace rs

static int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return fail(RS_E_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                      __FILE__, __LINE__);                                    \
  } while (0)

struct GraphKey {
  uint32_t njobs, gen_blocks, setup_blocks, threads, hsize, bsgs_blocks;
  bool operator==(const GraphKey& o) const {
    return njobs == o.njobs && gen_blocks == o.gen_blocks && setup_blocks == o.setup_blocks && threads == o.threads &&
           hsize == o.hsize && bsgs_blocks == o.bsgs_blocks;
  }
};

struct GraphExecPair {
  hipGraphExec_t exec[2];  // round preparation, BSGS
};

struct rs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t cus = 0;
  uint32_t lds_per_block = 0;
  char name[256] = {0};
  // batch staging: SegJob array and work queue shared by rs_sieve_batch calls
  SegJob* d_bjobs = nullptr;
  BatchState* d_bstate = nullptr;
  // global scratch of the BSGS overflow recovery (one table per workgroup)
  uint32_t* d_scratch = nullptr;
  size_t scratch_bytes = 0;
  // the device's static recovery scratch of rs_sieve_noalloc (its address on this device)
  uint32_t* static_scratch = nullptr;
  // captured rounds per batch shape
  std::vector<std::pair<GraphKey, GraphExecPair>> graphs;
};

struct rs_job {
  rs_ctx* ctx = nullptr;
  JobDev h{};
  rs_params prm{};
  int32_t klist[RS_MAXK] = {0};  // the reference's signed encoding (for k31_compat reports)
  uint32_t k[RS_MAXK] = {0};
  int32_t c[RS_MAXK] = {0};
  uint64_t nwords = 0;
  uint32_t seg_cap = 0;
  uint32_t bsgs_blocks = 0;
  // device buffers
  JobDev* djob = nullptr;
  uint32_t* dbitmap = nullptr;
  Counters* dcnt = nullptr;
  FactorRec* dfac = nullptr;
  SegJob* d_sj = nullptr;       // this job as a batch of one
  BatchState* d_bs = nullptr;
  SegJob h_sj{};                // pointers of d_sj (cursor fields are set on the device)
  uint64_t* dprimes = nullptr;
  PrimeRec* drecs = nullptr;
  KRec* dkrecs = nullptr;
  // timing of BSGS launches of the last rs_sieve call
  std::vector<hipEvent_t> ev;  // pairs
  size_t ev_used = 0;
  hipEvent_t ev_all[2] = {nullptr, nullptr};
  bool timed = false;
  bool timer_started = false;
  rs_job* parent = nullptr;     // set for a view: shares the parent's job constants, bitmap, counters, factors
  double acc_bsgs_s = 0;        // BSGS time of batch rounds (graph replays)
  uint64_t acc_launches = 0;
};

extern "C" const char* rs_last_error(void) { return g_err; }

extern "C" int rs_open(int device, rs_ctx** out) {
  if (!out) return fail(RS_E_ARG, "rs_open: null out");
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(RS_E_ARG, "rs_open: device %d not present (%d devices)", device, n);
  HIPCHK(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  rs_ctx* c = new rs_ctx;
  c->device = device;
  c->cus = (uint32_t)prop.multiProcessorCount;
  c->lds_per_block = (uint32_t)prop.sharedMemPerBlock;
  snprintf(c->name, sizeof(c->name), "%s", prop.gcnArchName);
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(RS_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  size_t st_lds = 0;
  e = bsgs_check_static_lds(&st_lds);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(RS_E_HIP, "rs_open: a BSGS kernel's static LDS reaches %zu bytes into its filter: %s", st_lds,
                hipGetErrorString(e));
  }
  e = static_scratch_address(device, &c->static_scratch);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(RS_E_HIP, "rs_open: address of the static scratch: %s", hipGetErrorString(e));
  }
  if (hipMalloc((void**)&c->d_bjobs, sizeof(SegJob) * RS_MAX_BATCH) != hipSuccess ||
      hipMalloc((void**)&c->d_bstate, sizeof(BatchState)) != hipSuccess) {
    (void)hipFree(c->d_bjobs);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(RS_E_NOMEM, "rs_open: batch buffers");
  }
  *out = c;
  return RS_OK;
}

extern "C" void rs_close(rs_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
  }
  for (auto& g : ctx->graphs) {
    (void)hipGraphExecDestroy(g.second.exec[0]);
    (void)hipGraphExecDestroy(g.second.exec[1]);
  }

  (void)hipFree(ctx->d_bjobs);
  (void)hipFree(ctx->d_bstate);
  (void)hipFree(ctx->d_scratch);
  delete ctx;
}

extern "C" int rs_device_info(rs_ctx* ctx, uint32_t* cus, uint32_t* lds, char* name, size_t name_len) {
  if (!ctx) return fail(RS_E_ARG, "rs_device_info: null ctx");
  if (cus) *cus = ctx->cus;
  if (lds) *lds = ctx->lds_per_block;
  if (name && name_len) snprintf(name, name_len, "%s", ctx->name);
  return RS_OK;
}

extern "C" int rs_params_reference(uint32_t nmin, uint32_t nmax, int32_t kcount, uint32_t lmemsize,
                                   uint32_t compute_units, rs_params* out) {
  if (!out || nmax < nmin || kcount < 1 || kcount > RS_MAXK) return fail(RS_E_ARG, "rs_params_reference: bad args");
  // cl_sieve.cpp:893-938 for a non-NVIDIA device
  const uint32_t L = nmax - nmin + 1;
  const uint32_t targetQ = (uint32_t)ceil(sqrt((double)L));
  uint32_t Q = 1024, hsize = 2048;
  while (Q < 4096 && Q < targetQ) {
    Q <<= 1;
    hsize <<= 1;
  }
  uint32_t lmem = hsize * 4 + 16;
  while (lmem > lmemsize) {
    Q >>= 1;
    hsize >>= 1;
    if (Q < 512) return fail(RS_E_ARG, "insufficient local memory");
    lmem = hsize * 4 + 16;
  }
  memset(out, 0, sizeof(*out));
  out->algo = RS_ALGO_THREE_LISTS;
  out->Q = Q;
  out->hsize = hsize;
  out->threads = 256;
  out->seg_primes = 256u * (compute_units ? compute_units : 1);
  out->blocks_per_cu = 0;
  out->k31_compat = 0;
  out->record_factors = 1;
  out->factor_cap = 10000000ull;  // main.cpp:167
  return RS_OK;
}

extern "C" int rs_params_derive(const rs_params* prm, uint32_t nmin, uint32_t nmax, uint32_t* M, uint32_t* QQ,
                                uint32_t* MM) {
  if (!prm || nmax < nmin || prm->Q == 0) return fail(RS_E_ARG, "rs_params_derive: bad args");
  const uint64_t L = (uint64_t)nmax - nmin + 1;
  const uint64_t qq = (uint64_t)prm->Q * 2;
  if (M) *M = (uint32_t)((L + prm->Q - 1) / prm->Q);
  if (QQ) *QQ = (uint32_t)qq;
  if (MM) *MM = (uint32_t)((L + qq - 1) / qq);
  return RS_OK;
}

static void job_free(rs_job* j) {
  if (!j) return;
  if (!j->parent) {
    (void)hipFree(j->djob);
    (void)hipFree(j->dbitmap);
    (void)hipFree(j->dcnt);
    (void)hipFree(j->dfac);
  }
  (void)hipFree(j->d_sj);
  (void)hipFree(j->d_bs);
  (void)hipFree(j->dprimes);
  (void)hipFree(j->drecs);
  (void)hipFree(j->dkrecs);
  for (auto e : j->ev) (void)hipEventDestroy(e);
  if (j->ev_all[0]) (void)hipEventDestroy(j->ev_all[0]);
  if (j->ev_all[1]) (void)hipEventDestroy(j->ev_all[1]);
  delete j;
}

extern "C" void rs_job_destroy(rs_job* job) {
  if (!job) return;
  (void)hipSetDevice(job->ctx->device);
  (void)hipStreamSynchronize(job->ctx->stream);
  job_free(job);
}

static bool pow2(uint32_t x) { return x && !(x & (x - 1)); }

// Checks of a BSGS configuration against a job shape and the device.
static int check_params(const rs_ctx* ctx, const rs_params* prm, const char* who) {
  if (prm->algo != RS_ALGO_THREE_LISTS && prm->algo != RS_ALGO_MERGED && prm->algo != RS_ALGO_CLASSES)
    return fail(RS_E_ARG, "%s: unknown algorithm %d", who, prm->algo);
  if (prm->algo == RS_ALGO_MERGED && (prm->Q & 1))
    return fail(RS_E_ARG, "%s: the merged table needs an even exponent span Q (got %u)", who, prm->Q);
  if (!(prm->threads == 128 || prm->threads == 256 || prm->threads == 512 || prm->threads == 1024))
    return fail(RS_E_ARG, "%s: threads %u not in {128,256,512,1024}", who, prm->threads);
  const uint32_t entries = prm->algo == RS_ALGO_MERGED ? prm->Q / 2 : prm->Q;  // baby steps in one table
  if (prm->algo == RS_ALGO_CLASSES && (prm->Q < 64 || prm->Q % 64))
    return fail(RS_E_ARG, "%s: residue-class tables need a multiple of 64 entries (got %u)", who, prm->Q);
  if (prm->Q < 16 || entries > 16384 || !pow2(prm->hsize) || prm->hsize < 64 || prm->hsize < entries)
    return fail(RS_E_ARG, "%s: need at most 16384 baby-step entries and a power-of-two filter of at least 64 "
                "words and at least one word per entry (Q=%u entries=%u hsize=%u)", who, prm->Q, entries, prm->hsize);
  if (prm->queue_cap > bsgs_queue_cap(prm->threads))
    return fail(RS_E_ARG, "%s: queue_cap %u above the %u entries of a %u-thread workgroup", who, prm->queue_cap,
                bsgs_queue_cap(prm->threads), prm->threads);
  if (bsgs_block_lds(prm->hsize, prm->threads) > ctx->lds_per_block)
    return fail(RS_E_ARG, "%s: hash of %u slots needs %zu LDS bytes, device has %u", who, prm->hsize,
                bsgs_block_lds(prm->hsize, prm->threads), ctx->lds_per_block);
  return RS_OK;
}

// Apply a checked configuration to the job's device constants.
static int apply_params(rs_job* j, const rs_params* prm) {
  JobDev& h = j->h;
  j->prm.algo = prm->algo;
  j->prm.Q = prm->Q;
  j->prm.hsize = prm->hsize;
  j->prm.threads = prm->threads;
  j->prm.blocks_per_cu = prm->blocks_per_cu;
  h.algo = (uint32_t)prm->algo;
  h.Q = prm->Q;
  rs_params_derive(prm, h.nmin, h.nmax, &h.M, &h.QQ, &h.MM);
  h.hsize = prm->hsize;
  h.hmask = prm->hsize - 1;
  h.threads = prm->threads;
  h.qcap = prm->queue_cap;  // 0: the kernel's queue size
  int bpc = 0;
  if (bsgs_blocks_per_cu(prm->threads, prm->hsize, &bpc) != hipSuccess || bpc < 1) bpc = 1;
  if (prm->blocks_per_cu && (int)prm->blocks_per_cu < bpc) bpc = (int)prm->blocks_per_cu;
  j->bsgs_blocks = (uint32_t)bpc * j->ctx->cus;
  return RS_OK;
}

extern "C" int rs_job_create(rs_ctx* ctx, uint32_t base, int32_t kcount, const uint32_t* k, const int32_t* c,
                             uint32_t nmin, uint32_t nmax, const uint32_t* bitmap, uint64_t words_per_seq,
                             const rs_params* prm, rs_job** out) {
  if (!ctx || !k || !c || !bitmap || !prm || !out) return fail(RS_E_ARG, "rs_job_create: null argument");
  if (kcount < 1 || kcount > RS_MAXK) return fail(RS_E_ARG, "rs_job_create: kcount %d outside [1, %d]", kcount, RS_MAXK);
  if (base < 2) return fail(RS_E_ARG, "rs_job_create: base %u < 2", base);
  if (nmax < nmin || nmax > 0x7fffffffu) return fail(RS_E_ARG, "rs_job_create: bad n range [%u, %u]", nmin, nmax);
  if (nmin & 1u) return fail(RS_E_ARG, "rs_job_create: nmin must be even (bitmap origin)");
  if (words_per_seq < (((uint64_t)nmax - nmin + 1) + 31) / 32)
    return fail(RS_E_ARG, "rs_job_create: words_per_seq too small for [%u, %u]", nmin, nmax);
  {
    const int rc = check_params(ctx, prm, "rs_job_create");
    if (rc) return rc;
  }
  for (int i = 0; i < kcount; i++) {
    if (k[i] == 0) return fail(RS_E_ARG, "rs_job_create: k[%d] = 0", i);
    if (c[i] != 1 && c[i] != -1) return fail(RS_E_ARG, "rs_job_create: c[%d] = %d, need +1 or -1", i, c[i]);
  }
  HIPCHK(hipSetDevice(ctx->device));
  rs_job* j = new rs_job;
  j->ctx = ctx;
  j->prm = *prm;
  JobDev& h = j->h;
  h.base = base;
  h.kcount = kcount;
  h.nmin = nmin;
  h.nmax = nmax;
  h.words_per_seq = words_per_seq;
  for (int i = 0; i < kcount; i++) {
    j->k[i] = k[i];
    j->c[i] = c[i];
    // inputfile.cpp:366: klist[i] = c > 0 ? (int)k : -(int)k, modulo 2^32
    const int32_t e = (c[i] > 0) ? (int32_t)k[i] : (int32_t)(0u - k[i]);
    j->klist[i] = e;
    if (prm->k31_compat) {
      // setup.cl:143 and 188 on the signed entry; the negation stays in int
      const int32_t a = (e < 0) ? (int32_t)(0u - (uint32_t)e) : e;
      h.ak[i] = (uint64_t)(int64_t)a;
      h.cpos[i] = e > 0;
    } else {
      h.ak[i] = k[i];
      h.cpos[i] = c[i] > 0;
    }
  }
  {
    rs_params ref;
    if (rs_params_reference(nmin, nmax, kcount, 65536, ctx->cus, &ref) != RS_OK) {
      job_free(j);
      return RS_E_ARG;
    }
    h.Qref = ref.Q;
  }
  h.record_factors = prm->record_factors ? 1 : 0;
  h.factor_cap = prm->factor_cap ? prm->factor_cap : 1;
  apply_params(j, prm);
  j->nwords = (uint64_t)kcount * words_per_seq;
  j->seg_cap = std::max<uint32_t>(4096, 2 * (prm->seg_primes ? prm->seg_primes : 65536));

#define ALLOC(ptr, bytes)                                                                        \
  do {                                                                                           \
    if (hipMalloc((void**)&(ptr), (bytes)) != hipSuccess) {                                      \
      job_free(j);                                                                               \
      return fail(RS_E_NOMEM, "hipMalloc of %zu bytes for %s failed", (size_t)(bytes), #ptr);    \
    }                                                                                            \
  } while (0)
  ALLOC(j->djob, sizeof(JobDev));
  ALLOC(j->dbitmap, j->nwords * 4);
  ALLOC(j->dcnt, sizeof(Counters));
  ALLOC(j->dfac, h.factor_cap * sizeof(FactorRec));
  ALLOC(j->d_sj, sizeof(SegJob));
  ALLOC(j->d_bs, sizeof(BatchState));
  ALLOC(j->dprimes, (size_t)j->seg_cap * 8);
  ALLOC(j->drecs, (size_t)j->seg_cap * sizeof(PrimeRec));
  ALLOC(j->dkrecs, (size_t)j->seg_cap * kcount * 2 * sizeof(KRec));  // room for the merged table's targets
#undef ALLOC
  SegJob& sj = j->h_sj;
  memset(&sj, 0, sizeof(sj));
  sj.job = j->djob;
  sj.recs = j->drecs;
  sj.krecs = j->dkrecs;
  sj.primes = j->dprimes;
  sj.bitmap = j->dbitmap;
  sj.fac = j->dfac;
  sj.cnt = j->dcnt;
  sj.cap = j->seg_cap;
  sj.threads = prm->threads;
  hipStream_t s = ctx->stream;
  if (hipMemcpyAsync(j->djob, &h, sizeof(JobDev), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(j->dbitmap, bitmap, j->nwords * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemsetAsync(j->dcnt, 0, sizeof(Counters), s) != hipSuccess ||
      hipMemcpyAsync(j->d_sj, &sj, sizeof(SegJob), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemsetAsync(j->d_bs, 0, sizeof(BatchState), s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
    job_free(j);
    return fail(RS_E_HIP, "rs_job_create: upload failed");
  }
  if (hipEventCreate(&j->ev_all[0]) != hipSuccess || hipEventCreate(&j->ev_all[1]) != hipSuccess) {
    job_free(j);
    return fail(RS_E_HIP, "rs_job_create: event creation failed");
  }
  *out = j;
  return RS_OK;
}

extern "C" int rs_job_sequence(const rs_job* job, uint32_t kidx, uint32_t* k, int32_t* c) {
  if (!job || kidx >= (uint32_t)job->h.kcount) return fail(RS_E_ARG, "rs_job_sequence: bad index");
  if (job->prm.k31_compat) {
    // cl_sieve.cpp:603-605 on the signed entry
    const int32_t e = job->klist[kidx];
    if (k) *k = (e < 0) ? (uint32_t)(0u - (uint32_t)e) : (uint32_t)e;
    if (c) *c = (e < 0) ? -1 : 1;
  } else {
    if (k) *k = job->k[kidx];
    if (c) *c = job->c[kidx];
  }
  return RS_OK;
}

// cl_sieve.cpp:693-735: p-range of a segment holding about 'target' primes
static uint64_t segment_range(uint64_t start, uint32_t target) {
  double s = (double)(start < 17 ? 17 : start);
  double C = (double)target + s / log(s);
  double x = C * (log(C) + log(log(C)));
  for (int i = 0; i < 10; i++) {
    double fx = x / log(x) - C;
    double dfx = (log(x) - 1) / (log(x) * log(x));
    x = x - fx / dfx;
  }
  double r = x - s;
  if (!(r > 1.0)) r = 1.0;
  if (r > 4294900000.0) r = 4294900000.0;
  // at small p the estimate is loose: never exceed what the buffer can hold
  // even if every integer coprime to 30 were prime
  return (uint64_t)r;
}

// integers per segment and the blocks the generator needs for them
static void segment_shape(const rs_job* j, uint64_t pmin, uint64_t* range, uint32_t* gen_blocks) {
  const uint32_t target = j->prm.seg_primes ? j->prm.seg_primes : 65536;
  uint64_t r = segment_range(pmin, target);
  // below 2^20 the estimate is loose: bound the segment by what the buffer
  // holds even if every integer coprime to 30 were prime (8 per 30, + 16)
  if (pmin < (1ull << 20)) r = std::min<uint64_t>(r, (uint64_t)(j->seg_cap - 64) * 30 / 8);
  if (r < 60) r = 60;
  *range = r;
  *gen_blocks = (uint32_t)(((r / 60 + 1) + 255) / 256);
}

static uint64_t rounds_for(uint64_t pmin, uint64_t pmax, uint64_t range) {
  if (pmax <= pmin) return 0;
  return (pmax - pmin + range - 1) / range;
}

static int record_event(rs_job* j, hipEvent_t* out) {
  if (j->ev_used + 1 > j->ev.size()) {
    for (int i = 0; i < 64; i++) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      j->ev.push_back(e);
    }
  }
  *out = j->ev[j->ev_used++];
  return RS_OK;
}

// Baby-step entries of the largest list of a job's BSGS shape.
static uint32_t job_max_entries(const JobDev& h) {
  return h.algo == 1 ? h.Q / 2 : h.algo == 2 ? h.Q : std::max(h.Q, h.QQ / 2);
}

// Size the overflow-recovery scratch for 'blocks' workgroups and lists of up
// to 'entries' entries, and point the launch's BatchState at it (stream
// ordered, before the launches that read it).  Growing it waits for the
// stream, so no enqueued kernel still uses the old buffer.
static int ensure_scratch(rs_ctx* ctx, BatchState* d_bs, uint32_t blocks, uint32_t entries) {
  const uint64_t words = recovery_table_words(entries);
  const size_t need = (size_t)blocks * words * 4;
  if (need > ctx->scratch_bytes) {
    HIPCHK(hipStreamSynchronize(ctx->stream));
    (void)hipFree(ctx->d_scratch);
    ctx->d_scratch = nullptr;
    ctx->scratch_bytes = 0;
    if (hipMalloc((void**)&ctx->d_scratch, need) != hipSuccess)
      return fail(RS_E_NOMEM, "BSGS recovery scratch of %zu bytes", need);
    ctx->scratch_bytes = need;
  }
  struct {
    uint32_t* p;
    uint64_t w;
  } hdr{ctx->d_scratch, words};
  static_assert(sizeof(hdr) == offsetof(BatchState, njobs), "BatchState header");
  HIPCHK(hipMemcpyAsync(d_bs, &hdr, sizeof(hdr), hipMemcpyHostToDevice, ctx->stream));
  return RS_OK;
}

extern "C" int rs_sieve(rs_job* j, uint64_t pmin, uint64_t pmax) {
  if (!j) return fail(RS_E_ARG, "rs_sieve: null job");
  if (pmin < 3) return fail(RS_E_ARG, "rs_sieve: pmin %llu < 3", (unsigned long long)pmin);
  if (pmax < pmin) return fail(RS_E_ARG, "rs_sieve: pmax < pmin");
  HIPCHK(hipSetDevice(j->ctx->device));
  hipStream_t s = j->ctx->stream;
  uint64_t range;
  uint32_t gen_blocks;
  segment_shape(j, pmin, &range, &gen_blocks);
  HIPCHK(launch_set_cursor(j->d_sj, pmin, pmax, j->h.threads, range, gen_blocks, s));
  {
    const int rc = ensure_scratch(j->ctx, j->d_bs, j->bsgs_blocks, job_max_entries(j->h));
    if (rc) return rc;
  }
  if (!j->timer_started) {
    HIPCHK(hipEventRecord(j->ev_all[0], s));
    j->timer_started = true;
  }
  const uint64_t R = rounds_for(pmin, pmax, range);
  const uint32_t setup_blocks = (j->seg_cap + 3) / 4;
  for (uint64_t r = 0; r < R; r++) {
    hipEvent_t e0, e1;
    int rc = record_event(j, &e0);
    if (!rc) rc = record_event(j, &e1);
    if (rc) return rc;
    HIPCHK(launch_round(j->d_sj, j->d_bs, 1, gen_blocks, setup_blocks, j->h.threads, j->h.hsize, j->bsgs_blocks, e0,
                        e1, s));
  }
  HIPCHK(hipEventRecord(j->ev_all[1], s));
  return RS_OK;
}

extern "C" int rs_sieve_noalloc(rs_job* j, uint64_t pmin, uint64_t pmax) {
  if (!j) return fail(RS_E_ARG, "rs_sieve_noalloc: null job");
  if (pmin < 3) return fail(RS_E_ARG, "rs_sieve_noalloc: pmin %llu < 3", (unsigned long long)pmin);
  if (pmax < pmin) return fail(RS_E_ARG, "rs_sieve_noalloc: pmax < pmin");
  rs_ctx* ctx = j->ctx;
  const uint64_t words = recovery_table_words(job_max_entries(j->h));
  const size_t need = (size_t)j->bsgs_blocks * words * 4;
  if (need > static_scratch_bytes())
    return fail(RS_E_ARG, "rs_sieve_noalloc: %u BSGS workgroups with recovery tables of %llu words need %zu bytes, "
                "above the %zu-byte static scratch (use rs_sieve, or a smaller table or fewer blocks per CU)",
                j->bsgs_blocks, (unsigned long long)words, need, static_scratch_bytes());
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  uint64_t range;
  uint32_t gen_blocks;
  segment_shape(j, pmin, &range, &gen_blocks);
  HIPCHK(launch_set_cursor(j->d_sj, pmin, pmax, j->h.threads, range, gen_blocks, s));
  HIPCHK(launch_set_scratch(j->d_bs, ctx->static_scratch, words, s));
  const uint64_t R = rounds_for(pmin, pmax, range);
  const uint32_t setup_blocks = (j->seg_cap + 3) / 4;
  for (uint64_t r = 0; r < R; r++)
    HIPCHK(launch_round(j->d_sj, j->d_bs, 1, gen_blocks, setup_blocks, j->h.threads, j->h.hsize, j->bsgs_blocks,
                        nullptr, nullptr, s));
  return RS_OK;
}

extern "C" int rs_sieve_batch(rs_job** jobs, int njobs, const uint64_t* pmin, const uint64_t* pmax) {
  if (!jobs || njobs < 1 || njobs > RS_MAX_BATCH || !pmin || !pmax)
    return fail(RS_E_ARG, "rs_sieve_batch: need 1..%d jobs", RS_MAX_BATCH);
  rs_ctx* ctx = jobs[0] ? jobs[0]->ctx : nullptr;
  if (!ctx) return fail(RS_E_ARG, "rs_sieve_batch: null job");
  uint32_t T = jobs[0]->h.threads, hsize = 0, gen_max = 0, cap_max = 0;
  uint64_t R = 0;
  std::vector<uint64_t> ranges(njobs);
  std::vector<uint32_t> gbs(njobs);
  uint32_t bpc_cap = 0;  // the smallest blocks_per_cu a job asks for (0: none)
  for (int i = 0; i < njobs; i++) {
    rs_job* j = jobs[i];
    if (!j || j->ctx != ctx) return fail(RS_E_ARG, "rs_sieve_batch: jobs must share one context");
    if (pmin[i] < 3 || pmax[i] < pmin[i]) return fail(RS_E_ARG, "rs_sieve_batch: bad range of job %d", i);
    if (j->h.threads != T)
      return fail(RS_E_ARG, "rs_sieve_batch: job %d has %u-thread BSGS workgroups, job 0 has %u (one launch shape "
                  "per batch)", i, j->h.threads, T);
    for (int k = 0; k < i; k++)
      if (jobs[k] == j)
        return fail(RS_E_ARG, "rs_sieve_batch: job %d appears twice (a job has one segment buffer; sieve more "
                    "ranges of it through rs_job_view)", i);
    if (j->prm.blocks_per_cu && (!bpc_cap || j->prm.blocks_per_cu < bpc_cap)) bpc_cap = j->prm.blocks_per_cu;
    uint64_t range;
    uint32_t gb;
    segment_shape(j, pmin[i], &range, &gb);
    ranges[i] = range;
    gbs[i] = gb;
    hsize = std::max(hsize, j->h.hsize);
    gen_max = std::max(gen_max, gb);
    cap_max = std::max(cap_max, j->seg_cap);
    R = std::max(R, rounds_for(pmin[i], pmax[i], range));
  }
  if (bsgs_block_lds(hsize, T) > ctx->lds_per_block) return fail(RS_E_ARG, "rs_sieve_batch: table too large");
  int bpc = 0;
  if (bsgs_blocks_per_cu(T, hsize, &bpc) != hipSuccess || bpc < 1) bpc = 1;
  if (bpc_cap && (int)bpc_cap < bpc) bpc = (int)bpc_cap;
  const uint32_t bsgs_blocks = (uint32_t)bpc * ctx->cus;
  const uint32_t setup_blocks = (cap_max + 3) / 4;
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  {
    uint32_t emax = 0;
    for (int i = 0; i < njobs; i++) emax = std::max(emax, job_max_entries(jobs[i]->h));
    const int rc = ensure_scratch(ctx, ctx->d_bstate, bsgs_blocks, emax);
    if (rc) return rc;
  }
  // the pointers of every job (device to device), then its cursor and shape
  for (int i = 0; i < njobs; i++) {
    HIPCHK(hipMemcpyAsync(ctx->d_bjobs + i, jobs[i]->d_sj, sizeof(SegJob), hipMemcpyDeviceToDevice, s));
    HIPCHK(launch_set_cursor(ctx->d_bjobs + i, pmin[i], pmax[i], T, ranges[i], gbs[i], s));
  }
  // one captured round per batch shape, replayed R times
  const GraphKey key{(uint32_t)njobs, gen_max, setup_blocks, T, hsize, bsgs_blocks};
  // The round as two captured graphs, preparation and BSGS, with stream
  // events between them so every replay's BSGS launch is timed.
  GraphExecPair* gp = nullptr;
  for (auto& g : ctx->graphs)
    if (g.first == key) gp = &g.second;
  if (!gp) {
    GraphExecPair pair{};
    for (int c = 0; c < 2; c++) {
      hipGraph_t graph;
      HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      hipError_t e = c == 0 ? launch_round_prep(ctx->d_bjobs, ctx->d_bstate, (uint32_t)njobs, gen_max, setup_blocks, s)
                            : launch_round_bsgs(ctx->d_bjobs, ctx->d_bstate, T, hsize, bsgs_blocks, s);
      hipError_t e2 = hipStreamEndCapture(s, &graph);
      if (e != hipSuccess || e2 != hipSuccess)
        return fail(RS_E_HIP, "rs_sieve_batch: capture failed: %s", hipGetErrorString(e ? e : e2));
      e = hipGraphInstantiate(&pair.exec[c], graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) return fail(RS_E_HIP, "rs_sieve_batch: instantiate failed: %s", hipGetErrorString(e));
    }
    ctx->graphs.push_back({key, pair});
    gp = &ctx->graphs.back().second;
  }
  rs_job* j0 = jobs[0];
  if (!j0->timer_started) {
    HIPCHK(hipEventRecord(j0->ev_all[0], s));
    j0->timer_started = true;
  }
  for (uint64_t r = 0; r < R; r++) {
    hipEvent_t e0, e1;
    int rc = record_event(j0, &e0);
    if (!rc) rc = record_event(j0, &e1);
    if (rc) return rc;
    HIPCHK(hipGraphLaunch(gp->exec[0], s));
    HIPCHK(hipEventRecord(e0, s));
    HIPCHK(hipGraphLaunch(gp->exec[1], s));
    HIPCHK(hipEventRecord(e1, s));
  }
  HIPCHK(hipEventRecord(j0->ev_all[1], s));
  return RS_OK;
}

extern "C" int rs_sync(rs_ctx* ctx) {
  if (!ctx) return fail(RS_E_ARG, "rs_sync: null ctx");
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return RS_OK;
}

extern "C" int rs_job_kernel_time(rs_job* j, double* bsgs_s, uint64_t* launches, double* total_s) {
  if (!j) return fail(RS_E_ARG, "rs_job_kernel_time: null job");
  HIPCHK(hipStreamSynchronize(j->ctx->stream));
  double t = 0;
  for (size_t i = 0; i + 1 < j->ev_used; i += 2) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, j->ev[i], j->ev[i + 1]));
    t += ms * 1e-3;
  }
  if (bsgs_s) *bsgs_s = t + j->acc_bsgs_s;
  if (launches) *launches = j->ev_used / 2 + j->acc_launches;
  if (total_s) {
    float ms = 0;
    if (j->timer_started) HIPCHK(hipEventElapsedTime(&ms, j->ev_all[0], j->ev_all[1]));
    *total_s = ms * 1e-3;
  }
  return RS_OK;
}

extern "C" int rs_job_timer_reset(rs_job* j) {
  if (!j) return fail(RS_E_ARG, "rs_job_timer_reset: null job");
  HIPCHK(hipStreamSynchronize(j->ctx->stream));
  j->ev_used = 0;
  j->timer_started = false;
  j->acc_bsgs_s = 0;
  j->acc_launches = 0;
  return RS_OK;
}

extern "C" int rs_counters_read(rs_job* j, rs_counters* out) {
  if (!j || !out) return fail(RS_E_ARG, "rs_counters_read: null argument");
  static_assert(sizeof(rs_counters) == sizeof(Counters), "counter layout");
  HIPCHK(hipSetDevice(j->ctx->device));
  HIPCHK(hipStreamSynchronize(j->ctx->stream));
  HIPCHK(hipMemcpy(out, j->dcnt, sizeof(Counters), hipMemcpyDeviceToHost));
  return RS_OK;
}

extern "C" int rs_factors_drain(rs_job* j, rs_factor* out, uint64_t cap, uint64_t* n) {
  if (!j || !n || (cap && !out)) return fail(RS_E_ARG, "rs_factors_drain: null argument");
  static_assert(sizeof(rs_factor) == sizeof(FactorRec), "factor layout");
  HIPCHK(hipSetDevice(j->ctx->device));
  HIPCHK(hipStreamSynchronize(j->ctx->stream));
  Counters c;
  HIPCHK(hipMemcpy(&c, j->dcnt, sizeof(Counters), hipMemcpyDeviceToHost));
  const uint64_t stored = std::min<uint64_t>(c.nfactors, j->h.factor_cap);
  const uint64_t ncopy = std::min<uint64_t>(stored, cap);
  if (ncopy) HIPCHK(hipMemcpy(out, j->dfac, ncopy * sizeof(FactorRec), hipMemcpyDeviceToHost));
  *n = ncopy;
  if (c.nfactors > j->h.factor_cap)
    return fail(RS_E_OVERFLOW, "number of results (%llu) overflowed the device factor list (%llu)",
                (unsigned long long)c.nfactors, (unsigned long long)j->h.factor_cap);
  if (stored > cap)
    return fail(RS_E_OVERFLOW, "host buffer of %llu records is smaller than the %llu stored",
                (unsigned long long)cap, (unsigned long long)stored);
  uint64_t zero = 0;
  HIPCHK(hipMemcpy(&j->dcnt->nfactors, &zero, 8, hipMemcpyHostToDevice));
  return RS_OK;
}

extern "C" int rs_bitmap_read(rs_job* j, uint32_t* out, uint64_t nwords) {
  if (!j || !out || nwords != j->nwords) return fail(RS_E_ARG, "rs_bitmap_read: need %llu words",
                                                     (unsigned long long)(j ? j->nwords : 0));
  HIPCHK(hipSetDevice(j->ctx->device));
  HIPCHK(hipStreamSynchronize(j->ctx->stream));
  HIPCHK(hipMemcpy(out, j->dbitmap, nwords * 4, hipMemcpyDeviceToHost));
  return RS_OK;
}

extern "C" int rs_bitmap_write(rs_job* j, const uint32_t* in, uint64_t nwords) {
  if (!j || !in || nwords != j->nwords) return fail(RS_E_ARG, "rs_bitmap_write: need %llu words",
                                                    (unsigned long long)(j ? j->nwords : 0));
  HIPCHK(hipSetDevice(j->ctx->device));
  HIPCHK(hipStreamSynchronize(j->ctx->stream));
  HIPCHK(hipMemcpy(j->dbitmap, in, nwords * 4, hipMemcpyHostToDevice));
  return RS_OK;
}

extern "C" int rs_primegen(rs_ctx* ctx, uint64_t pmin, uint64_t pmax, uint64_t* out, uint64_t cap, uint64_t* count) {
  if (!ctx || !count || (cap && !out)) return fail(RS_E_ARG, "rs_primegen: null argument");
  if (pmin < 3 || pmax < pmin) return fail(RS_E_ARG, "rs_primegen: need 3 <= pmin <= pmax");
  HIPCHK(hipSetDevice(ctx->device));
  const uint32_t segcap = 1u << 20;
  uint64_t* dp = nullptr;
  SegJob* dj = nullptr;
  BatchState* dbs = nullptr;
  if (hipMalloc((void**)&dp, (size_t)segcap * 8) != hipSuccess || hipMalloc((void**)&dj, sizeof(SegJob)) != hipSuccess ||
      hipMalloc((void**)&dbs, sizeof(BatchState)) != hipSuccess) {
    (void)hipFree(dp);
    (void)hipFree(dj);
    return fail(RS_E_NOMEM, "rs_primegen: allocation failed");
  }
  // a segment can hold at most 8/30 of its integers (+ 16 at the edges)
  const uint64_t range = (uint64_t)(segcap - 64) * 30 / 8;
  const uint32_t gen_blocks = (uint32_t)(((range / 60 + 1) + 255) / 256);
  SegJob h;
  memset(&h, 0, sizeof(h));
  h.primes = dp;
  h.cap = segcap;
  hipStream_t s = ctx->stream;
  int rc = RS_OK;
  if (hipMemcpy(dj, &h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess ||
      launch_set_cursor(dj, pmin, pmax, 256, range, gen_blocks, s) != hipSuccess)
    rc = fail(RS_E_HIP, "rs_primegen: launch failed");
  uint64_t total = 0;
  std::vector<uint64_t> tmp;
  const uint64_t R = rounds_for(pmin, pmax, range);
  for (uint64_t r = 0; r < R && rc == RS_OK; r++) {
    SegJob hs;
    if (launch_primegen_round(dj, dbs, gen_blocks, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess ||
        hipMemcpy(&hs, dj, sizeof(hs), hipMemcpyDeviceToHost) != hipSuccess) {
      rc = fail(RS_E_HIP, "rs_primegen: round failed");
      break;
    }
    const uint64_t n = hs.nprimes;
    if (n > segcap) {
      rc = fail(RS_E_OVERFLOW, "rs_primegen: segment overflow");
      break;
    }
    tmp.resize(n);
    if (n && hipMemcpy(tmp.data(), dp, n * 8, hipMemcpyDeviceToHost) != hipSuccess) {
      rc = fail(RS_E_HIP, "rs_primegen: copy failed");
      break;
    }
    for (uint64_t i = 0; i < n; i++) {
      if (total < cap) out[total] = tmp[i];
      total++;
    }
  }
  (void)hipFree(dp);
  (void)hipFree(dj);
  (void)hipFree(dbs);
  *count = total;
  return rc;
}

extern "C" int rs_context_stream(rs_ctx* ctx, void** stream) {
  if (!ctx || !stream) return fail(RS_E_ARG, "rs_context_stream: null");
  *stream = (void*)ctx->stream;
  return RS_OK;
}

extern "C" int rs_job_view(rs_job* parent, rs_job** out) {
  if (!parent || !out) return fail(RS_E_ARG, "rs_job_view: null");
  if (parent->parent) return fail(RS_E_ARG, "rs_job_view: a view of a view");
  HIPCHK(hipSetDevice(parent->ctx->device));
  rs_job* j = new rs_job;
  j->ctx = parent->ctx;
  j->h = parent->h;
  j->prm = parent->prm;
  memcpy(j->klist, parent->klist, sizeof(j->klist));
  memcpy(j->k, parent->k, sizeof(j->k));
  memcpy(j->c, parent->c, sizeof(j->c));
  j->nwords = parent->nwords;
  j->seg_cap = parent->seg_cap;
  j->bsgs_blocks = parent->bsgs_blocks;
  j->parent = parent;
  j->djob = parent->djob;
  j->dbitmap = parent->dbitmap;
  j->dcnt = parent->dcnt;
  j->dfac = parent->dfac;
  const int K = parent->h.kcount;
  if (hipMalloc((void**)&j->dprimes, (size_t)j->seg_cap * 8) != hipSuccess ||
      hipMalloc((void**)&j->drecs, (size_t)j->seg_cap * sizeof(PrimeRec)) != hipSuccess ||
      hipMalloc((void**)&j->dkrecs, (size_t)j->seg_cap * K * 2 * sizeof(KRec)) != hipSuccess ||
      hipMalloc((void**)&j->d_sj, sizeof(SegJob)) != hipSuccess ||
      hipMalloc((void**)&j->d_bs, sizeof(BatchState)) != hipSuccess ||
      hipEventCreate(&j->ev_all[0]) != hipSuccess || hipEventCreate(&j->ev_all[1]) != hipSuccess) {
    job_free(j);
    return fail(RS_E_NOMEM, "rs_job_view: allocation");
  }
  j->h_sj = parent->h_sj;
  j->h_sj.primes = j->dprimes;
  j->h_sj.recs = j->drecs;
  j->h_sj.krecs = j->dkrecs;
  if (hipMemcpy(j->d_sj, &j->h_sj, sizeof(SegJob), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(j->d_bs, 0, sizeof(BatchState)) != hipSuccess) {
    job_free(j);
    return fail(RS_E_HIP, "rs_job_view: upload");
  }
  *out = j;
  return RS_OK;
}

extern "C" int rs_job_reconfigure(rs_job* j, const rs_params* prm) {
  if (!j || !prm) return fail(RS_E_ARG, "rs_job_reconfigure: null");
  if (j->parent) return fail(RS_E_ARG, "rs_job_reconfigure: reconfigure the parent job, not a view");
  int rc = check_params(j->ctx, prm, "rs_job_reconfigure");
  if (rc) return rc;
  HIPCHK(hipSetDevice(j->ctx->device));
  HIPCHK(hipStreamSynchronize(j->ctx->stream));
  const uint32_t want = std::max<uint32_t>(4096, 2 * (prm->seg_primes ? prm->seg_primes : 65536));
  if (want > j->seg_cap) {
    uint64_t* np = nullptr;
    PrimeRec* nr = nullptr;
    KRec* nk = nullptr;
    if (hipMalloc((void**)&np, (size_t)want * 8) != hipSuccess ||
        hipMalloc((void**)&nr, (size_t)want * sizeof(PrimeRec)) != hipSuccess ||
        hipMalloc((void**)&nk, (size_t)want * j->h.kcount * 2 * sizeof(KRec)) != hipSuccess) {
      (void)hipFree(np);
      (void)hipFree(nr);
      return fail(RS_E_NOMEM, "rs_job_reconfigure: segment buffers");
    }
    (void)hipFree(j->dprimes);
    (void)hipFree(j->drecs);
    (void)hipFree(j->dkrecs);
    j->dprimes = np;
    j->drecs = nr;
    j->dkrecs = nk;
    j->seg_cap = want;
    j->h_sj.primes = np;
    j->h_sj.recs = nr;
    j->h_sj.krecs = nk;
    j->h_sj.cap = want;
    HIPCHK(hipMemcpy(j->d_sj, &j->h_sj, sizeof(SegJob), hipMemcpyHostToDevice));
  }
  j->prm.seg_primes = prm->seg_primes;
  apply_params(j, prm);
  HIPCHK(hipMemcpy(j->djob, &j->h, sizeof(JobDev), hipMemcpyHostToDevice));
  return RS_OK;
}

// The candidate set of the tuner and the cost model: slot 0 is the reference
// split; then the merged table at 1024..16384 entries, filter words 1, 2 or 4
// per entry (where the LDS holds them), and 256, 512 or 1024 threads per
// workgroup; then merged spans sized so the giant steps fill F = 1..4 whole
// rounds of the workgroup (the Q of these depends on the n-range); then the
// residue-class tables.
extern "C" int rs_tune_candidates(rs_ctx* ctx, uint32_t nmin, uint32_t nmax, int32_t kcount, rs_params* out, int cap,
                                  int* n) {
  if (!ctx || !out || !n || cap < 1) return fail(RS_E_ARG, "rs_tune_candidates: bad args");
  rs_params ref;
  int rc = rs_params_reference(nmin, nmax, kcount, 65536, ctx->cus, &ref);
  if (rc) return rc;
  int m = 0;
  out[m++] = ref;
  const uint32_t entries[5] = {1024, 2048, 4096, 8192, 16384};
  const uint32_t load[3] = {1, 2, 4};
  const uint32_t threads[3] = {256, 512, 1024};
  for (uint32_t e : entries)
    for (uint32_t l : load)
      for (uint32_t t : threads) {
        rs_params p = ref;
        p.algo = RS_ALGO_MERGED;
        p.Q = 2 * e;
        p.hsize = e * l;
        p.threads = t;
        if (check_params(ctx, &p, "rs_tune_candidates") != RS_OK) continue;
        if (m < cap) out[m] = p;
        m++;
      }
  // Spans whose giant-step count M is an exact multiple of the workgroup size
  // (F full rounds, no partial last round), filter load 1 or 1/2.  These 24
  // slots are always listed, so a slot means the same shape for every job; a
  // slot the n-range or the device cannot use has threads = 0 (rs_autotune
  // then reports it as not run).
  const uint64_t L = (uint64_t)nmax - nmin + 1;
  for (uint32_t t : threads)
    for (uint32_t F = 1; F <= 4; F++)
      for (uint32_t l = 1; l <= 2; l++) {
        const uint64_t Q = 2 * ((L + 2ull * t * F - 1) / (2ull * t * F));
        const uint64_t e = Q / 2;
        uint32_t hs = 64;
        while (hs < e && hs < (1u << 30)) hs <<= 1;
        rs_params p = ref;
        p.algo = RS_ALGO_MERGED;
        p.Q = (uint32_t)std::min<uint64_t>(Q, 1u << 30);
        p.hsize = hs * l;
        p.threads = t;
        if (e < 512 || e > 16384 || (L + Q - 1) / Q != (uint64_t)t * F ||
            check_params(ctx, &p, "rs_tune_candidates") != RS_OK)
          p.threads = 0;
        if (m < cap) out[m] = p;
        m++;
      }
  // Residue-class tables of at most 1024..16384 entries, filter load 1 or
  // 1/2, 256 or 512 threads (20 fixed slots; threads = 0 where the LDS
  // cannot hold them).
  for (uint32_t e : entries)
    for (uint32_t l = 1; l <= 2; l++)
      for (uint32_t t : {256u, 512u}) {
        rs_params p = ref;
        p.algo = RS_ALGO_CLASSES;
        p.Q = e;
        p.hsize = e * l;
        p.threads = t;
        if (check_params(ctx, &p, "rs_tune_candidates") != RS_OK) p.threads = 0;
        if (m < cap) out[m] = p;
        m++;
      }
  *n = m;
  rs::set_error("");
  return RS_OK;
}

extern "C" int rs_autotune(rs_job* j, uint64_t pmin, uint64_t width, const rs_params* cand, int ncand, double* seconds,
                           uint64_t* primes, int* best) {
  if (!j || !cand || ncand < 1 || !seconds || !primes || !best || width == 0 || pmin < 3)
    return fail(RS_E_ARG, "rs_autotune: bad args");
  HIPCHK(hipSetDevice(j->ctx->device));
  hipStream_t s = j->ctx->stream;
  HIPCHK(hipStreamSynchronize(s));
  // snapshot the state the samples change
  uint32_t* saved_bm = nullptr;
  HIPCHK(hipMalloc((void**)&saved_bm, j->nwords * 4));
  Counters saved_cnt;
  int rc = RS_OK;
  if (hipMemcpy(saved_bm, j->dbitmap, j->nwords * 4, hipMemcpyDeviceToDevice) != hipSuccess ||
      hipMemcpy(&saved_cnt, j->dcnt, sizeof(Counters), hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(RS_E_HIP, "rs_autotune: snapshot");
  const rs_params orig = j->prm;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(RS_E_HIP, "rs_autotune: events");
  int ib = -1;
  double tb = 0;
  for (int i = -1; i < ncand && !rc; i++) {
    // i == -1: one untimed pass with the first candidate (code-object load, caches)
    rc = rs_job_reconfigure(j, &cand[i < 0 ? 0 : i]);
    if (rc) {
      if (i >= 0) {
        seconds[i] = -1.0;
        primes[i] = 0;
        rc = RS_OK;
        continue;
      }
      break;
    }
    const uint64_t pmax = (pmin + width < pmin) ? ~0ull : pmin + width;
    if (hipEventRecord(e0, s) != hipSuccess) rc = fail(RS_E_HIP, "rs_autotune: event");
    if (!rc) rc = rs_sieve(j, pmin, pmax);
    if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
      rc = fail(RS_E_HIP, "rs_autotune: sync");
    Counters c;
    if (!rc && hipMemcpy(&c, j->dcnt, sizeof(Counters), hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(RS_E_HIP, "rs_autotune: counters");
    float ms = 0;
    if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = fail(RS_E_HIP, "rs_autotune: timing");
    // restore
    if (!rc && (hipMemcpy(j->dbitmap, saved_bm, j->nwords * 4, hipMemcpyDeviceToDevice) != hipSuccess ||
                hipMemcpy(j->dcnt, &saved_cnt, sizeof(Counters), hipMemcpyHostToDevice) != hipSuccess))
      rc = fail(RS_E_HIP, "rs_autotune: restore");
    if (rc || i < 0) continue;
    seconds[i] = ms * 1e-3;
    primes[i] = c.primecount - saved_cnt.primecount;
    if (ib < 0 || seconds[i] < tb) {
      ib = i;
      tb = seconds[i];
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(saved_bm);
  j->ev_used = 0;
  j->timer_started = false;
  if (rc) {
    rs_job_reconfigure(j, &orig);
    return rc;
  }
  *best = ib;
  if (ib >= 0) return rs_job_reconfigure(j, &cand[ib]);
  return rs_job_reconfigure(j, &orig);
}

// The batch split of one job: for each option i the job is reconfigured to
// seg_primes[i] target primes per segment and [pmin, pmin + width) is cut
// into views[i] equal parts sieved as one batch (the job and views[i] - 1
// views, one persistent BSGS launch per round), timed with stream events;
// the bitmap and counters are restored after every option, the views
// destroyed, and the job left on its original configuration.
extern "C" int rs_autotune_split(rs_job* j, uint64_t pmin, uint64_t width, const uint32_t* views,
                                 const uint32_t* seg_primes, int nopt, double* seconds, uint64_t* primes, int* best) {
  if (!j || !views || !seg_primes || nopt < 1 || !seconds || !primes || !best || width == 0 || pmin < 3)
    return fail(RS_E_ARG, "rs_autotune_split: bad args");
  if (j->parent) return fail(RS_E_ARG, "rs_autotune_split: tune the parent job, not a view");
  for (int i = 0; i < nopt; i++)
    if (views[i] < 1 || views[i] > RS_MAX_BATCH || seg_primes[i] < 1024)
      return fail(RS_E_ARG, "rs_autotune_split: option %d: views %u (1..%d), seg_primes %u (>= 1024)", i, views[i],
                  RS_MAX_BATCH, seg_primes[i]);
  HIPCHK(hipSetDevice(j->ctx->device));
  hipStream_t s = j->ctx->stream;
  HIPCHK(hipStreamSynchronize(s));
  uint32_t* saved_bm = nullptr;
  HIPCHK(hipMalloc((void**)&saved_bm, j->nwords * 4));
  Counters saved_cnt;
  int rc = RS_OK;
  if (hipMemcpy(saved_bm, j->dbitmap, j->nwords * 4, hipMemcpyDeviceToDevice) != hipSuccess ||
      hipMemcpy(&saved_cnt, j->dcnt, sizeof(Counters), hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(RS_E_HIP, "rs_autotune_split: snapshot");
  const rs_params orig = j->prm;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(RS_E_HIP, "rs_autotune_split: events");
  const uint64_t pmax = (pmin + width < pmin) ? ~0ull : pmin + width;
  int ib = -1;
  double tb = 0;
  std::vector<rs_job*> handles;
  std::vector<uint64_t> lo, hi;
  for (int i = -1; i < nopt && !rc; i++) {
    // i == -1: one untimed pass of the first option (code-object load, caches)
    const int o = i < 0 ? 0 : i;
    rs_params prm = orig;
    prm.seg_primes = seg_primes[o];
    rc = rs_job_reconfigure(j, &prm);
    if (rc) break;
    const uint32_t V = views[o];
    handles.assign(1, j);
    for (uint32_t v = 1; v < V && !rc; v++) {
      rs_job* h = nullptr;
      rc = rs_job_view(j, &h);
      if (!rc) handles.push_back(h);
    }
    if (!rc) {
      lo.resize(V);
      hi.resize(V);
      const uint64_t part = (pmax - pmin) / V;
      for (uint32_t v = 0; v < V; v++) {
        lo[v] = pmin + v * part;
        hi[v] = v + 1 < V ? pmin + (v + 1) * part : pmax;
      }
      // capture this batch shape's graphs outside the timed region: one round
      // over a one-integer range per view has the same launch shape
      std::vector<uint64_t> lo1(lo);
      for (uint32_t v = 0; v < V; v++) lo1[v] = lo[v] + 1;
      rc = rs_sieve_batch(handles.data(), (int)V, lo.data(), lo1.data());
      if (!rc && hipEventRecord(e0, s) != hipSuccess) rc = fail(RS_E_HIP, "rs_autotune_split: event");
      if (!rc) rc = rs_sieve_batch(handles.data(), (int)V, lo.data(), hi.data());
      if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
        rc = fail(RS_E_HIP, "rs_autotune_split: sync");
    }
    Counters c;
    if (!rc && hipMemcpy(&c, j->dcnt, sizeof(Counters), hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(RS_E_HIP, "rs_autotune_split: counters");
    float ms = 0;
    if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = fail(RS_E_HIP, "rs_autotune_split: timing");
    for (size_t v = 1; v < handles.size(); v++) rs_job_destroy(handles[v]);
    handles.clear();
    if (!rc && (hipMemcpy(j->dbitmap, saved_bm, j->nwords * 4, hipMemcpyDeviceToDevice) != hipSuccess ||
                hipMemcpy(j->dcnt, &saved_cnt, sizeof(Counters), hipMemcpyHostToDevice) != hipSuccess))
      rc = fail(RS_E_HIP, "rs_autotune_split: restore");
    if (rc || i < 0) continue;
    seconds[i] = ms * 1e-3;
    primes[i] = c.primecount - saved_cnt.primecount;
    if (ib < 0 || seconds[i] / std::max<double>(1.0, (double)primes[i]) < tb) {
      ib = i;
      tb = seconds[i] / std::max<double>(1.0, (double)primes[i]);
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(saved_bm);
  j->ev_used = 0;
  j->timer_started = false;
  const int rc2 = rs_job_reconfigure(j, &orig);
  if (rc) return rc;
  *best = ib;
  return rc2;
}
