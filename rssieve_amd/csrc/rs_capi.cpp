// rs_capi.cpp -- host side of the C-ABI library: device context, jobs, the
// segment loop and counter / factor read-back.  Replaces the reference host's
// OpenCL plumbing (simpleCL.c) and the device half of cl_sieve.cpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rssieve.h"
#include "rs_internal.h"

namespace rs {
hipError_t launch_seg_begin(SegState* seg, hipStream_t s);
hipError_t launch_smallprimes(uint64_t low, uint64_t high, uint64_t* out, uint32_t cap, SegState* seg, hipStream_t s);
hipError_t launch_primegen(uint64_t low, uint64_t high, int32_t widx, uint64_t* out, uint32_t cap, SegState* seg,
                           hipStream_t s);
hipError_t launch_setup(const JobDev* job, const uint64_t* primes, const SegState* seg, PrimeRec* recs, KRec* krecs,
                        Counters* cnt, uint32_t cap, hipStream_t s);
hipError_t bsgs_blocks_per_cu(uint32_t threads, uint32_t hsize, int* out);
hipError_t launch_bsgs(uint32_t threads, uint32_t hsize, uint32_t nblocks, const JobDev* job, const PrimeRec* recs,
                       const KRec* krecs, SegState* seg, uint32_t* bitmap, FactorRec* fac, Counters* cnt,
                       uint32_t cap, hipStream_t s);
size_t bsgs_lds_bytes(uint32_t hsize);
}  // namespace rs

using namespace rs;

static thread_local std::string g_err;

namespace rs {
void set_error(const std::string& msg) { g_err = msg; }
}  // namespace rs

static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return fail(RS_E_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                      __FILE__, __LINE__);                                    \
  } while (0)

struct rs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t cus = 0;
  uint32_t lds_per_block = 0;
  char name[256] = {0};
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
  SegState* dseg = nullptr;
  uint64_t* dprimes = nullptr;
  PrimeRec* drecs = nullptr;
  KRec* dkrecs = nullptr;
  // timing of BSGS launches of the last rs_sieve call
  std::vector<hipEvent_t> ev;  // pairs
  size_t ev_used = 0;
  hipEvent_t ev_all[2] = {nullptr, nullptr};
  bool timed = false;
};

extern "C" const char* rs_last_error(void) { return g_err.c_str(); }

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
  (void)hipFree(j->djob);
  (void)hipFree(j->dbitmap);
  (void)hipFree(j->dcnt);
  (void)hipFree(j->dfac);
  (void)hipFree(j->dseg);
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
  if (prm->algo != RS_ALGO_THREE_LISTS) return fail(RS_E_ARG, "rs_job_create: algorithm %d not available", prm->algo);
  if (!(prm->threads == 128 || prm->threads == 256 || prm->threads == 512 || prm->threads == 1024))
    return fail(RS_E_ARG, "rs_job_create: threads %u not in {128,256,512,1024}", prm->threads);
  if (prm->Q < 16 || !pow2(prm->hsize) || prm->hsize <= prm->Q || prm->Q > 32768)
    return fail(RS_E_ARG, "rs_job_create: need Q in [16, 32768] and a power-of-two hsize > Q (Q=%u hsize=%u)",
                prm->Q, prm->hsize);
  if (bsgs_lds_bytes(prm->hsize) + 64 > ctx->lds_per_block)
    return fail(RS_E_ARG, "rs_job_create: hash of %u slots needs %zu LDS bytes, device has %u", prm->hsize,
                bsgs_lds_bytes(prm->hsize), ctx->lds_per_block);
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
  h.algo = (uint32_t)prm->algo;
  h.Q = prm->Q;
  rs_params_derive(prm, nmin, nmax, &h.M, &h.QQ, &h.MM);
  h.hsize = prm->hsize;
  h.hmask = prm->hsize - 1;
  h.threads = prm->threads;
  h.record_factors = prm->record_factors ? 1 : 0;
  h.factor_cap = prm->factor_cap ? prm->factor_cap : 1;
  if (h.QQ / 2 >= prm->hsize) {
    job_free(j);
    return fail(RS_E_ARG, "rs_job_create: hsize %u too small for %u parity baby steps", prm->hsize, h.QQ / 2);
  }
  if (h.QQ > 65536) {
    job_free(j);
    return fail(RS_E_ARG, "rs_job_create: QQ = %u exceeds the 16-bit table index", h.QQ);
  }
  j->nwords = (uint64_t)kcount * words_per_seq;
  j->seg_cap = std::max<uint32_t>(4096, 2 * (prm->seg_primes ? prm->seg_primes : 65536));

  int bpc = 0;
  if (bsgs_blocks_per_cu(prm->threads, prm->hsize, &bpc) != hipSuccess || bpc < 1) bpc = 1;
  if (prm->blocks_per_cu && (int)prm->blocks_per_cu < bpc) bpc = (int)prm->blocks_per_cu;
  j->bsgs_blocks = (uint32_t)bpc * ctx->cus;

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
  ALLOC(j->dseg, sizeof(SegState));
  ALLOC(j->dprimes, (size_t)j->seg_cap * 8);
  ALLOC(j->drecs, (size_t)j->seg_cap * sizeof(PrimeRec));
  ALLOC(j->dkrecs, (size_t)j->seg_cap * kcount * sizeof(KRec));
#undef ALLOC
  hipStream_t s = ctx->stream;
  if (hipMemcpyAsync(j->djob, &h, sizeof(JobDev), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(j->dbitmap, bitmap, j->nwords * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemsetAsync(j->dcnt, 0, sizeof(Counters), s) != hipSuccess ||
      hipMemsetAsync(j->dseg, 0, sizeof(SegState), s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
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

extern "C" void rs_wheel_offset(uint64_t* start, int32_t* index);

static int enqueue_segment(rs_job* j, uint64_t p, uint64_t stop) {
  hipStream_t s = j->ctx->stream;
  HIPCHK(launch_seg_begin(j->dseg, s));
  if (p < 114) {
    const uint64_t stop_sm = stop > 114 ? 114 : stop;
    HIPCHK(launch_smallprimes(p, stop_sm, j->dprimes, j->seg_cap, j->dseg, s));
    p = stop_sm;
  }
  if (p < stop) {
    uint64_t ks = p;
    int32_t widx = 0;
    rs_wheel_offset(&ks, &widx);
    if (ks < stop && ks >= p) HIPCHK(launch_primegen(ks, stop, widx, j->dprimes, j->seg_cap, j->dseg, s));
  }
  HIPCHK(launch_setup(j->djob, j->dprimes, j->dseg, j->drecs, j->dkrecs, j->dcnt, j->seg_cap, s));
  if (j->timed) {
    if (j->ev_used + 2 > j->ev.size()) {
      for (int i = 0; i < 64; i++) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        j->ev.push_back(e);
      }
    }
    HIPCHK(hipEventRecord(j->ev[j->ev_used], s));
  }
  HIPCHK(launch_bsgs(j->h.threads, j->h.hsize, j->bsgs_blocks, j->djob, j->drecs, j->dkrecs, j->dseg, j->dbitmap,
                     j->dfac, j->dcnt, j->seg_cap, s));
  if (j->timed) {
    HIPCHK(hipEventRecord(j->ev[j->ev_used + 1], s));
    j->ev_used += 2;
  }
  return RS_OK;
}

extern "C" int rs_sieve(rs_job* j, uint64_t pmin, uint64_t pmax) {
  if (!j) return fail(RS_E_ARG, "rs_sieve: null job");
  if (pmin < 3) return fail(RS_E_ARG, "rs_sieve: pmin %llu < 3", (unsigned long long)pmin);
  if (pmax < pmin) return fail(RS_E_ARG, "rs_sieve: pmax < pmin");
  HIPCHK(hipSetDevice(j->ctx->device));
  const uint32_t target = j->prm.seg_primes ? j->prm.seg_primes : 65536;
  uint64_t range = segment_range(pmin, target);
  // the buffer holds seg_cap primes; with every integer coprime to 30 a prime
  // a segment of r integers has at most 8r/30 + 16 of them
  if (pmin < (1ull << 20)) range = std::min<uint64_t>(range, (uint64_t)(j->seg_cap - 16) * 30 / 8);
  if (range < 60) range = 60;
  j->ev_used = 0;
  j->timed = true;
  HIPCHK(hipEventRecord(j->ev_all[0], j->ctx->stream));
  uint64_t p = pmin;
  while (p < pmax) {
    uint64_t stop = p + range;
    if (stop > pmax || stop < p) stop = pmax;
    int rc = enqueue_segment(j, p, stop);
    if (rc) return rc;
    p = stop;
  }
  HIPCHK(hipEventRecord(j->ev_all[1], j->ctx->stream));
  return RS_OK;
}

extern "C" int rs_sieve_batch(rs_job** jobs, int njobs, const uint64_t* pmin, const uint64_t* pmax) {
  if (!jobs || njobs < 1 || !pmin || !pmax) return fail(RS_E_ARG, "rs_sieve_batch: bad args");
  for (int i = 0; i < njobs; i++) {
    int rc = rs_sieve(jobs[i], pmin[i], pmax[i]);
    if (rc) return rc;
  }
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
  if (bsgs_s) *bsgs_s = t;
  if (launches) *launches = j->ev_used / 2;
  if (total_s) {
    float ms = 0;
    if (j->ev_used) HIPCHK(hipEventElapsedTime(&ms, j->ev_all[0], j->ev_all[1]));
    *total_s = ms * 1e-3;
  }
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
  uint64_t *dp = nullptr;
  SegState* ds = nullptr;
  HIPCHK(hipMalloc((void**)&dp, (size_t)segcap * 8));
  if (hipMalloc((void**)&ds, sizeof(SegState)) != hipSuccess) {
    (void)hipFree(dp);
    return fail(RS_E_NOMEM, "rs_primegen: allocation failed");
  }
  const uint64_t range = (uint64_t)(segcap - 16) * 30 / 8;  // cannot overflow the buffer
  uint64_t total = 0;
  int rc = RS_OK;
  std::vector<uint64_t> tmp;
  for (uint64_t p = pmin; p < pmax && rc == RS_OK;) {
    uint64_t stop = p + range;
    if (stop > pmax || stop < p) stop = pmax;
    hipStream_t s = ctx->stream;
    uint64_t pp = p;
    if (launch_seg_begin(ds, s) != hipSuccess) rc = fail(RS_E_HIP, "rs_primegen: launch failed");
    if (rc == RS_OK && pp < 114) {
      const uint64_t stop_sm = stop > 114 ? 114 : stop;
      if (launch_smallprimes(pp, stop_sm, dp, segcap, ds, s) != hipSuccess) rc = fail(RS_E_HIP, "launch failed");
      pp = stop_sm;
    }
    if (rc == RS_OK && pp < stop) {
      uint64_t ks = pp;
      int32_t widx = 0;
      rs_wheel_offset(&ks, &widx);
      if (ks < stop && ks >= pp && launch_primegen(ks, stop, widx, dp, segcap, ds, s) != hipSuccess)
        rc = fail(RS_E_HIP, "rs_primegen: launch failed");
    }
    SegState hs;
    if (rc == RS_OK && (hipStreamSynchronize(s) != hipSuccess ||
                        hipMemcpy(&hs, ds, sizeof(hs), hipMemcpyDeviceToHost) != hipSuccess))
      rc = fail(RS_E_HIP, "rs_primegen: sync failed");
    if (rc == RS_OK) {
      const uint64_t n = hs.nprimes;
      if (n > segcap) {
        rc = fail(RS_E_OVERFLOW, "rs_primegen: segment overflow");
        break;
      }
      tmp.resize(n);
      if (n && hipMemcpy(tmp.data(), dp, n * 8, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(RS_E_HIP, "rs_primegen: copy failed");
      for (uint64_t i = 0; i < n && rc == RS_OK; i++) {
        if (total < cap) out[total] = tmp[i];
        total++;
      }
    }
    p = stop;
  }
  (void)hipFree(dp);
  (void)hipFree(ds);
  *count = total;
  return rc;
}
