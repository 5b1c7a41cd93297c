// rs_model.cpp -- the learned cost model from C: load the model file the
// training script writes, and pick a job's BSGS configuration and batch
// split with it.  The networks run on the library's own MLP kernels
// (rs_mlp_forward); a job without a usable model keeps the reference split
// (rs_params_reference), which the callers use as the fallback.
//
// Model file (little-endian): "RSCM", u32 version = 1, u32 heads; per head
// u32 kind (0: BSGS candidates, 1: batch split), u32 n_in, hidden, n_out,
// f32 mu[n_in], f32 sd[n_in], f64 ymu, u32 nparams, f32 params[nparams] (the
// rs_mlp parameter order), and for a split head u32 nopt, u32 views[nopt],
// u32 seg_primes[nopt].  The job features are log2(base), log2(sequences),
// log2(terms + 1), log2(n-span), log2(p), normalised by mu and sd.  The
// candidates' head maps them to the log seconds per prime of each candidate
// (n_in 5, n_out = candidates); the split head maps them and one option,
// log2(views), log2(primes per segment), to that option's log seconds per
// prime relative to the job's mean (n_in 7, n_out 1), evaluated for every
// option.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rssieve.h"
#include "../../include/rssieve_mlp.h"

namespace rs {
void set_error(const std::string& msg);
}

namespace {

int mfail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  rs::set_error(buf);
  return code;
}

constexpr int kFeatures = 5;

struct Head {
  uint32_t n_in = 0, hidden = 0, n_out = 0;
  std::vector<float> mu, sd;
  double ymu = 0;
  rs_mlp* mlp = nullptr;
  std::vector<uint32_t> views, seg_primes;  // split head options
};

template <class T>
bool get(FILE* f, T* v, size_t n = 1) {
  return fread(v, sizeof(T), n, f) == n;
}

// predicted log(seconds per prime) of every candidate (candidates' head) or
// of every split option (split head: one row per option) for one job
int predict(Head& h, bool split, uint32_t base, int32_t kcount, uint64_t nterms, uint32_t span, double log2p,
            std::vector<float>* out) {
  if (base < 2 || kcount < 1 || span < 1) return mfail(RS_E_ARG, "rs_model: bad job shape");
  const double f[kFeatures] = {std::log2((double)base), std::log2((double)kcount), std::log2((double)nterms + 1.0),
                               std::log2((double)span), log2p};
  const uint32_t rows = split ? (uint32_t)h.views.size() : 1u;
  std::vector<float> x((size_t)rows * h.n_in);
  for (uint32_t r = 0; r < rows; r++) {
    float* xr = x.data() + (size_t)r * h.n_in;
    for (int i = 0; i < kFeatures; i++) xr[i] = (float)f[i];
    if (split) {
      xr[kFeatures] = (float)std::log2((double)h.views[r]);
      xr[kFeatures + 1] = (float)std::log2((double)h.seg_primes[r]);
    }
    for (uint32_t i = 0; i < h.n_in; i++) xr[i] = (xr[i] - h.mu[i]) / h.sd[i];
  }
  out->assign((size_t)rows * h.n_out, 0.0f);
  const int rc = rs_mlp_forward(h.mlp, x.data(), (int)rows, out->data());
  if (rc) return rc;
  for (auto& y : *out) y = (float)(y + h.ymu);
  return RS_OK;
}

}  // namespace

struct rs_model {
  Head cand, split;
  bool has_cand = false, has_split = false;
};

extern "C" void rs_model_free(rs_model* m) {
  if (!m) return;
  rs_mlp_destroy(m->cand.mlp);
  rs_mlp_destroy(m->split.mlp);
  delete m;
}

extern "C" int rs_model_load(rs_ctx* ctx, const char* path, rs_model** out) {
  if (!ctx || !path || !out) return mfail(RS_E_ARG, "rs_model_load: null argument");
  FILE* f = fopen(path, "rb");
  if (!f) return mfail(RS_E_IO, "rs_model_load: cannot open %s", path);
  rs_model* m = new rs_model;
  char magic[4];
  uint32_t version = 0, heads = 0;
  int rc = RS_OK;
  if (!get(f, magic, 4) || memcmp(magic, "RSCM", 4) || !get(f, &version) || version != 1 || !get(f, &heads) ||
      heads < 1 || heads > 2)
    rc = mfail(RS_E_IO, "rs_model_load: %s is not a version-1 cost model file", path);
  for (uint32_t k = 0; k < heads && !rc; k++) {
    uint32_t kind = 0;
    Head h;
    uint32_t np = 0;
    if (!get(f, &kind) || kind > 1 || !get(f, &h.n_in) || !get(f, &h.hidden) || !get(f, &h.n_out) ||
        h.n_in != (kind == 0 ? (uint32_t)kFeatures : (uint32_t)kFeatures + 2) || h.hidden < 1 || h.hidden > 4096 ||
        h.n_out < 1 || h.n_out > 4096 || (kind == 1 && h.n_out != 1)) {
      rc = mfail(RS_E_IO, "rs_model_load: bad head %u in %s", k, path);
      break;
    }
    h.mu.resize(h.n_in);
    h.sd.resize(h.n_in);
    std::vector<float> params;
    if (!get(f, h.mu.data(), h.n_in) || !get(f, h.sd.data(), h.n_in) || !get(f, &h.ymu) || !get(f, &np)) {
      rc = mfail(RS_E_IO, "rs_model_load: truncated head %u in %s", k, path);
      break;
    }
    const uint64_t want = (uint64_t)h.n_in * h.hidden + h.hidden + (uint64_t)h.hidden * h.hidden + h.hidden +
                          (uint64_t)h.hidden * h.n_out + h.n_out;
    if (np != want) {
      rc = mfail(RS_E_IO, "rs_model_load: head %u has %u parameters, its shape needs %llu", k, np,
                 (unsigned long long)want);
      break;
    }
    params.resize(np);
    if (!get(f, params.data(), np)) {
      rc = mfail(RS_E_IO, "rs_model_load: truncated parameters in %s", path);
      break;
    }
    if (kind == 1) {
      uint32_t nopt = 0;
      if (!get(f, &nopt) || nopt < 1 || nopt > 1024) {
        rc = mfail(RS_E_IO, "rs_model_load: bad split option count in %s", path);
        break;
      }
      h.views.resize(nopt);
      h.seg_primes.resize(nopt);
      if (!get(f, h.views.data(), nopt) || !get(f, h.seg_primes.data(), nopt)) {
        rc = mfail(RS_E_IO, "rs_model_load: truncated split options in %s", path);
        break;
      }
      for (uint32_t i = 0; i < nopt && !rc; i++)
        if (h.views[i] < 1 || h.seg_primes[i] < 1) rc = mfail(RS_E_IO, "rs_model_load: bad split option %u", i);
      if (rc) break;
    }
    for (uint32_t i = 0; i < h.n_in; i++)
      if (!(h.sd[i] > 0.0f)) rc = mfail(RS_E_IO, "rs_model_load: feature scale %u is not positive", i);
    if (rc) break;
    rc = rs_mlp_create(ctx, (int)h.n_in, (int)h.hidden, (int)h.n_out, 1, &h.mlp);
    if (!rc) rc = rs_mlp_set_params(h.mlp, params.data());
    if (rc) {
      rs_mlp_destroy(h.mlp);
      break;
    }
    Head& dst = kind == 0 ? m->cand : m->split;
    bool& has = kind == 0 ? m->has_cand : m->has_split;
    if (has) {
      rs_mlp_destroy(h.mlp);
      rc = mfail(RS_E_IO, "rs_model_load: two heads of kind %u in %s", kind, path);
      break;
    }
    dst = std::move(h);
    has = true;
  }
  fclose(f);
  if (rc) {
    rs_model_free(m);
    return rc;
  }
  *out = m;
  return RS_OK;
}

extern "C" int rs_model_rank(rs_model* m, uint32_t base, int32_t kcount, uint64_t nterms, uint32_t span,
                             double log2p, const rs_params* cand, int ncand, uint32_t threads_only, int k, int* idx,
                             int* nout) {
  if (!m || !cand || !idx || !nout || k < 1) return mfail(RS_E_ARG, "rs_model_rank: null argument or k < 1");
  if (!m->has_cand) return mfail(RS_E_ARG, "rs_model_rank: the model has no candidate head");
  if ((uint32_t)ncand != m->cand.n_out)
    return mfail(RS_E_ARG, "rs_model_rank: %d candidates, the model predicts %u", ncand, m->cand.n_out);
  std::vector<float> y;
  const int rc = predict(m->cand, false, base, kcount, nterms, span, log2p, &y);
  if (rc) return rc;
  std::vector<int> ok;
  for (int i = 0; i < ncand; i++)
    if (cand[i].threads && (!threads_only || cand[i].threads == threads_only)) ok.push_back(i);
  if (ok.empty()) {  // nothing usable: the reference split (candidate 0)
    idx[0] = 0;
    *nout = 1;
    return RS_OK;
  }
  // fastest first; equal predictions keep the lower index
  std::stable_sort(ok.begin(), ok.end(), [&](int a, int b) { return y[a] < y[b]; });
  const int n = std::min<int>(k, (int)ok.size());
  for (int i = 0; i < n; i++) idx[i] = ok[i];
  *nout = n;
  return RS_OK;
}

extern "C" int rs_model_choose(rs_model* m, uint32_t base, int32_t kcount, uint64_t nterms, uint32_t span,
                               double log2p, const rs_params* cand, int ncand, uint32_t threads_only, int* idx) {
  if (!idx) return mfail(RS_E_ARG, "rs_model_choose: null argument");
  int n = 0;
  return rs_model_rank(m, base, kcount, nterms, span, log2p, cand, ncand, threads_only, 1, idx, &n);
}

extern "C" int rs_model_split(rs_model* m, uint32_t base, int32_t kcount, uint64_t nterms, uint32_t span,
                              double log2p, uint32_t* views, uint32_t* seg_primes) {
  if (!m || !views || !seg_primes) return mfail(RS_E_ARG, "rs_model_split: null argument");
  if (!m->has_split) return mfail(RS_E_ARG, "rs_model_split: the model has no batch-split head");
  std::vector<float> y;
  const int rc = predict(m->split, true, base, kcount, nterms, span, log2p, &y);
  if (rc) return rc;
  uint32_t best = 0;
  for (uint32_t i = 1; i < (uint32_t)y.size(); i++)
    if (y[i] < y[best]) best = i;
  *views = m->split.views[best];
  *seg_primes = m->split.seg_primes[best];
  return RS_OK;
}
