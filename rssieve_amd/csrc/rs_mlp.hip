// rs_mlp.hip -- the learned cost model: a small fp32 MLP (inputs -> H -> H ->
// outputs, ReLU) trained and evaluated by hand-written HIP kernels.  It
// predicts, for a job descriptor, a log time per prime for every candidate
// BSGS configuration; the lowest prediction picks the configuration.  There
// is no counterpart in the reference.
//
// Layout: row-major activations [n][width]; weights W[in][out], bias[out].
// Loss: masked mean squared error / 2 over the unmasked (sample, output)
// entries.  Optimiser: Adam (Kingma & Ba 2015).
#include <hip/hip_runtime.h>

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

#define MCHK(expr)                                                                                       \
  do {                                                                                                   \
    hipError_t e_ = (expr);                                                                              \
    if (e_ != hipSuccess) return mfail(RS_E_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                                       __LINE__);                                                       \
  } while (0)

// out[i][j] = act(sum_k in[i][k] W[k][j] + b[j]); one thread per (i, j),
// reading its weight column straight from global memory (the layers are at
// most 64 wide, so the column stays in cache).
__global__ void k_dense_fwd(const float* __restrict__ in, const float* __restrict__ W, const float* __restrict__ b,
                            float* __restrict__ z, float* __restrict__ out, int n, int a, int bw, int relu) {
  const int j = blockIdx.y * blockDim.x + threadIdx.x;  // output unit
  const int i = blockIdx.x;                             // sample
  if (i >= n || j >= bw) return;
  float acc = b[j];
  for (int k = 0; k < a; k++) acc = fmaf(in[(size_t)i * a + k], W[(size_t)k * bw + j], acc);
  if (z) z[(size_t)i * bw + j] = acc;
  out[(size_t)i * bw + j] = relu ? fmaxf(acc, 0.0f) : acc;
}

// dy = mask * (y - t) / count, and the per-entry squared error for the loss
__global__ void k_mse_grad(const float* __restrict__ y, const float* __restrict__ t, const float* __restrict__ m,
                           float* __restrict__ dy, float* __restrict__ sq, int total, float inv_count) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const float mk = m ? m[e] : 1.0f;
  const float r = (y[e] - t[e]) * mk;
  dy[e] = r * inv_count;
  sq[e] = r * r;
}

// din[i][k] = (sum_j dout[i][j] W[k][j]) * (z[i][k] > 0)
__global__ void k_dense_bwd_in(const float* __restrict__ dout, const float* __restrict__ W,
                               const float* __restrict__ zin, float* __restrict__ din, int n, int a, int bw) {
  const int k = blockIdx.y * blockDim.x + threadIdx.x;
  const int i = blockIdx.x;
  if (i >= n || k >= a) return;
  float acc = 0.0f;
  for (int j = 0; j < bw; j++) acc = fmaf(dout[(size_t)i * bw + j], W[(size_t)k * bw + j], acc);
  din[(size_t)i * a + k] = (zin[(size_t)i * a + k] > 0.0f) ? acc : 0.0f;
}

// dW[k][j] = sum_i in[i][k] dout[i][j];  db[j] = sum_i dout[i][j]  (k == a row)
__global__ void k_dense_bwd_w(const float* __restrict__ in, const float* __restrict__ dout, float* __restrict__ dW,
                              float* __restrict__ db, int n, int a, int bw) {
  const int j = blockIdx.y * blockDim.x + threadIdx.x;
  const int k = blockIdx.x;  // 0..a (a = bias row)
  if (j >= bw || k > a) return;
  float acc = 0.0f;
  if (k < a) {
    for (int i = 0; i < n; i++) acc = fmaf(in[(size_t)i * a + k], dout[(size_t)i * bw + j], acc);
    dW[(size_t)k * bw + j] = acc;
  } else {
    for (int i = 0; i < n; i++) acc += dout[(size_t)i * bw + j];
    db[j] = acc;
  }
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int count, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count) return;
  const float gi = g[e];
  const float mi = b1 * m[e] + (1.0f - b1) * gi;
  const float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
  m[e] = mi;
  v[e] = vi;
  p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// gather rows idx[0..nb) of src (width w) into dst
__global__ void k_gather_rows(const float* __restrict__ src, const int* __restrict__ idx, float* __restrict__ dst,
                              int nb, int w) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nb * w) return;
  const int r = e / w, c = e % w;
  dst[e] = src[(size_t)idx[r] * w + c];
}

}  // namespace

struct rs_mlp {
  rs_ctx* ctx = nullptr;
  hipStream_t s = nullptr;
  int n_in = 0, H = 0, C = 0;
  int cap = 0;  // activation rows allocated
  // parameters, gradients and Adam moments: one contiguous buffer each, in
  // the order W1, b1, W2, b2, W3, b3
  float *P = nullptr, *G = nullptr, *Mm = nullptr, *Vv = nullptr;
  size_t nparam = 0;
  size_t off[6] = {0};
  int64_t step = 0;
  // activations
  float *X = nullptr, *Z1 = nullptr, *A1 = nullptr, *Z2 = nullptr, *A2 = nullptr, *Y = nullptr;
  float *T = nullptr, *Mk = nullptr, *dY = nullptr, *dA2 = nullptr, *dA1 = nullptr, *SQ = nullptr;
  int* idx = nullptr;
};

static void mlp_free_act(rs_mlp* m) {
  float* bufs[] = {m->X, m->Z1, m->A1, m->Z2, m->A2, m->Y, m->T, m->Mk, m->dY, m->dA2, m->dA1, m->SQ};
  for (float* b : bufs) (void)hipFree(b);
  (void)hipFree(m->idx);
  m->X = m->Z1 = m->A1 = m->Z2 = m->A2 = m->Y = m->T = m->Mk = m->dY = m->dA2 = m->dA1 = m->SQ = nullptr;
  m->idx = nullptr;
  m->cap = 0;
}

static int mlp_reserve(rs_mlp* m, int n) {
  if (n <= m->cap) return RS_OK;
  mlp_free_act(m);
  const size_t N = (size_t)n;
  float** fb[] = {&m->X, &m->Z1, &m->A1, &m->Z2, &m->A2, &m->Y, &m->T, &m->Mk, &m->dY, &m->dA2, &m->dA1, &m->SQ};
  const size_t w[] = {(size_t)m->n_in, (size_t)m->H, (size_t)m->H, (size_t)m->H, (size_t)m->H, (size_t)m->C,
                      (size_t)m->C,    (size_t)m->C, (size_t)m->C, (size_t)m->H, (size_t)m->H, (size_t)m->C};
  for (int i = 0; i < 12; i++) MCHK(hipMalloc((void**)fb[i], N * w[i] * sizeof(float)));
  MCHK(hipMalloc((void**)&m->idx, N * sizeof(int)));
  m->cap = n;
  return RS_OK;
}

extern "C" int rs_mlp_create(rs_ctx* ctx, int n_in, int n_hidden, int n_out, uint64_t seed, rs_mlp** out) {
  if (!ctx || !out || n_in < 1 || n_hidden < 1 || n_out < 1 || n_in > 1024 || n_hidden > 1024 || n_out > 1024)
    return mfail(RS_E_ARG, "rs_mlp_create: bad shape");
  hipStream_t s;
  if (rs_context_stream(ctx, (void**)&s) != RS_OK) return RS_E_ARG;
  rs_mlp* m = new rs_mlp;
  m->ctx = ctx;
  m->s = s;
  m->n_in = n_in;
  m->H = n_hidden;
  m->C = n_out;
  const size_t sz[6] = {(size_t)n_in * n_hidden, (size_t)n_hidden, (size_t)n_hidden * n_hidden, (size_t)n_hidden,
                        (size_t)n_hidden * n_out, (size_t)n_out};
  size_t o = 0;
  for (int i = 0; i < 6; i++) {
    m->off[i] = o;
    o += sz[i];
  }
  m->nparam = o;
  if (hipMalloc((void**)&m->P, o * 4) != hipSuccess || hipMalloc((void**)&m->G, o * 4) != hipSuccess ||
      hipMalloc((void**)&m->Mm, o * 4) != hipSuccess || hipMalloc((void**)&m->Vv, o * 4) != hipSuccess) {
    rs_mlp_destroy(m);
    return mfail(RS_E_NOMEM, "rs_mlp_create: allocation");
  }
  // He-uniform initialisation from a splitmix64 stream
  std::vector<float> h(o, 0.0f);
  uint64_t st = seed;
  auto next = [&]() {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  const int fan_in[3] = {n_in, n_hidden, n_hidden};
  for (int l = 0; l < 3; l++) {
    const float lim = sqrtf(6.0f / (float)fan_in[l]);
    for (size_t e = 0; e < sz[2 * l]; e++) h[m->off[2 * l] + e] = (float)((next() >> 11) * 0x1.0p-53 * 2.0 - 1.0) * lim;
  }
  if (hipMemcpy(m->P, h.data(), o * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->Mm, 0, o * 4) != hipSuccess || hipMemset(m->Vv, 0, o * 4) != hipSuccess) {
    rs_mlp_destroy(m);
    return mfail(RS_E_HIP, "rs_mlp_create: upload");
  }
  *out = m;
  return RS_OK;
}

extern "C" void rs_mlp_destroy(rs_mlp* m) {
  if (!m) return;
  (void)hipStreamSynchronize(m->s);
  mlp_free_act(m);
  (void)hipFree(m->P);
  (void)hipFree(m->G);
  (void)hipFree(m->Mm);
  (void)hipFree(m->Vv);
  delete m;
}

extern "C" int rs_mlp_num_params(const rs_mlp* m, uint64_t* n) {
  if (!m || !n) return mfail(RS_E_ARG, "rs_mlp_num_params: null");
  *n = m->nparam;
  return RS_OK;
}

extern "C" int rs_mlp_get_params(rs_mlp* m, float* out) {
  if (!m || !out) return mfail(RS_E_ARG, "rs_mlp_get_params: null");
  MCHK(hipStreamSynchronize(m->s));
  MCHK(hipMemcpy(out, m->P, m->nparam * 4, hipMemcpyDeviceToHost));
  return RS_OK;
}

extern "C" int rs_mlp_set_params(rs_mlp* m, const float* in) {
  if (!m || !in) return mfail(RS_E_ARG, "rs_mlp_set_params: null");
  MCHK(hipStreamSynchronize(m->s));
  MCHK(hipMemcpy(m->P, in, m->nparam * 4, hipMemcpyHostToDevice));
  MCHK(hipMemset(m->Mm, 0, m->nparam * 4));
  MCHK(hipMemset(m->Vv, 0, m->nparam * 4));
  m->step = 0;
  return RS_OK;
}

static int forward_dev(rs_mlp* m, const float* X, int n) {
  const float* P = m->P;
  const int H = m->H, C = m->C;
  const dim3 blk(64);
  hipLaunchKernelGGL(k_dense_fwd, dim3(n, (H + 63) / 64), blk, 0, m->s, X, P + m->off[0], P + m->off[1], m->Z1, m->A1,
                     n, m->n_in, H, 1);
  hipLaunchKernelGGL(k_dense_fwd, dim3(n, (H + 63) / 64), blk, 0, m->s, m->A1, P + m->off[2], P + m->off[3], m->Z2,
                     m->A2, n, H, H, 1);
  hipLaunchKernelGGL(k_dense_fwd, dim3(n, (C + 63) / 64), blk, 0, m->s, m->A2, P + m->off[4], P + m->off[5],
                     (float*)nullptr, m->Y, n, H, C, 0);
  MCHK(hipGetLastError());
  return RS_OK;
}

// loss gradient into m->G for rows already in X/T/Mk; returns the loss
static int backward_dev(rs_mlp* m, const float* X, const float* Tt, const float* Mk, int n, float mask_count,
                        float* loss) {
  const int H = m->H, C = m->C, a = m->n_in;
  const int total = n * C;
  hipLaunchKernelGGL(k_mse_grad, dim3((total + 255) / 256), dim3(256), 0, m->s, m->Y, Tt, Mk, m->dY, m->SQ, total,
                     1.0f / mask_count);
  float* G = m->G;
  const float* P = m->P;
  const dim3 blk(64);
  // layer 3
  hipLaunchKernelGGL(k_dense_bwd_w, dim3(H + 1, (C + 63) / 64), blk, 0, m->s, m->A2, m->dY, G + m->off[4],
                     G + m->off[5], n, H, C);
  hipLaunchKernelGGL(k_dense_bwd_in, dim3(n, (H + 63) / 64), blk, 0, m->s, m->dY, P + m->off[4], m->Z2, m->dA2, n, H,
                     C);
  // layer 2
  hipLaunchKernelGGL(k_dense_bwd_w, dim3(H + 1, (H + 63) / 64), blk, 0, m->s, m->A1, m->dA2, G + m->off[2],
                     G + m->off[3], n, H, H);
  hipLaunchKernelGGL(k_dense_bwd_in, dim3(n, (H + 63) / 64), blk, 0, m->s, m->dA2, P + m->off[2], m->Z1, m->dA1, n, H,
                     H);
  // layer 1
  hipLaunchKernelGGL(k_dense_bwd_w, dim3(a + 1, (H + 63) / 64), blk, 0, m->s, X, m->dA1, G + m->off[0], G + m->off[1],
                     n, a, H);
  MCHK(hipGetLastError());
  if (loss) {
    std::vector<float> sq((size_t)total);
    MCHK(hipMemcpyAsync(sq.data(), m->SQ, (size_t)total * 4, hipMemcpyDeviceToHost, m->s));
    MCHK(hipStreamSynchronize(m->s));
    double acc = 0;
    for (float v : sq) acc += v;
    *loss = (float)(0.5 * acc / mask_count);
  }
  return RS_OK;
}

static float mask_count(const float* M, int n, int C) {
  if (!M) return (float)n * C;
  double c = 0;
  for (size_t e = 0; e < (size_t)n * C; e++) c += M[e];
  return c > 0 ? (float)c : 1.0f;
}

extern "C" int rs_mlp_forward(rs_mlp* m, const float* X, int n, float* Y) {
  if (!m || !X || !Y || n < 1) return mfail(RS_E_ARG, "rs_mlp_forward: bad args");
  int rc = mlp_reserve(m, n);
  if (rc) return rc;
  MCHK(hipMemcpyAsync(m->X, X, (size_t)n * m->n_in * 4, hipMemcpyHostToDevice, m->s));
  rc = forward_dev(m, m->X, n);
  if (rc) return rc;
  MCHK(hipMemcpyAsync(Y, m->Y, (size_t)n * m->C * 4, hipMemcpyDeviceToHost, m->s));
  MCHK(hipStreamSynchronize(m->s));
  return RS_OK;
}

extern "C" int rs_mlp_loss_grad(rs_mlp* m, const float* X, const float* T, const float* M, int n, float* loss,
                                float* grads) {
  if (!m || !X || !T || n < 1) return mfail(RS_E_ARG, "rs_mlp_loss_grad: bad args");
  int rc = mlp_reserve(m, n);
  if (rc) return rc;
  const size_t nc = (size_t)n * m->C;
  MCHK(hipMemcpyAsync(m->X, X, (size_t)n * m->n_in * 4, hipMemcpyHostToDevice, m->s));
  MCHK(hipMemcpyAsync(m->T, T, nc * 4, hipMemcpyHostToDevice, m->s));
  if (M) MCHK(hipMemcpyAsync(m->Mk, M, nc * 4, hipMemcpyHostToDevice, m->s));
  rc = forward_dev(m, m->X, n);
  if (!rc) rc = backward_dev(m, m->X, m->T, M ? m->Mk : nullptr, n, mask_count(M, n, m->C), loss);
  if (rc) return rc;
  if (grads) {
    MCHK(hipMemcpyAsync(grads, m->G, m->nparam * 4, hipMemcpyDeviceToHost, m->s));
    MCHK(hipStreamSynchronize(m->s));
  }
  return RS_OK;
}

extern "C" int rs_mlp_adam_step(rs_mlp* m, float lr) {
  if (!m) return mfail(RS_E_ARG, "rs_mlp_adam_step: null");
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  m->step++;
  const float c1 = 1.0f - powf(b1, (float)m->step), c2 = 1.0f - powf(b2, (float)m->step);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((m->nparam + 255) / 256)), dim3(256), 0, m->s, m->P, m->G, m->Mm, m->Vv,
                     (int)m->nparam, lr, b1, b2, eps, c1, c2);
  MCHK(hipGetLastError());
  return RS_OK;
}

// Mini-batch Adam over the whole set, 'epochs' passes in a seeded order.
extern "C" int rs_mlp_train(rs_mlp* m, const float* X, const float* T, const float* M, int n, int epochs, int batch,
                            float lr, uint64_t seed, float* final_loss) {
  if (!m || !X || !T || n < 1 || epochs < 1 || batch < 1) return mfail(RS_E_ARG, "rs_mlp_train: bad args");
  const int a = m->n_in, C = m->C;
  if (batch > n) batch = n;
  // the whole set lives on the device; batches are gathered rows
  float *dX = nullptr, *dT = nullptr, *dM = nullptr, *bX = nullptr, *bT = nullptr, *bM = nullptr;
  int* didx = nullptr;
  auto cleanup = [&]() {
    (void)hipFree(dX);
    (void)hipFree(dT);
    (void)hipFree(dM);
    (void)hipFree(bX);
    (void)hipFree(bT);
    (void)hipFree(bM);
    (void)hipFree(didx);
  };
  const size_t nc = (size_t)n * C;
  if (hipMalloc((void**)&dX, (size_t)n * a * 4) != hipSuccess || hipMalloc((void**)&dT, nc * 4) != hipSuccess ||
      hipMalloc((void**)&dM, nc * 4) != hipSuccess || hipMalloc((void**)&bX, (size_t)batch * a * 4) != hipSuccess ||
      hipMalloc((void**)&bT, (size_t)batch * C * 4) != hipSuccess ||
      hipMalloc((void**)&bM, (size_t)batch * C * 4) != hipSuccess || hipMalloc((void**)&didx, n * sizeof(int)) != hipSuccess) {
    cleanup();
    return mfail(RS_E_NOMEM, "rs_mlp_train: allocation");
  }
  std::vector<float> ones;
  if (!M) ones.assign(nc, 1.0f);
  int rc = mlp_reserve(m, batch);
  if (!rc && (hipMemcpy(dX, X, (size_t)n * a * 4, hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(dT, T, nc * 4, hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(dM, M ? M : ones.data(), nc * 4, hipMemcpyHostToDevice) != hipSuccess))
    rc = mfail(RS_E_HIP, "rs_mlp_train: upload");
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  uint64_t st = seed;
  auto rnd = [&]() {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    return (uint32_t)(st >> 33);
  };
  std::vector<float> hm((size_t)batch * C);
  for (int ep = 0; ep < epochs && !rc; ep++) {
    for (int i = n - 1; i > 0; i--) {
      const int jx = (int)(rnd() % (uint32_t)(i + 1));
      std::swap(order[i], order[jx]);
    }
    if (hipMemcpy(didx, order.data(), n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
      rc = mfail(RS_E_HIP, "rs_mlp_train: order upload");
      break;
    }
    for (int b0 = 0; b0 < n && !rc; b0 += batch) {
      const int nb = std::min(batch, n - b0);
      hipLaunchKernelGGL(k_gather_rows, dim3((nb * a + 255) / 256), dim3(256), 0, m->s, dX, didx + b0, bX, nb, a);
      hipLaunchKernelGGL(k_gather_rows, dim3((nb * C + 255) / 256), dim3(256), 0, m->s, dT, didx + b0, bT, nb, C);
      hipLaunchKernelGGL(k_gather_rows, dim3((nb * C + 255) / 256), dim3(256), 0, m->s, dM, didx + b0, bM, nb, C);
      // count of unmasked entries of this batch (host side from the mask)
      float cnt = 0;
      for (int r = 0; r < nb; r++)
        for (int c = 0; c < C; c++) cnt += M ? M[(size_t)order[b0 + r] * C + c] : 1.0f;
      if (cnt <= 0) continue;
      rc = forward_dev(m, bX, nb);
      if (!rc) rc = backward_dev(m, bX, bT, bM, nb, cnt, nullptr);
      if (!rc) rc = rs_mlp_adam_step(m, lr);
    }
  }
  if (!rc && final_loss) {
    // loss over the whole set
    std::vector<float> Y(nc);
    cleanup();
    rc = rs_mlp_forward(m, X, n, Y.data());
    if (!rc) {
      double acc = 0, cnt = 0;
      for (size_t e = 0; e < nc; e++) {
        const double mk = M ? M[e] : 1.0;
        const double r = (Y[e] - T[e]) * mk;
        acc += r * r;
        cnt += mk;
      }
      *final_loss = (float)(0.5 * acc / (cnt > 0 ? cnt : 1));
    }
    return rc;
  }
  (void)hipStreamSynchronize(m->s);
  cleanup();
  return rc;
}
