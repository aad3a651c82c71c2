// selector.hip -- the learned kernel selector, on the device.
//
//   fs_csr_features   16 row-length statistics from fs_csr_analyze's integer
//                     statistics (same definitions as the CPU oracle)
//   fs_mlp_forward    16 -> H -> H -> K perceptron with ReLU, one block per sample
//   fs_mlp_backward   gradients of the masked mean-squared error; per-sample
//                     deltas first, then one thread per parameter summing over
//                     the batch in sample order (deterministic, no atomics)
//   fs_adam_step      bias-corrected Adam update
//   fs_mlp_select     standardise features, forward one sample, write the
//                     argmin output (the predicted-fastest variant)
//
// Sums are carried in fp64 and stored in fp32.  No counterpart in the
// reference, which picks launch parameters from static tables
// (spmv_csr.py:128-153, spmm_csr_opt_alg2.py:175-304).
#include <math.h>

#include "fs_host.hpp"

namespace fs {

constexpr int kMlpThreads = 64;

struct MlpView {
  int in, hid, out;
  const float* W1; const float* b1; const float* W2; const float* b2; const float* W3; const float* b3;
};

__host__ __device__ inline int64_t mlp_param_count(int in, int hid, int out) {
  return (int64_t)in * hid + hid + (int64_t)hid * hid + hid + (int64_t)hid * out + out;
}

__host__ __device__ inline MlpView mlp_view(const fs_mlp* n) {
  MlpView v;
  v.in = n->in_dim; v.hid = n->hidden; v.out = n->out_dim;
  const float* p = n->params;
  v.W1 = p; p += (int64_t)v.in * v.hid;
  v.b1 = p; p += v.hid;
  v.W2 = p; p += (int64_t)v.hid * v.hid;
  v.b2 = p; p += v.hid;
  v.W3 = p; p += (int64_t)v.hid * v.out;
  v.b3 = p;
  return v;
}

// --------------------------------------------------------------- features

__global__ void features_kernel(const int64_t* __restrict__ st, int64_t n_rows, int64_t n_cols,
                                int64_t nnz, float* __restrict__ f) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double n = (double)n_rows;
  const double mean = n_rows > 0 ? (double)nnz / n : 0.0;
  double var = n_rows > 0 ? (double)st[7] / n - mean * mean : 0.0;
  if (var < 0.0) var = 0.0;
  const double sd = sqrt(var);
  f[0] = (float)log2(n + 1.0);
  f[1] = (float)log2((double)n_cols + 1.0);
  f[2] = (float)log2((double)nnz + 1.0);
  f[3] = (float)mean;
  f[4] = (float)sd;
  f[5] = (float)(mean > 0.0 ? sd / mean : 0.0);
  f[6] = (float)log2((double)st[5] + 1.0);
  f[7] = (float)(n_rows > 0 ? st[6] : 0);
  f[8] = (float)(n_rows > 0 ? (double)st[8] / n : 0.0);
  for (int b = 0; b < 6; ++b) f[9 + b] = (float)(n_rows > 0 ? (double)st[9 + b] / n : 0.0);
  f[15] = (float)(nnz > 0 ? (double)st[15] / (double)nnz : 0.0);
}

// ---------------------------------------------------------------- forward

// One block (64 threads) per sample; hidden width <= 64 * 4, out <= 64 * 4.
__global__ __launch_bounds__(kMlpThreads) void mlp_forward_kernel(MlpView m, const float* __restrict__ X,
                                                                  float* __restrict__ H1,
                                                                  float* __restrict__ H2,
                                                                  float* __restrict__ Y) {
  extern __shared__ float sh[];  // x[in], h1[hid], h2[hid]
  const int64_t s = blockIdx.x;
  float* xs = sh;
  float* h1 = sh + m.in;
  float* h2 = h1 + m.hid;
  for (int i = threadIdx.x; i < m.in; i += kMlpThreads) xs[i] = X[s * m.in + i];
  __syncthreads();
  for (int j = threadIdx.x; j < m.hid; j += kMlpThreads) {
    double a = m.b1[j];
    for (int i = 0; i < m.in; ++i) a += (double)xs[i] * (double)m.W1[(int64_t)i * m.hid + j];
    const float v = a > 0.0 ? (float)a : 0.f;
    h1[j] = v;
    if (H1) H1[s * m.hid + j] = v;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < m.hid; j += kMlpThreads) {
    double a = m.b2[j];
    for (int i = 0; i < m.hid; ++i) a += (double)h1[i] * (double)m.W2[(int64_t)i * m.hid + j];
    const float v = a > 0.0 ? (float)a : 0.f;
    h2[j] = v;
    if (H2) H2[s * m.hid + j] = v;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < m.out; k += kMlpThreads) {
    double a = m.b3[k];
    for (int i = 0; i < m.hid; ++i) a += (double)h2[i] * (double)m.W3[(int64_t)i * m.out + k];
    Y[s * m.out + k] = (float)a;
  }
}

// --------------------------------------------------------------- backward

// Per-sample deltas: d3 = scale * M (Y - T); d2 = relu'(h2) W3 d3; d1 = relu'(h1) W2 d2.
__global__ __launch_bounds__(kMlpThreads) void mlp_delta_kernel(MlpView m, const float* __restrict__ H1,
                                                                const float* __restrict__ H2,
                                                                const float* __restrict__ Y,
                                                                const float* __restrict__ T,
                                                                const float* __restrict__ Mk,
                                                                double scale, float* __restrict__ D3,
                                                                float* __restrict__ D2,
                                                                float* __restrict__ D1,
                                                                double* __restrict__ sample_loss) {
  extern __shared__ float sh[];  // d3[out], d2[hid]
  const int64_t s = blockIdx.x;
  float* d3 = sh;
  float* d2 = sh + m.out;
  for (int k = threadIdx.x; k < m.out; k += kMlpThreads) {
    const double e = (double)Mk[s * m.out + k] * ((double)Y[s * m.out + k] - (double)T[s * m.out + k]);
    const float d = (float)(e * scale);
    d3[k] = d;
    D3[s * m.out + k] = d;
  }
  if (threadIdx.x == 0) {
    double l = 0.0;
    for (int k = 0; k < m.out; ++k) {
      const double e = (double)Mk[s * m.out + k] * ((double)Y[s * m.out + k] - (double)T[s * m.out + k]);
      l += 0.5 * e * e;
    }
    sample_loss[s] = l;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < m.hid; i += kMlpThreads) {
    double a = 0.0;
    for (int k = 0; k < m.out; ++k) a += (double)m.W3[(int64_t)i * m.out + k] * (double)d3[k];
    const float v = H2[s * m.hid + i] > 0.f ? (float)a : 0.f;
    d2[i] = v;
    D2[s * m.hid + i] = v;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < m.hid; i += kMlpThreads) {
    double a = 0.0;
    for (int j = 0; j < m.hid; ++j) a += (double)m.W2[(int64_t)i * m.hid + j] * (double)d2[j];
    D1[s * m.hid + i] = H1[s * m.hid + i] > 0.f ? (float)a : 0.f;
  }
}

// One thread per parameter: g = sum over samples (in order) of input * delta.
__global__ void mlp_grad_kernel(MlpView m, int64_t batch, const float* __restrict__ X,
                                const float* __restrict__ H1, const float* __restrict__ H2,
                                const float* __restrict__ D1, const float* __restrict__ D2,
                                const float* __restrict__ D3, const double* __restrict__ sample_loss,
                                double scale, float* __restrict__ grad, float* __restrict__ loss) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nW1 = (int64_t)m.in * m.hid, nW2 = (int64_t)m.hid * m.hid, nW3 = (int64_t)m.hid * m.out;
  const int64_t o_b1 = nW1, o_W2 = o_b1 + m.hid, o_b2 = o_W2 + nW2, o_W3 = o_b2 + m.hid,
                o_b3 = o_W3 + nW3, total = o_b3 + m.out;
  if (t == total && loss) {
    double l = 0.0;
    for (int64_t s = 0; s < batch; ++s) l += sample_loss[s];
    *loss = (float)(l * scale);
    return;
  }
  if (t >= total) return;
  const float* A;  // per-sample input of this parameter (null for biases)
  const float* D;  // per-sample delta
  int a_stride, d_stride, ai, dj;
  if (t < o_b1) { A = X; a_stride = m.in; ai = (int)(t / m.hid); D = D1; d_stride = m.hid; dj = (int)(t % m.hid); }
  else if (t < o_W2) { A = nullptr; a_stride = 0; ai = 0; D = D1; d_stride = m.hid; dj = (int)(t - o_b1); }
  else if (t < o_b2) { const int64_t u = t - o_W2; A = H1; a_stride = m.hid; ai = (int)(u / m.hid); D = D2; d_stride = m.hid; dj = (int)(u % m.hid); }
  else if (t < o_W3) { A = nullptr; a_stride = 0; ai = 0; D = D2; d_stride = m.hid; dj = (int)(t - o_b2); }
  else if (t < o_b3) { const int64_t u = t - o_W3; A = H2; a_stride = m.hid; ai = (int)(u / m.out); D = D3; d_stride = m.out; dj = (int)(u % m.out); }
  else { A = nullptr; a_stride = 0; ai = 0; D = D3; d_stride = m.out; dj = (int)(t - o_b3); }
  double g = 0.0;
  if (A) {
    for (int64_t s = 0; s < batch; ++s) g += (double)A[s * a_stride + ai] * (double)D[s * d_stride + dj];
  } else {
    for (int64_t s = 0; s < batch; ++s) g += (double)D[s * d_stride + dj];
  }
  grad[t] = (float)g;
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ mo,
                            float* __restrict__ vo, int64_t n, double lr, double b1, double b2,
                            double eps, double c1, double c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double gi = g[i];
  const double m = b1 * (double)mo[i] + (1.0 - b1) * gi;
  const double v = b2 * (double)vo[i] + (1.0 - b2) * gi * gi;
  mo[i] = (float)m;
  vo[i] = (float)v;
  p[i] = (float)((double)p[i] - lr * (m / c1) / (sqrt(v / c2) + eps));
}

// Standardise, forward one sample, argmin.
__global__ __launch_bounds__(kMlpThreads) void mlp_select_kernel(MlpView m, const float* __restrict__ feat,
                                                                 const float* __restrict__ mu,
                                                                 const float* __restrict__ sd,
                                                                 int32_t* __restrict__ choice,
                                                                 float* __restrict__ scores) {
  extern __shared__ float sh[];  // x[in], h1[hid], h2[hid], y[out]
  float* xs = sh;
  float* h1 = xs + m.in;
  float* h2 = h1 + m.hid;
  float* ys = h2 + m.hid;
  for (int i = threadIdx.x; i < m.in; i += kMlpThreads) {
    const float s = sd[i] > 0.f ? sd[i] : 1.f;
    xs[i] = (feat[i] - mu[i]) / s;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < m.hid; j += kMlpThreads) {
    double a = m.b1[j];
    for (int i = 0; i < m.in; ++i) a += (double)xs[i] * (double)m.W1[(int64_t)i * m.hid + j];
    h1[j] = a > 0.0 ? (float)a : 0.f;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < m.hid; j += kMlpThreads) {
    double a = m.b2[j];
    for (int i = 0; i < m.hid; ++i) a += (double)h1[i] * (double)m.W2[(int64_t)i * m.hid + j];
    h2[j] = a > 0.0 ? (float)a : 0.f;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < m.out; k += kMlpThreads) {
    double a = m.b3[k];
    for (int i = 0; i < m.hid; ++i) a += (double)h2[i] * (double)m.W3[(int64_t)i * m.out + k];
    ys[k] = (float)a;
    if (scores) scores[k] = (float)a;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int best = 0;
    for (int k = 1; k < m.out; ++k)
      if (ys[k] < ys[best]) best = k;
    *choice = best;
  }
}

static int check_mlp(const fs_mlp* n, const char* who) {
  if (!n || !n->params) return fail(FS_ERR_VALUE, "%s: network is null", who);
  if (n->in_dim <= 0 || n->hidden <= 0 || n->out_dim <= 0 || n->in_dim > 256 || n->hidden > 256 ||
      n->out_dim > 256)
    return fail(FS_ERR_VALUE, "%s: layer sizes must be in 1..256", who);
  return FS_OK;
}

}  // namespace fs

using namespace fs;

extern "C" int64_t fs_mlp_param_count(int32_t in_dim, int32_t hidden, int32_t out_dim) {
  if (in_dim <= 0 || hidden <= 0 || out_dim <= 0) return -1;
  return mlp_param_count(in_dim, hidden, out_dim);
}

extern "C" int fs_csr_features(const int64_t* stats, int64_t n_rows, int64_t n_cols, int64_t nnz,
                               float* features, void* stream) {
  if (!stats || !features) return fail(FS_ERR_VALUE, "fs_csr_features: null pointer");
  if (n_rows < 0 || n_cols < 0 || nnz < 0) return fail(FS_ERR_VALUE, "fs_csr_features: negative size");
  features_kernel<<<1, 64, 0, as_stream(stream)>>>(stats, n_rows, n_cols, nnz, features);
  return check_launch("fs_csr_features");
}

extern "C" int fs_mlp_forward(const fs_mlp* net, const float* X, int64_t batch, float* H1, float* H2,
                              float* Y, void* stream) {
  int rc = check_mlp(net, "fs_mlp_forward");
  if (rc) return rc;
  if (batch < 0 || batch > 0x7fffffffLL) return fail(FS_ERR_VALUE, "fs_mlp_forward: bad batch");
  if (batch == 0) return FS_OK;
  if (!X || !Y) return fail(FS_ERR_VALUE, "fs_mlp_forward: X or Y is null");
  MlpView m = mlp_view(net);
  const size_t shm = sizeof(float) * (m.in + 2 * m.hid);
  mlp_forward_kernel<<<(unsigned)batch, kMlpThreads, shm, as_stream(stream)>>>(m, X, H1, H2, Y);
  return check_launch("fs_mlp_forward");
}

extern "C" int fs_mlp_backward(const fs_mlp* net, const float* X, const float* H1, const float* H2,
                               const float* Y, const float* T, const float* M, int64_t batch,
                               float* grad, float* loss, void* workspace, int64_t workspace_bytes,
                               void* stream) {
  int rc = check_mlp(net, "fs_mlp_backward");
  if (rc) return rc;
  if (batch <= 0 || batch > 0x7fffffffLL) return fail(FS_ERR_VALUE, "fs_mlp_backward: bad batch");
  if (!X || !H1 || !H2 || !Y || !T || !M || !grad)
    return fail(FS_ERR_VALUE, "fs_mlp_backward: null pointer");
  MlpView m = mlp_view(net);
  const int64_t need = fs_mlp_workspace_bytes(net->in_dim, net->hidden, net->out_dim, batch);
  if (!workspace || workspace_bytes < need)
    return fail(FS_ERR_VALUE, "fs_mlp_backward: workspace too small (%lld < %lld bytes)",
                (long long)workspace_bytes, (long long)need);
  double* sample_loss = static_cast<double*>(workspace);
  float* D3 = reinterpret_cast<float*>(sample_loss + batch);
  float* D2 = D3 + batch * m.out;
  float* D1 = D2 + batch * m.hid;
  const double scale = 1.0 / ((double)batch * (double)m.out);
  hipStream_t st = as_stream(stream);
  const size_t shm = sizeof(float) * (m.out + m.hid);
  mlp_delta_kernel<<<(unsigned)batch, kMlpThreads, shm, st>>>(m, H1, H2, Y, T, M, scale, D3, D2, D1,
                                                              sample_loss);
  rc = check_launch("fs_mlp_backward");
  if (rc) return rc;
  const int64_t total = mlp_param_count(m.in, m.hid, m.out) + 1;  // + the loss thread
  mlp_grad_kernel<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(m, batch, X, H1, H2, D1, D2, D3,
                                                                  sample_loss, scale, grad, loss);
  return check_launch("fs_mlp_backward");
}

extern "C" int64_t fs_mlp_workspace_bytes(int32_t in_dim, int32_t hidden, int32_t out_dim,
                                          int64_t batch) {
  (void)in_dim;
  if (hidden <= 0 || out_dim <= 0 || batch < 0) return -1;
  return batch * (int64_t)sizeof(double) + batch * (int64_t)(out_dim + 2 * hidden) * (int64_t)sizeof(float);
}

extern "C" int fs_adam_step(float* params, const float* grad, float* m, float* v, int64_t n, double lr,
                            double beta1, double beta2, double eps, int64_t step, void* stream) {
  if (n < 0) return fail(FS_ERR_VALUE, "fs_adam_step: negative size");
  if (n == 0) return FS_OK;
  if (!params || !grad || !m || !v) return fail(FS_ERR_VALUE, "fs_adam_step: null pointer");
  if (step < 1) return fail(FS_ERR_VALUE, "fs_adam_step: step counts from 1");
  const double c1 = 1.0 - pow(beta1, (double)step), c2 = 1.0 - pow(beta2, (double)step);
  adam_kernel<<<(unsigned)((n + 255) / 256), 256, 0, as_stream(stream)>>>(params, grad, m, v, n, lr, beta1,
                                                                          beta2, eps, c1, c2);
  return check_launch("fs_adam_step");
}

extern "C" int fs_mlp_select(const fs_mlp* net, const float* features, const float* mu,
                             const float* sd, int32_t* choice, float* scores, void* stream) {
  int rc = check_mlp(net, "fs_mlp_select");
  if (rc) return rc;
  if (!features || !mu || !sd || !choice) return fail(FS_ERR_VALUE, "fs_mlp_select: null pointer");
  MlpView m = mlp_view(net);
  const size_t shm = sizeof(float) * (m.in + 2 * m.hid + m.out);
  mlp_select_kernel<<<1, kMlpThreads, shm, as_stream(stream)>>>(m, features, mu, sd, choice, scores);
  return check_launch("fs_mlp_select");
}
