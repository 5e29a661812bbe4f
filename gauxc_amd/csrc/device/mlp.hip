// Cost-model network of the task scheduler, trained and evaluated on the
// device with kernels of this file (FP32).
//
//   x (nin) -> [Linear -> ReLU] x nlayers (width nhidden) -> Linear -> y (nout)
//
// Row-major activations [n][width]; layer l has W_l [out][in] and b_l [out],
// stored one after the other in a single parameter vector (W1, b1, W2, b2,
// ...).  Loss: mean squared error over the n x nout outputs.  Optimiser:
// Adam (bias-corrected).  The dense products use one LDS-tiled kernel:
//   C[M][N] (+)= op(A) op(B), op = identity or transpose, row-major.
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <memory>
#include <random>
#include <vector>

namespace xcq {
namespace {

constexpr int kT = 32;  // tile edge of the dense-product kernel

// C = alpha * op(A) * op(B) + beta * C.  A is M x K (or K x M when ta),
// B is K x N (or N x K when tb); leading dimensions are the row lengths.
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, float alpha,
                                              const float* __restrict__ A, int lda, int ta,
                                              const float* __restrict__ B, int ldb, int tb,
                                              float beta, float* __restrict__ C, int ldc) {
  __shared__ float As[kT][kT + 1];
  __shared__ float Bs[kT][kT + 1];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const int m0 = blockIdx.y * kT, n0 = blockIdx.x * kT;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += kT) {
    for (int r = ty; r < kT; r += 8) {
      // As[r][tx] = op(A)[m0 + r][k0 + tx]
      const int m = m0 + r, k = k0 + tx;
      float a = 0.f;
      if (m < M && k < K) a = ta ? A[(int64_t)k * lda + m] : A[(int64_t)m * lda + k];
      As[r][tx] = a;
      // Bs[r][tx] = op(B)[k0 + r][n0 + tx]
      const int kb = k0 + r, n = n0 + tx;
      float b = 0.f;
      if (kb < K && n < N) b = tb ? B[(int64_t)n * ldb + kb] : B[(int64_t)kb * ldb + n];
      Bs[r][tx] = b;
    }
    __syncthreads();
#pragma unroll 8
    for (int k = 0; k < kT; ++k) {
      const float b = Bs[k][tx];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] += As[ty + 8 * q][k] * b;
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int m = m0 + ty + 8 * q, n = n0 + tx;
    if (m < M && n < N) {
      float* c = C + (int64_t)m * ldc + n;
      *c = alpha * acc[q] + (beta != 0.f ? beta * *c : 0.f);
    }
  }
}

// Z[n][j] += b[j]; A = relu(Z) when relu (A may alias Z when !relu).
__global__ __launch_bounds__(256) void k_bias_act(int64_t n, int w, const float* __restrict__ b,
                                                  float* __restrict__ Z, float* __restrict__ A,
                                                  int relu) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n * w) return;
  const float z = Z[i] + b[i % w];
  Z[i] = z;
  A[i] = relu ? fmaxf(z, 0.f) : z;
}

// dZ = dA * (Z > 0) in place on dA.
__global__ __launch_bounds__(256) void k_relu_bwd(int64_t cnt, const float* __restrict__ Z,
                                                  float* __restrict__ dA) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= cnt) return;
  if (!(Z[i] > 0.f)) dA[i] = 0.f;
}

// db[j] = sum_n dZ[n][j]: one workgroup per column, fixed-order reduction.
__global__ __launch_bounds__(256) void k_col_sum(int64_t n, int w, const float* __restrict__ dZ,
                                                 float* __restrict__ db) {
  __shared__ float s[256];
  const int j = blockIdx.x;
  float acc = 0.f;
  for (int64_t i = threadIdx.x; i < n; i += 256) acc += dZ[i * w + j];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[j] = s[0];
}

// dY = 2 (Y - T) / (n nout); per-block partial sums of (Y - T)^2 into part.
__global__ __launch_bounds__(256) void k_mse(int64_t cnt, const float* __restrict__ Y,
                                             const float* __restrict__ T, float* __restrict__ dY,
                                             float* __restrict__ part) {
  __shared__ float s[256];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  float d = 0.f;
  if (i < cnt) {
    d = Y[i] - T[i];
    dY[i] = 2.f * d / (float)cnt;
  }
  s[threadIdx.x] = d * d;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Adam step on the parameter vector.
__global__ __launch_bounds__(256) void k_adam(int64_t n, float* __restrict__ p,
                                              const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, float lr, float b1, float b2,
                                              float c1, float c2, float eps) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// rows[k] = src[idx[k]] (row length w): minibatch gather.
__global__ __launch_bounds__(256) void k_gather_rows(int64_t n, int w,
                                                     const int32_t* __restrict__ idx,
                                                     const float* __restrict__ src,
                                                     float* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n * w) return;
  const int64_t r = i / w, c = i % w;
  dst[i] = src[(int64_t)idx[r] * w + c];
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace
}  // namespace xcq

using namespace xcq;

struct xcq_mlp_s {
  int device = 0;
  hipStream_t stream = nullptr;
  int nin = 0, nh = 0, nl = 0, nout = 0;
  std::vector<int> widths;          // nin, nh, ..., nh, nout
  std::vector<int64_t> w_off, b_off;  // offsets of W_l, b_l in the parameter vector
  int64_t nparam = 0;
  float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;
  int64_t adam_t = 0;
  // activation workspace for up to cap rows
  int64_t cap = 0;
  std::vector<float*> Z, A;  // per layer (A[0] = input)
  float *d_T = nullptr, *d_dA = nullptr, *d_dB = nullptr, *d_part = nullptr;
  float *d_X = nullptr, *d_Y = nullptr;  // full training set (train calls)
  int32_t* d_idx = nullptr;
  int64_t data_cap = 0;

  ~xcq_mlp_s() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    free_work();
    for (float* q : {d_p, d_g, d_m, d_v, d_X, d_Y})
      if (q) (void)hipFree(q);
    if (d_idx) (void)hipFree(d_idx);
    if (stream) (void)hipStreamDestroy(stream);
  }
  void free_work() {
    for (auto* q : Z)
      if (q) (void)hipFree(q);
    for (size_t l = 1; l < A.size(); ++l)
      if (A[l]) (void)hipFree(A[l]);
    if (!A.empty() && A[0]) (void)hipFree(A[0]);
    Z.clear();
    A.clear();
    for (float* q : {d_T, d_dA, d_dB, d_part})
      if (q) (void)hipFree(q);
    d_T = d_dA = d_dB = d_part = nullptr;
    cap = 0;
  }
  int ensure(int64_t n) {
    if (n <= cap) return XCQ_SUCCESS;
    free_work();
    const int L = (int)widths.size() - 1;
    A.assign(L + 1, nullptr);
    Z.assign(L + 1, nullptr);
    int wmax = 0;
    for (int w : widths) wmax = std::max(wmax, w);
    for (int l = 0; l <= L; ++l) {
      XCQ_HIP(hipMalloc(&A[l], sizeof(float) * n * widths[l]));
      if (l > 0) XCQ_HIP(hipMalloc(&Z[l], sizeof(float) * n * widths[l]));
    }
    XCQ_HIP(hipMalloc(&d_T, sizeof(float) * n * nout));
    XCQ_HIP(hipMalloc(&d_dA, sizeof(float) * n * wmax));
    XCQ_HIP(hipMalloc(&d_dB, sizeof(float) * n * wmax));
    XCQ_HIP(hipMalloc(&d_part, sizeof(float) * (n * nout / 256 + 2)));
    cap = n;
    return XCQ_SUCCESS;
  }
};

namespace {

int gemm(hipStream_t s, int M, int N, int K, float alpha, const float* A, int lda, int ta,
         const float* B, int ldb, int tb, float beta, float* C, int ldc) {
  if (M == 0 || N == 0) return XCQ_SUCCESS;
  hipLaunchKernelGGL(k_gemm, dim3((N + kT - 1) / kT, (M + kT - 1) / kT), dim3(256), 0, s, M, N,
                     K, alpha, A, lda, ta, B, ldb, tb, beta, C, ldc);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

// Forward pass on the n rows already in m->A[0]; output in m->A[L].
int forward(xcq_mlp_s* m, int64_t n) {
  const int L = (int)m->widths.size() - 1;
  for (int l = 1; l <= L; ++l) {
    const int in = m->widths[l - 1], out = m->widths[l];
    int rc = gemm(m->stream, (int)n, out, in, 1.f, m->A[l - 1], in, 0, m->d_p + m->w_off[l - 1],
                  in, 1, 0.f, m->Z[l], out);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bias_act, dim3(nblk(n * out)), dim3(256), 0, m->stream, n, out,
                       m->d_p + m->b_off[l - 1], m->Z[l], m->A[l], l < L ? 1 : 0);
    XCQ_HIP(hipGetLastError());
  }
  return XCQ_SUCCESS;
}

// Loss and gradient into d_g for the n rows in A[0] with targets in d_T.
int backward(xcq_mlp_s* m, int64_t n, float* loss) {
  const int L = (int)m->widths.size() - 1;
  const int64_t cnt = n * m->nout;
  float* dcur = m->d_dA;
  float* dnext = m->d_dB;
  const unsigned nb = nblk(cnt);
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, m->stream, cnt, m->A[L], m->d_T, dcur,
                     m->d_part);
  XCQ_HIP(hipGetLastError());
  for (int l = L; l >= 1; --l) {
    const int in = m->widths[l - 1], out = m->widths[l];
    if (l < L) {
      hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(n * out)), dim3(256), 0, m->stream, n * out,
                         m->Z[l], dcur);
      XCQ_HIP(hipGetLastError());
    }
    // dW_l[out][in] = dZ^T (out x n) * A_{l-1} (n x in)
    int rc = gemm(m->stream, out, in, (int)n, 1.f, dcur, out, 1, m->A[l - 1], in, 0, 0.f,
                  m->d_g + m->w_off[l - 1], in);
    if (rc) return rc;
    hipLaunchKernelGGL(k_col_sum, dim3(out), dim3(256), 0, m->stream, n, out, dcur,
                       m->d_g + m->b_off[l - 1]);
    XCQ_HIP(hipGetLastError());
    if (l > 1) {
      // dA_{l-1}[n][in] = dZ (n x out) * W_l (out x in)
      rc = gemm(m->stream, (int)n, in, out, 1.f, dcur, out, 0, m->d_p + m->w_off[l - 1], in, 0,
                0.f, dnext, in);
      if (rc) return rc;
      std::swap(dcur, dnext);
    }
  }
  if (loss) {
    std::vector<float> part(nb);
    XCQ_HIP(hipMemcpyAsync(part.data(), m->d_part, 4 * (size_t)nb, hipMemcpyDeviceToHost,
                           m->stream));
    XCQ_HIP(hipStreamSynchronize(m->stream));
    double s = 0.;
    for (float x : part) s += x;
    *loss = (float)(s / (double)cnt);
  }
  return XCQ_SUCCESS;
}

}  // namespace

extern "C" {

int xcq_mlp_create(int device, int32_t nin, int32_t nhidden, int32_t nlayers, int32_t nout,
                   uint64_t seed, xcq_mlp* out) {
  XCQ_REQUIRE(out, XCQ_ERROR_INVALID_ARGUMENT, "null output");
  *out = nullptr;
  XCQ_REQUIRE(nin > 0 && nhidden > 0 && nlayers >= 0 && nout > 0, XCQ_ERROR_INVALID_ARGUMENT,
              "bad network shape");
  int n = 0;
  if (xcq_device_count(&n) != XCQ_SUCCESS || device < 0 || device >= n) {
    set_error("no HIP device available");
    return XCQ_ERROR_NO_DEVICE;
  }
  XCQ_HIP(hipSetDevice(device));
  auto m = std::make_unique<xcq_mlp_s>();
  m->device = device;
  m->nin = nin;
  m->nh = nhidden;
  m->nl = nlayers;
  m->nout = nout;
  m->widths.push_back(nin);
  for (int l = 0; l < nlayers; ++l) m->widths.push_back(nhidden);
  m->widths.push_back(nout);
  const int L = (int)m->widths.size() - 1;
  for (int l = 1; l <= L; ++l) {
    m->w_off.push_back(m->nparam);
    m->nparam += (int64_t)m->widths[l] * m->widths[l - 1];
    m->b_off.push_back(m->nparam);
    m->nparam += m->widths[l];
  }
  // He-uniform initialisation of the weights, zero biases.
  std::vector<float> h((size_t)m->nparam, 0.f);
  std::mt19937_64 rng(seed);
  for (int l = 1; l <= L; ++l) {
    const float a = std::sqrt(6.f / (float)m->widths[l - 1]);
    std::uniform_real_distribution<float> U(-a, a);
    for (int64_t k = 0; k < (int64_t)m->widths[l] * m->widths[l - 1]; ++k)
      h[m->w_off[l - 1] + k] = U(rng);
  }
  XCQ_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
  for (float** q : {&m->d_p, &m->d_g, &m->d_m, &m->d_v})
    XCQ_HIP(hipMalloc(q, sizeof(float) * m->nparam));
  XCQ_HIP(hipMemcpy(m->d_p, h.data(), sizeof(float) * m->nparam, hipMemcpyHostToDevice));
  XCQ_HIP(hipMemset(m->d_m, 0, sizeof(float) * m->nparam));
  XCQ_HIP(hipMemset(m->d_v, 0, sizeof(float) * m->nparam));
  *out = m.release();
  return XCQ_SUCCESS;
}

int xcq_mlp_destroy(xcq_mlp m) {
  delete m;
  return XCQ_SUCCESS;
}

int xcq_mlp_num_params(xcq_mlp m, int64_t* n) {
  XCQ_REQUIRE(m && n, XCQ_ERROR_INVALID_ARGUMENT, "null argument");
  *n = m->nparam;
  return XCQ_SUCCESS;
}

int xcq_mlp_get_params(xcq_mlp m, float* host) {
  XCQ_REQUIRE(m && host, XCQ_ERROR_INVALID_ARGUMENT, "null argument");
  XCQ_HIP(hipSetDevice(m->device));
  XCQ_HIP(hipStreamSynchronize(m->stream));
  XCQ_HIP(hipMemcpy(host, m->d_p, sizeof(float) * m->nparam, hipMemcpyDeviceToHost));
  return XCQ_SUCCESS;
}

int xcq_mlp_set_params(xcq_mlp m, const float* host) {
  XCQ_REQUIRE(m && host, XCQ_ERROR_INVALID_ARGUMENT, "null argument");
  XCQ_HIP(hipSetDevice(m->device));
  XCQ_HIP(hipStreamSynchronize(m->stream));
  XCQ_HIP(hipMemcpy(m->d_p, host, sizeof(float) * m->nparam, hipMemcpyHostToDevice));
  XCQ_HIP(hipMemset(m->d_m, 0, sizeof(float) * m->nparam));
  XCQ_HIP(hipMemset(m->d_v, 0, sizeof(float) * m->nparam));
  m->adam_t = 0;
  return XCQ_SUCCESS;
}

int xcq_mlp_forward(xcq_mlp m, int64_t n, const float* x, float* y) {
  XCQ_REQUIRE(m && (n == 0 || (x && y)), XCQ_ERROR_INVALID_ARGUMENT, "null argument");
  XCQ_REQUIRE(n >= 0 && n < (1ll << 31), XCQ_ERROR_INVALID_ARGUMENT, "row count out of range");
  if (n == 0) return XCQ_SUCCESS;
  XCQ_HIP(hipSetDevice(m->device));
  int rc = m->ensure(n);
  if (rc) return rc;
  XCQ_HIP(hipMemcpyAsync(m->A[0], x, sizeof(float) * n * m->nin, hipMemcpyHostToDevice,
                         m->stream));
  if ((rc = forward(m, n))) return rc;
  const int L = (int)m->widths.size() - 1;
  XCQ_HIP(hipMemcpyAsync(y, m->A[L], sizeof(float) * n * m->nout, hipMemcpyDeviceToHost,
                         m->stream));
  XCQ_HIP(hipStreamSynchronize(m->stream));
  return XCQ_SUCCESS;
}

int xcq_mlp_gradients(xcq_mlp m, int64_t n, const float* x, const float* t, float* loss,
                      float* grad) {
  XCQ_REQUIRE(m && n > 0 && x && t, XCQ_ERROR_INVALID_ARGUMENT, "null argument");
  XCQ_REQUIRE(n < (1ll << 31), XCQ_ERROR_INVALID_ARGUMENT, "row count out of range");
  XCQ_HIP(hipSetDevice(m->device));
  int rc = m->ensure(n);
  if (rc) return rc;
  XCQ_HIP(hipMemcpyAsync(m->A[0], x, sizeof(float) * n * m->nin, hipMemcpyHostToDevice,
                         m->stream));
  XCQ_HIP(hipMemcpyAsync(m->d_T, t, sizeof(float) * n * m->nout, hipMemcpyHostToDevice,
                         m->stream));
  if ((rc = forward(m, n))) return rc;
  if ((rc = backward(m, n, loss))) return rc;
  if (grad)
    XCQ_HIP(hipMemcpyAsync(grad, m->d_g, sizeof(float) * m->nparam, hipMemcpyDeviceToHost,
                           m->stream));
  XCQ_HIP(hipStreamSynchronize(m->stream));
  return XCQ_SUCCESS;
}

int xcq_mlp_train(xcq_mlp m, int64_t n, const float* x, const float* t, int32_t epochs,
                  int32_t batch, float lr, uint64_t seed, float* final_loss) {
  XCQ_REQUIRE(m && n > 0 && x && t, XCQ_ERROR_INVALID_ARGUMENT, "null argument");
  XCQ_REQUIRE(epochs >= 0 && batch > 0 && lr > 0.f, XCQ_ERROR_INVALID_ARGUMENT,
              "bad training settings");
  XCQ_REQUIRE(n < (1ll << 31), XCQ_ERROR_INVALID_ARGUMENT, "row count out of range");
  XCQ_HIP(hipSetDevice(m->device));
  const int64_t bsz = std::min<int64_t>(batch, n);
  int rc = m->ensure(bsz);
  if (rc) return rc;
  if (n > m->data_cap) {
    for (float** q : {&m->d_X, &m->d_Y})
      if (*q) {
        (void)hipFree(*q);
        *q = nullptr;
      }
    if (m->d_idx) (void)hipFree(m->d_idx);
    m->d_idx = nullptr;
    XCQ_HIP(hipMalloc(&m->d_X, sizeof(float) * n * m->nin));
    XCQ_HIP(hipMalloc(&m->d_Y, sizeof(float) * n * m->nout));
    XCQ_HIP(hipMalloc(&m->d_idx, sizeof(int32_t) * n));
    m->data_cap = n;
  }
  XCQ_HIP(hipMemcpyAsync(m->d_X, x, sizeof(float) * n * m->nin, hipMemcpyHostToDevice,
                         m->stream));
  XCQ_HIP(hipMemcpyAsync(m->d_Y, t, sizeof(float) * n * m->nout, hipMemcpyHostToDevice,
                         m->stream));
  std::vector<int32_t> perm(n);
  for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
  std::mt19937_64 rng(seed);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float last = 0.f;
  for (int e = 0; e < epochs; ++e) {
    std::shuffle(perm.begin(), perm.end(), rng);
    XCQ_HIP(hipMemcpyAsync(m->d_idx, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice,
                           m->stream));
    for (int64_t b0 = 0; b0 + bsz <= n; b0 += bsz) {
      hipLaunchKernelGGL(k_gather_rows, dim3(nblk(bsz * m->nin)), dim3(256), 0, m->stream, bsz,
                         m->nin, m->d_idx + b0, m->d_X, m->A[0]);
      hipLaunchKernelGGL(k_gather_rows, dim3(nblk(bsz * m->nout)), dim3(256), 0, m->stream, bsz,
                         m->nout, m->d_idx + b0, m->d_Y, m->d_T);
      XCQ_HIP(hipGetLastError());
      if ((rc = forward(m, bsz))) return rc;
      const bool last_step = e == epochs - 1 && b0 + 2 * bsz > n;
      if ((rc = backward(m, bsz, last_step ? &last : nullptr))) return rc;
      ++m->adam_t;
      const float c1 = 1.f - std::pow(b1, (float)m->adam_t);
      const float c2 = 1.f - std::pow(b2, (float)m->adam_t);
      hipLaunchKernelGGL(k_adam, dim3(nblk(m->nparam)), dim3(256), 0, m->stream, m->nparam,
                         m->d_p, m->d_g, m->d_m, m->d_v, lr, b1, b2, c1, c2, eps);
      XCQ_HIP(hipGetLastError());
    }
  }
  XCQ_HIP(hipStreamSynchronize(m->stream));
  if (final_loss) *final_loss = last;
  return XCQ_SUCCESS;
}

}  // extern "C"
