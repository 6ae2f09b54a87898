// Learned-prefetch MLP on gfx950: FP32 GEMMs on the matrix cores
// (v_mfma_f32_16x16x4f32), fused bias/ReLU epilogues, binary cross-entropy,
// Adam, and bit-mask input expansion, driven from the C ABI in
// include/fcs/fcs_mlp.h. No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fcs/fcs_mlp.h"
#include "fcs_internal.h"

typedef float fcs_f4 __attribute__((ext_vector_type(4)));

namespace {

// ------------------------------------------------------------------ GEMM
// C[M x N] = op(A)[M x K] * op(B)[K x N], row-major with leading dimensions.
// op(A)(m, k) = TA ? A[k * lda + m] : A[m * lda + k]
// op(B)(k, n) = TB ? B[n * ldb + k] : B[k * ldb + n]
// A 256-thread workgroup computes a 64 x 64 tile; each wave a 32 x 32 quarter
// as 2 x 2 MFMA 16x16x4 tiles. K advances 16 at a time through LDS. With
// split-K (gridDim.z > 1) each workgroup adds its partial product into C
// atomically (C zeroed beforehand).
constexpr int TM = 64, TN = 64, TK = 16;

template <bool TA, bool TB, bool ATOMIC>
__global__ __launch_bounds__(256) void gemm_kernel(const float* __restrict__ A, const float* __restrict__ B,
                                                   float* __restrict__ C, int M, int N, int K, int lda, int ldb,
                                                   int ldc, int k_split) {
  __shared__ float As[TK][TM + 4];
  __shared__ float Bs[TK][TN + 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
  const int kb = blockIdx.z * k_split, ke = min(K, kb + k_split);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  fcs_f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = fcs_f4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += TK) {
#pragma unroll
    for (int e = tid; e < TM * TK; e += 256) {
      int mm, kk;
      if (TA) { mm = e % TM; kk = e / TM; } else { kk = e % TK; mm = e / TK; }
      const int gm = m0 + mm, gk = k0 + kk;
      float v = 0.f;
      if (gm < M && gk < ke) v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
      As[kk][mm] = v;
    }
#pragma unroll
    for (int e = tid; e < TN * TK; e += 256) {
      int nn, kk;
      if (TB) { kk = e % TK; nn = e / TK; } else { nn = e % TN; kk = e / TN; }
      const int gn = n0 + nn, gk = k0 + kk;
      float v = 0.f;
      if (gn < N && gk < ke) v = TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
      Bs[kk][nn] = v;
    }
    __syncthreads();
    // MFMA 16x16x4: lane l supplies A(l % 16, l / 16) and B(l / 16, l % 16)
#pragma unroll
    for (int ks = 0; ks < TK; ks += 4) {
      const int kl = ks + (lane >> 4);
      float a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[kl][wm + 16 * i + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[kl][wn + 16 * j + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // result: lane l holds rows 4 * (l / 16) + r, column l % 16 of each tile
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
        const int gn = n0 + wn + 16 * j + (lane & 15);
        if (gm < M && gn < N) {
          if (ATOMIC) atomicAdd(&C[(size_t)gm * ldc + gn], acc[i][j][r]);
          else C[(size_t)gm * ldc + gn] = acc[i][j][r];
        }
      }
}

__global__ void zero_f32_kernel(float* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// C[m][n] = act(C[m][n] + bias[n])
__global__ void bias_act_kernel(float* C, const float* __restrict__ bias, int M, int N, int relu) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)M * N) return;
  float v = C[i] + bias[i % N];
  C[i] = (relu && v < 0.f) ? 0.f : v;
}

// dH *= (H > 0)
__global__ void relu_grad_kernel(float* dH, const float* __restrict__ H, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(H[i] > 0.f)) dH[i] = 0.f;
}

// out[n] = sum over m of D[m][n]
__global__ void colsum_kernel(const float* __restrict__ D, float* out, int M, int N) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int m = 0; m < M; ++m) s += D[(size_t)m * N + n];
  out[n] = s;
}

// dZ = (sigmoid(Z) - Y) * scale; loss += sum of max(z,0) - z*y + log1p(exp(-|z|))
__global__ void bce_kernel(const float* __restrict__ Z, const float* __restrict__ Y, float* dZ, double* loss,
                           size_t n, float scale) {
  __shared__ double red[256];
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (i < n) {
    const float z = Z[i], y = Y[i];
    const float p = 1.f / (1.f + expf(-z));
    dZ[i] = (p - y) * scale;
    l = (double)(fmaxf(z, 0.f) - z * y + log1pf(expf(-fabsf(z))));
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

__global__ void adam_kernel(float* p, const float* __restrict__ g, float* m, float* v, size_t n, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// X[b] = [poses[idx[b]], bits of in_bits row idx[b]]; Y[b] = bits of tgt row
__global__ void expand_kernel(const float* __restrict__ poses, const uint32_t* __restrict__ bits,
                              const uint32_t* __restrict__ idx, float* X, int B, int n_pose, int n_cols, int wpr) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)B * n_cols) return;
  const int b = (int)(i / n_cols), c = (int)(i % n_cols);
  const size_t s = idx[b];
  float v;
  if (c < n_pose) {
    v = poses[s * n_pose + c];
  } else {
    const int q = c - n_pose;
    v = (float)((bits[s * wpr + (q >> 5)] >> (q & 31)) & 1u);
  }
  X[i] = v;
}

// X[0] = [pose, mask != 0]
__global__ void mask_input_kernel(const float* __restrict__ pose, const uint32_t* __restrict__ mask, float* X,
                                  int n_pose, int n_cols) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cols) return;
  X[c] = c < n_pose ? pose[c] : (mask[c - n_pose] ? 1.f : 0.f);
}

__global__ void sigmoid_kernel(const float* __restrict__ Z, float* P, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) P[i] = 1.f / (1.f + expf(-Z[i]));
}

inline unsigned nblk(size_t n, unsigned t = 256) { return (unsigned)((n + t - 1) / t); }

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace

struct fcs_mlp {
  fcs_ctx* ctx = nullptr;
  uint32_t n_in = 0, h = 0, n_out = 0, max_batch = 0;
  uint64_t np = 0;
  size_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0;
  float* d_p = nullptr;   // parameters
  float* d_g = nullptr;   // gradients
  float* d_m = nullptr;   // Adam moments
  float* d_v = nullptr;
  int step = 0;
  float* d_x = nullptr;   // max_batch x n_in
  float* d_y = nullptr;   // max_batch x n_out
  float* d_h1 = nullptr;  // max_batch x h
  float* d_h2 = nullptr;
  float* d_z = nullptr;   // max_batch x n_out (logits; probabilities after predict)
  float* d_dz = nullptr;
  float* d_dh2 = nullptr;
  float* d_dh1 = nullptr;
  float* d_pose = nullptr;  // staging for one pose
  double* d_loss = nullptr;
};

namespace {

template <bool TA, bool TB>
hipError_t gemm(fcs_mlp* m, const float* A, const float* B, float* C, int M, int N, int K, int lda, int ldb,
                int ldc) {
  hipStream_t s = m->ctx->stream;
  const int gx = (N + TN - 1) / TN, gy = (M + TM - 1) / TM;
  const int tiles = gx * gy;
  const int want = 2 * m->ctx->num_cu;
  int splits = 1;
  if (tiles < want && K >= 4 * TK * 4) splits = std::min((want + tiles - 1) / tiles, K / (4 * TK));
  int k_split = (K + splits - 1) / splits;
  k_split = (k_split + TK - 1) / TK * TK;
  splits = (K + k_split - 1) / k_split;
  if (splits <= 1) {
    hipLaunchKernelGGL((gemm_kernel<TA, TB, false>), dim3(gx, gy, 1), dim3(256), 0, s, A, B, C, M, N, K, lda, ldb,
                       ldc, K);
  } else {
    hipLaunchKernelGGL(zero_f32_kernel, dim3(nblk((size_t)M * ldc)), dim3(256), 0, s, C, (size_t)M * ldc);
    hipLaunchKernelGGL((gemm_kernel<TA, TB, true>), dim3(gx, gy, splits), dim3(256), 0, s, A, B, C, M, N, K, lda,
                       ldb, ldc, k_split);
  }
  return hipGetLastError();
}

// logits for the batch in d_x
int forward_dev(fcs_mlp* m, uint32_t B) {
  fcs_ctx* c = m->ctx;
  hipStream_t s = c->stream;
  const int n_in = m->n_in, h = m->h, n_out = m->n_out;
  FCS_HIP_TRY(c, (gemm<false, false>(m, m->d_x, m->d_p + m->oW1, m->d_h1, B, h, n_in, n_in, h, h)));
  hipLaunchKernelGGL(bias_act_kernel, dim3(nblk((size_t)B * h)), dim3(256), 0, s, m->d_h1, m->d_p + m->ob1, (int)B,
                     h, 1);
  FCS_HIP_TRY(c, (gemm<false, false>(m, m->d_h1, m->d_p + m->oW2, m->d_h2, B, h, h, h, h, h)));
  hipLaunchKernelGGL(bias_act_kernel, dim3(nblk((size_t)B * h)), dim3(256), 0, s, m->d_h2, m->d_p + m->ob2, (int)B,
                     h, 1);
  FCS_HIP_TRY(c, (gemm<false, false>(m, m->d_h2, m->d_p + m->oW3, m->d_z, B, n_out, h, h, n_out, n_out)));
  hipLaunchKernelGGL(bias_act_kernel, dim3(nblk((size_t)B * n_out)), dim3(256), 0, s, m->d_z, m->d_p + m->ob3,
                     (int)B, n_out, 0);
  FCS_HIP_TRY(c, hipGetLastError());
  return 0;
}

// loss (accumulated into d_loss) and gradients for the batch in d_x / d_y
int backward_dev(fcs_mlp* m, uint32_t B) {
  fcs_ctx* c = m->ctx;
  hipStream_t s = c->stream;
  const int n_in = m->n_in, h = m->h, n_out = m->n_out;
  if (int rc = forward_dev(m, B)) return rc;
  const size_t nz = (size_t)B * n_out;
  hipLaunchKernelGGL(bce_kernel, dim3(nblk(nz)), dim3(256), 0, s, m->d_z, m->d_y, m->d_dz, m->d_loss, nz,
                     1.f / (float)nz);
  float* g = m->d_g;
  // output layer
  FCS_HIP_TRY(c, (gemm<true, false>(m, m->d_h2, m->d_dz, g + m->oW3, h, n_out, B, h, n_out, n_out)));
  hipLaunchKernelGGL(colsum_kernel, dim3(nblk(n_out)), dim3(256), 0, s, m->d_dz, g + m->ob3, (int)B, n_out);
  FCS_HIP_TRY(c, (gemm<false, true>(m, m->d_dz, m->d_p + m->oW3, m->d_dh2, B, h, n_out, n_out, n_out, h)));
  hipLaunchKernelGGL(relu_grad_kernel, dim3(nblk((size_t)B * h)), dim3(256), 0, s, m->d_dh2, m->d_h2, (size_t)B * h);
  // hidden layer 2
  FCS_HIP_TRY(c, (gemm<true, false>(m, m->d_h1, m->d_dh2, g + m->oW2, h, h, B, h, h, h)));
  hipLaunchKernelGGL(colsum_kernel, dim3(nblk(h)), dim3(256), 0, s, m->d_dh2, g + m->ob2, (int)B, h);
  FCS_HIP_TRY(c, (gemm<false, true>(m, m->d_dh2, m->d_p + m->oW2, m->d_dh1, B, h, h, h, h, h)));
  hipLaunchKernelGGL(relu_grad_kernel, dim3(nblk((size_t)B * h)), dim3(256), 0, s, m->d_dh1, m->d_h1, (size_t)B * h);
  // hidden layer 1
  FCS_HIP_TRY(c, (gemm<true, false>(m, m->d_x, m->d_dh1, g + m->oW1, n_in, h, B, n_in, h, h)));
  hipLaunchKernelGGL(colsum_kernel, dim3(nblk(h)), dim3(256), 0, s, m->d_dh1, g + m->ob1, (int)B, h);
  FCS_HIP_TRY(c, hipGetLastError());
  return 0;
}

int adam_dev(fcs_mlp* m, float lr, float b1, float b2, float eps) {
  ++m->step;
  const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
  hipLaunchKernelGGL(adam_kernel, dim3(nblk(m->np)), dim3(256), 0, m->ctx->stream, m->d_p, m->d_g, m->d_m, m->d_v,
                     (size_t)m->np, lr, b1, b2, eps, c1, c2);
  FCS_HIP_TRY(m->ctx, hipGetLastError());
  return 0;
}

int zero_loss(fcs_mlp* m) {
  const double z = 0.0;
  FCS_HIP_TRY(m->ctx, hipMemcpyAsync(m->d_loss, &z, sizeof z, hipMemcpyHostToDevice, m->ctx->stream));
  FCS_HIP_TRY(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return 0;
}

}  // namespace

// Device-side prediction for the prefetcher: probabilities of one sample land
// in a device buffer owned by the network (valid until its next call).
int fcs_mlp_predict_dev(fcs_mlp* m, const float* pose, uint32_t n_pose, const uint32_t* d_mask, const float** d_probs) {
  if (!m || !m->ctx || !pose || !d_mask || n_pose > m->n_in) return FCS_E_INVALID;
  fcs_ctx* c = m->ctx;
  FCS_HIP_TRY(c, hipSetDevice(c->device));
  FCS_HIP_TRY(c, hipMemcpyAsync(m->d_pose, pose, n_pose * sizeof(float), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(mask_input_kernel, dim3(nblk(m->n_in)), dim3(256), 0, c->stream, m->d_pose, d_mask, m->d_x,
                     (int)n_pose, (int)m->n_in);
  if (int rc = forward_dev(m, 1)) return rc;
  hipLaunchKernelGGL(sigmoid_kernel, dim3(nblk(m->n_out)), dim3(256), 0, c->stream, m->d_z, m->d_z,
                     (size_t)m->n_out);
  FCS_HIP_TRY(c, hipGetLastError());
  if (d_probs) *d_probs = m->d_z;
  return 0;
}

void fcs_mlp_release_device(fcs_mlp* m) {
  fcs_ctx* c = m->ctx;
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  float* ptrs[] = {m->d_p, m->d_g, m->d_m, m->d_v, m->d_x, m->d_y, m->d_h1, m->d_h2, m->d_z, m->d_dz, m->d_dh2,
                   m->d_dh1, m->d_pose};
  for (float* q : ptrs)
    if (q) (void)hipFree(q);
  if (m->d_loss) (void)hipFree(m->d_loss);
  m->d_p = m->d_g = m->d_m = m->d_v = m->d_x = m->d_y = m->d_h1 = m->d_h2 = m->d_z = m->d_dz = m->d_dh2 =
      m->d_dh1 = m->d_pose = nullptr;
  m->d_loss = nullptr;
  auto& ms = c->mlps;
  ms.erase(std::remove(ms.begin(), ms.end(), m), ms.end());
  m->ctx = nullptr;
}

uint32_t fcs_mlp_outputs(const fcs_mlp* m) { return m ? m->n_out : 0; }
uint32_t fcs_mlp_inputs(const fcs_mlp* m) { return m ? m->n_in : 0; }

extern "C" {

int fcs_mlp_create(fcs_ctx* ctx, uint32_t n_in, uint32_t n_hidden, uint32_t n_out, uint32_t max_batch, uint64_t seed,
                   fcs_mlp** out) {
  if (!ctx || !out || !n_in || !n_hidden || !n_out || !max_batch || max_batch > 4096) return FCS_E_INVALID;
  *out = nullptr;
  FCS_HIP_TRY(ctx, hipSetDevice(ctx->device));
  fcs_mlp* m = new fcs_mlp();
  m->ctx = ctx;
  m->n_in = n_in;
  m->h = n_hidden;
  m->n_out = n_out;
  m->max_batch = max_batch;
  m->oW1 = 0;
  m->ob1 = m->oW1 + (size_t)n_in * n_hidden;
  m->oW2 = m->ob1 + n_hidden;
  m->ob2 = m->oW2 + (size_t)n_hidden * n_hidden;
  m->oW3 = m->ob2 + n_hidden;
  m->ob3 = m->oW3 + (size_t)n_hidden * n_out;
  m->np = m->ob3 + n_out;
  const size_t B = max_batch;
  struct Buf {
    float** p;
    size_t n;
  } bufs[] = {{&m->d_p, m->np}, {&m->d_g, m->np}, {&m->d_m, m->np}, {&m->d_v, m->np},
              {&m->d_x, B * n_in}, {&m->d_y, B * n_out}, {&m->d_h1, B * n_hidden}, {&m->d_h2, B * n_hidden},
              {&m->d_z, B * n_out}, {&m->d_dz, B * n_out}, {&m->d_dh2, B * n_hidden}, {&m->d_dh1, B * n_hidden},
              {&m->d_pose, n_in}};
  for (auto& b : bufs) {
    if (hipMalloc((void**)b.p, b.n * sizeof(float)) != hipSuccess) {
      fcs_mlp_destroy(m);
      return fcs_fail(ctx, FCS_E_NOMEM, "fcs_mlp_create: device allocation failed");
    }
  }
  if (hipMalloc((void**)&m->d_loss, sizeof(double)) != hipSuccess) {
    fcs_mlp_destroy(m);
    return fcs_fail(ctx, FCS_E_NOMEM, "fcs_mlp_create: device allocation failed");
  }
  // He-uniform weights, zero biases
  std::vector<float> p(m->np, 0.f);
  uint64_t st = seed;
  auto fill = [&](size_t off, size_t n, uint32_t fan_in) {
    const float a = std::sqrt(6.f / (float)fan_in);
    for (size_t i = 0; i < n; ++i) {
      const double u = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
      p[off + i] = (float)((2.0 * u - 1.0) * a);
    }
  };
  fill(m->oW1, (size_t)n_in * n_hidden, n_in);
  fill(m->oW2, (size_t)n_hidden * n_hidden, n_hidden);
  fill(m->oW3, (size_t)n_hidden * n_out, n_hidden);
  if (int rc = fcs_mlp_set_params(m, p.data())) {
    fcs_mlp_destroy(m);
    return rc;
  }
  std::vector<float> zero(m->np, 0.f);
  if (hipMemcpy(m->d_m, zero.data(), m->np * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->d_v, zero.data(), m->np * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->d_g, zero.data(), m->np * 4, hipMemcpyHostToDevice) != hipSuccess) {
    fcs_mlp_destroy(m);
    return fcs_fail(ctx, FCS_E_HIP, "fcs_mlp_create: upload failed");
  }
  ctx->mlps.push_back(m);
  *out = m;
  return 0;
}

void fcs_mlp_destroy(fcs_mlp* m) {
  if (!m) return;
  fcs_mlp_release_device(m);
  delete m;
}

uint64_t fcs_mlp_num_params(const fcs_mlp* m) { return m ? m->np : 0; }

int fcs_mlp_get_params(fcs_mlp* m, float* host) {
  if (!m || !m->ctx || !host) return FCS_E_INVALID;
  FCS_HIP_TRY(m->ctx, hipStreamSynchronize(m->ctx->stream));
  FCS_HIP_TRY(m->ctx, hipMemcpy(host, m->d_p, m->np * 4, hipMemcpyDeviceToHost));
  return 0;
}

int fcs_mlp_set_params(fcs_mlp* m, const float* host) {
  if (!m || !m->ctx || !host) return FCS_E_INVALID;
  FCS_HIP_TRY(m->ctx, hipStreamSynchronize(m->ctx->stream));
  FCS_HIP_TRY(m->ctx, hipMemcpy(m->d_p, host, m->np * 4, hipMemcpyHostToDevice));
  return 0;
}

int fcs_mlp_get_grads(fcs_mlp* m, float* host) {
  if (!m || !m->ctx || !host) return FCS_E_INVALID;
  FCS_HIP_TRY(m->ctx, hipStreamSynchronize(m->ctx->stream));
  FCS_HIP_TRY(m->ctx, hipMemcpy(host, m->d_g, m->np * 4, hipMemcpyDeviceToHost));
  return 0;
}

int fcs_mlp_forward(fcs_mlp* m, const float* x, uint32_t batch, float* logits) {
  if (!m || !m->ctx || !x || batch == 0 || batch > m->max_batch) return FCS_E_INVALID;
  fcs_ctx* c = m->ctx;
  FCS_HIP_TRY(c, hipSetDevice(c->device));
  FCS_HIP_TRY(c, hipMemcpyAsync(m->d_x, x, (size_t)batch * m->n_in * 4, hipMemcpyHostToDevice, c->stream));
  if (int rc = forward_dev(m, batch)) return rc;
  if (logits)
    FCS_HIP_TRY(c, hipMemcpyAsync(logits, m->d_z, (size_t)batch * m->n_out * 4, hipMemcpyDeviceToHost, c->stream));
  FCS_HIP_TRY(c, hipStreamSynchronize(c->stream));
  return 0;
}

int fcs_mlp_backward(fcs_mlp* m, const float* x, const float* y, uint32_t batch, float* loss) {
  if (!m || !m->ctx || !x || !y || batch == 0 || batch > m->max_batch) return FCS_E_INVALID;
  fcs_ctx* c = m->ctx;
  FCS_HIP_TRY(c, hipSetDevice(c->device));
  if (int rc = zero_loss(m)) return rc;
  FCS_HIP_TRY(c, hipMemcpyAsync(m->d_x, x, (size_t)batch * m->n_in * 4, hipMemcpyHostToDevice, c->stream));
  FCS_HIP_TRY(c, hipMemcpyAsync(m->d_y, y, (size_t)batch * m->n_out * 4, hipMemcpyHostToDevice, c->stream));
  if (int rc = backward_dev(m, batch)) return rc;
  double l = 0;
  FCS_HIP_TRY(c, hipMemcpyAsync(&l, m->d_loss, sizeof l, hipMemcpyDeviceToHost, c->stream));
  FCS_HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (loss) *loss = (float)(l / ((double)batch * m->n_out));
  return 0;
}

int fcs_mlp_adam(fcs_mlp* m, float lr, float beta1, float beta2, float eps) {
  if (!m || !m->ctx) return FCS_E_INVALID;
  FCS_HIP_TRY(m->ctx, hipSetDevice(m->ctx->device));
  if (int rc = adam_dev(m, lr, beta1, beta2, eps)) return rc;
  FCS_HIP_TRY(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return 0;
}

int fcs_mlp_train_masks(fcs_mlp* m, const float* poses, uint32_t n_pose, const uint32_t* in_bits,
                        const uint32_t* tgt_bits, uint32_t n_samples, uint32_t epochs, uint32_t batch, float lr,
                        uint64_t seed, float* epoch_loss) {
  if (!m || !m->ctx || !poses || !in_bits || !tgt_bits || n_samples == 0 || batch == 0 || batch > m->max_batch ||
      n_pose > m->n_in)
    return FCS_E_INVALID;
  fcs_ctx* c = m->ctx;
  FCS_HIP_TRY(c, hipSetDevice(c->device));
  const uint32_t n_mask = m->n_in - n_pose;
  const uint32_t wpr_in = (n_mask + 31) / 32, wpr_out = (m->n_out + 31) / 32;
  float* d_poses = nullptr;
  uint32_t *d_in = nullptr, *d_tgt = nullptr, *d_perm = nullptr;
  auto cleanup = [&]() {
    if (d_poses) (void)hipFree(d_poses);
    if (d_in) (void)hipFree(d_in);
    if (d_tgt) (void)hipFree(d_tgt);
    if (d_perm) (void)hipFree(d_perm);
  };
  if (hipMalloc((void**)&d_poses, (size_t)n_samples * std::max(n_pose, 1u) * 4) != hipSuccess ||
      hipMalloc((void**)&d_in, (size_t)n_samples * wpr_in * 4 + 4) != hipSuccess ||
      hipMalloc((void**)&d_tgt, (size_t)n_samples * wpr_out * 4) != hipSuccess ||
      hipMalloc((void**)&d_perm, (size_t)n_samples * 4) != hipSuccess) {
    cleanup();
    return fcs_fail(c, FCS_E_NOMEM, "fcs_mlp_train_masks: device allocation failed");
  }
  int rc = 0;
  do {
    if (n_pose && hipMemcpy(d_poses, poses, (size_t)n_samples * n_pose * 4, hipMemcpyHostToDevice) != hipSuccess) {
      rc = fcs_fail(c, FCS_E_HIP, "upload poses");
      break;
    }
    if (hipMemcpy(d_in, in_bits, (size_t)n_samples * wpr_in * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_tgt, tgt_bits, (size_t)n_samples * wpr_out * 4, hipMemcpyHostToDevice) != hipSuccess) {
      rc = fcs_fail(c, FCS_E_HIP, "upload masks");
      break;
    }
    std::vector<uint32_t> perm(n_samples);
    uint64_t st = seed;
    for (uint32_t e = 0; e < epochs && rc == 0; ++e) {
      for (uint32_t i = 0; i < n_samples; ++i) perm[i] = i;
      for (uint32_t i = n_samples - 1; i > 0; --i) std::swap(perm[i], perm[splitmix(st) % (i + 1)]);
      if (hipMemcpy(d_perm, perm.data(), (size_t)n_samples * 4, hipMemcpyHostToDevice) != hipSuccess) {
        rc = fcs_fail(c, FCS_E_HIP, "upload order");
        break;
      }
      if ((rc = zero_loss(m))) break;
      double seen = 0;
      for (uint32_t s0 = 0; s0 < n_samples && rc == 0; s0 += batch) {
        const uint32_t B = std::min(batch, n_samples - s0);
        hipLaunchKernelGGL(expand_kernel, dim3(nblk((size_t)B * m->n_in)), dim3(256), 0, c->stream, d_poses, d_in,
                           d_perm + s0, m->d_x, (int)B, (int)n_pose, (int)m->n_in, (int)wpr_in);
        hipLaunchKernelGGL(expand_kernel, dim3(nblk((size_t)B * m->n_out)), dim3(256), 0, c->stream, d_poses,
                           d_tgt, d_perm + s0, m->d_y, (int)B, 0, (int)m->n_out, (int)wpr_out);
        if ((rc = backward_dev(m, B))) break;
        if ((rc = adam_dev(m, lr, 0.9f, 0.999f, 1e-8f))) break;
        seen += (double)B * m->n_out;
      }
      if (rc) break;
      double l = 0;
      if (hipMemcpyAsync(&l, m->d_loss, sizeof l, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
          hipStreamSynchronize(c->stream) != hipSuccess) {
        rc = fcs_fail(c, FCS_E_HIP, "training step failed");
        break;
      }
      // mean over every (sample, output) element of the epoch
      if (epoch_loss) epoch_loss[e] = (float)(l / seen);
    }
  } while (0);
  (void)hipStreamSynchronize(c->stream);
  cleanup();
  return rc;
}

int fcs_mlp_predict_device_mask(fcs_mlp* m, const float* pose, uint32_t n_pose, const uint32_t* d_mask_u32,
                                float* probs) {
  const float* d = nullptr;
  if (int rc = fcs_mlp_predict_dev(m, pose, n_pose, d_mask_u32, &d)) return rc;
  if (probs) FCS_HIP_TRY(m->ctx, hipMemcpyAsync(probs, d, (size_t)m->n_out * 4, hipMemcpyDeviceToHost, m->ctx->stream));
  FCS_HIP_TRY(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return 0;
}

}  // extern "C"
