// Learned warm start on MI355X: the kernels of a small MLP (features -> 256 -> 256 -> per-spin logits, GELU) with
// BF16 operands and FP32 accumulation on the matrix cores, its backward pass, BCE loss, Adam, and the conversion of
// logits into per-spin start thresholds for the annealer. The reference has no learned start; the start it falls back
// to is its own random start (arpeggio/sim/arpeggio_fx.cu:48-51).
//
// GEMM: C(m, n) = sum_k A(m, k) B(k, n) with A(m, k) = A[m*sam + k*sak] and B(k, n) = B[k*sbk + n*sbn] (bf16, any
// strides, so X W, dY W^T and X^T dY all use one kernel). A workgroup (4 waves) computes a 64 x 64 tile of C; the
// 64 x 32 slice of A and the 32 x 64 slice of B are staged in LDS (B stored n-major so that each lane's fragment of
// 8 consecutive k is one 16-byte read), and each wave accumulates a 16 x 64 strip with v_mfma_f32_16x16x32_bf16:
// lane l feeds A[i = l%16][k = 8(l/16) .. +7] and B[k = 8(l/16) .. +7][j = l%16], and holds C[4(l/16) + v][l%16].
// Epilogue: + bias[n], optional pre-activation store, optional GELU (tanh form), fp32 and/or bf16 store.
#include "kernels.h"

namespace arp {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even (inputs are finite)
  const uint32_t u = __builtin_bit_cast(uint32_t, f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }

// GELU, tanh form: 0.5 x (1 + tanh(c (x + 0.044715 x^3))), c = sqrt(2/pi)
__device__ __forceinline__ float gelu_f(float x) {
  const float c = 0.7978845608028654f;
  return 0.5f * x * (1.f + tanhf(c * (x + 0.044715f * x * x * x)));
}

__device__ __forceinline__ float gelu_grad(float x) {
  const float c = 0.7978845608028654f;
  const float t = tanhf(c * (x + 0.044715f * x * x * x));
  return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * c * (1.f + 3.f * 0.044715f * x * x);
}

__global__ __launch_bounds__(256) void gemm_bf16_kernel(const uint16_t* __restrict__ A, long sam, long sak,
                                                        const uint16_t* __restrict__ B, long sbk, long sbn, int M,
                                                        int N, int K, const float* __restrict__ bias, int act,
                                                        float* __restrict__ Cf, float* __restrict__ Cpre,
                                                        uint16_t* __restrict__ Cb, long ldc) {
  __shared__ __align__(16) uint16_t As[64 * 32];  // [m][k]
  __shared__ __align__(16) uint16_t Bs[64 * 32];  // [n][k]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
  floatx4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 32) {
    // stage A (64 x 32): 256 threads x 8 elements; walk the unit-stride dimension with consecutive threads
    if (sak == 1) {
      const int m = tid >> 2, kc = (tid & 3) * 8;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int gm = m0 + m, gk = k0 + kc + e;
        As[m * 32 + kc + e] = (gm < M && gk < K) ? A[gm * sam + gk] : (uint16_t)0;
      }
    } else {
      const int k = tid >> 3, mc = (tid & 7) * 8;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int gm = m0 + mc + e, gk = k0 + k;
        As[(mc + e) * 32 + k] = (gm < M && gk < K) ? A[gm * sam + gk * sak] : (uint16_t)0;
      }
    }
    if (sbk == 1) {
      const int n = tid >> 2, kc = (tid & 3) * 8;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int gn = n0 + n, gk = k0 + kc + e;
        Bs[n * 32 + kc + e] = (gn < N && gk < K) ? B[gk + gn * sbn] : (uint16_t)0;
      }
    } else {
      const int k = tid >> 3, nc = (tid & 7) * 8;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int gn = n0 + nc + e, gk = k0 + k;
        Bs[(nc + e) * 32 + k] = (gn < N && gk < K) ? B[gk * sbk + gn * sbn] : (uint16_t)0;
      }
    }
    __syncthreads();
    const int li = lane & 15, lg = lane >> 4;
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(As + (wv * 16 + li) * 32 + lg * 8);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const bf16x8 bb = *reinterpret_cast<const bf16x8*>(Bs + (t * 16 + li) * 32 + lg * 8);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc[t], 0, 0, 0);
    }
    __syncthreads();
  }
  const int li = lane & 15, lg = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int n = n0 + t * 16 + li;
    if (n >= N) continue;
    const float bn = bias ? bias[n] : 0.f;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int m = m0 + wv * 16 + lg * 4 + v;
      if (m >= M) continue;
      const float x = acc[t][v] + bn;
      const size_t o = (size_t)m * ldc + n;
      if (Cpre) Cpre[o] = x;
      const float yv = act == 1 ? gelu_f(x) : x;
      if (Cf) Cf[o] = yv;
      if (Cb) Cb[o] = f2bf(yv);
    }
  }
}

// dA = dH * gelu'(Apre) (fp32 in), bf16 and/or fp32 out; n elements
__global__ void gelu_bwd_kernel(const float* __restrict__ dH, const float* __restrict__ Apre, long n,
                                uint16_t* __restrict__ dAb, float* __restrict__ dAf) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float d = dH[i] * gelu_grad(Apre[i]);
  if (dAb) dAb[i] = f2bf(d);
  if (dAf) dAf[i] = d;
}

// Binary cross-entropy on logits z (M x N, row stride ld) against spins y in {-1, +1} (int8, M x N contiguous):
// loss = mean over elements of softplus(z) - t z, t = (y + 1)/2; dz = (sigmoid(z) - t) / (M N) as bf16 and fp32.
// loss_acc gets one partial sum per block (atomic add).
__global__ void bce_kernel(const float* __restrict__ z, long ld, const int8_t* __restrict__ y, int M, int N,
                           uint16_t* __restrict__ dzb, float* __restrict__ dzf, float* __restrict__ loss_acc) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long tot = (long)M * N;
  float l = 0.f;
  if (i < tot) {
    const int m = (int)(i / N), n = (int)(i - (long)m * N);
    const float x = z[m * ld + n];
    const float t = y[i] > 0 ? 1.f : 0.f;
    const float sp = x > 0.f ? x + log1pf(expf(-x)) : log1pf(expf(x));
    l = sp - t * x;
    const float sg = 1.f / (1.f + expf(-x));
    const float g = (sg - t) / (float)tot;
    if (dzb) dzb[i] = f2bf(g);
    if (dzf) dzf[i] = g;
  }
  for (int o = 32; o; o >>= 1) l += __shfl_xor(l, o);
  if ((threadIdx.x & 63) == 0 && loss_acc) atomicAdd(loss_acc, l / (float)tot);
}

// out[n] = sum_m X[m*ld + n] (fp32), one thread per column, sequential over m
__global__ void colsum_kernel(const float* __restrict__ X, long ld, int M, int N, float* __restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int m = 0; m < M; ++m) s += X[(long)m * ld + n];
  out[n] = s;
}

// Adam (Kingma & Ba) with bias correction on fp32 master weights; writes the bf16 copy used by the GEMMs.
__global__ void adam_kernel(float* __restrict__ W, const float* __restrict__ G, float* __restrict__ mo,
                            float* __restrict__ ve, long n, float lr, float b1, float b2, float eps, float c1,
                            float c2, uint16_t* __restrict__ Wb) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = G[i];
  const float m = b1 * mo[i] + (1.f - b1) * g;
  const float v = b2 * ve[i] + (1.f - b2) * g * g;
  mo[i] = m;
  ve[i] = v;
  const float w = W[i] - lr * (m / c1) / (sqrtf(v / c2) + eps);
  W[i] = w;
  if (Wb) Wb[i] = f2bf(w);
}

__global__ void cast_bf16_kernel(const float* __restrict__ x, long n, uint16_t* __restrict__ y) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// Start thresholds: spin +1 iff the start word x >= 2^32 - thr, thr = round(sigmoid(z * inv_temp) 2^32) clamped to
// [1, 2^32 - 1]; sigmoid = 1/2 gives thr = 2^31, the plain random start.
__global__ void start_thr_kernel(const float* __restrict__ z, long ld, int M, int N, float inv_temp,
                                 uint32_t* __restrict__ thr) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)M * N) return;
  const int m = (int)(i / N), n = (int)(i - (long)m * N);
  const double p = 1.0 / (1.0 + exp(-(double)z[m * ld + n] * (double)inv_temp));
  double t = rint(p * 4294967296.0);
  t = t < 1.0 ? 1.0 : (t > 4294967295.0 ? 4294967295.0 : t);
  thr[i] = (uint32_t)t;
}

static inline dim3 grid1(long n) { return dim3((unsigned)((n + 255) / 256)); }

hipError_t launch_gemm_bf16(const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn, int M, int N,
                            int K, const float* bias, int act, float* Cf, float* Cpre, uint16_t* Cb, long ldc,
                            hipStream_t st) {
  dim3 grid((M + 63) / 64, (N + 63) / 64);
  hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, st, A, sam, sak, B, sbk, sbn, M, N, K, bias, act, Cf, Cpre,
                     Cb, ldc);
  return hipGetLastError();
}

hipError_t launch_gelu_bwd(const float* dH, const float* Apre, long n, uint16_t* dAb, float* dAf, hipStream_t st) {
  hipLaunchKernelGGL(gelu_bwd_kernel, grid1(n), dim3(256), 0, st, dH, Apre, n, dAb, dAf);
  return hipGetLastError();
}

hipError_t launch_bce(const float* z, long ld, const int8_t* y, int M, int N, uint16_t* dzb, float* dzf,
                      float* loss_acc, hipStream_t st) {
  hipLaunchKernelGGL(bce_kernel, grid1((long)M * N), dim3(256), 0, st, z, ld, y, M, N, dzb, dzf, loss_acc);
  return hipGetLastError();
}

hipError_t launch_colsum(const float* X, long ld, int M, int N, float* out, hipStream_t st) {
  hipLaunchKernelGGL(colsum_kernel, grid1(N), dim3(256), 0, st, X, ld, M, N, out);
  return hipGetLastError();
}

hipError_t launch_adam(float* W, const float* G, float* m, float* v, long n, float lr, float b1, float b2, float eps,
                       int step, uint16_t* Wb, hipStream_t st) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, grid1(n), dim3(256), 0, st, W, G, m, v, n, lr, b1, b2, eps, c1, c2, Wb);
  return hipGetLastError();
}

hipError_t launch_cast_bf16(const float* x, long n, uint16_t* y, hipStream_t st) {
  hipLaunchKernelGGL(cast_bf16_kernel, grid1(n), dim3(256), 0, st, x, n, y);
  return hipGetLastError();
}

hipError_t launch_start_thr(const float* z, long ld, int M, int N, float inv_temp, uint32_t* thr, hipStream_t st) {
  hipLaunchKernelGGL(start_thr_kernel, grid1((long)M * N), dim3(256), 0, st, z, ld, M, N, inv_temp, thr);
  return hipGetLastError();
}

}  // namespace arp
