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

// tanh u = sign(u) (1 - 2 / (e^(2|u|) + 1)), e^(2|u|) = 2^(2|u| log2 e) from the hardware exp2 and the reciprocal:
// about 1e-6 relative, far below the bf16 rounding of the activations; large |u| saturates to +-1.
__device__ __forceinline__ float tanh_fast(float u) {
  const float e = __builtin_amdgcn_exp2f(fminf(2.8853900817779268f * fabsf(u), 64.f));  // 2 log2(e) |u|
  return copysignf(1.f - 2.f * __builtin_amdgcn_rcpf(e + 1.f), u);
}

// GELU, tanh form: 0.5 x (1 + tanh(c (x + 0.044715 x^3))), c = sqrt(2/pi)
__device__ __forceinline__ float gelu_f(float x) {
  const float c = 0.7978845608028654f;
  return 0.5f * x * (1.f + tanh_fast(c * (x + 0.044715f * x * x * x)));
}

__device__ __forceinline__ float gelu_grad(float x) {
  const float c = 0.7978845608028654f;
  const float t = tanh_fast(c * (x + 0.044715f * x * x * x));
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

// Fused inference: features -> thresholds for a tile of 64 rows in one workgroup. The tile's features are staged in
// LDS once; each hidden layer's 64 x 256 output is computed by the four waves (64 columns each) from the LDS
// activations and n-major weights read from L2 (one 16-byte load per lane per 32-wide k-step), then written back
// over the activations as bf16 after GELU; the last layer's logits go straight to start thresholds. Every output
// element accumulates the same bf16 products in the same order as gemm_bf16_kernel (k ascending in steps of 32, the
// same lane split of k), so the thresholds equal the three-GEMM path's bit for bit.
template <int NC>
__device__ __forceinline__ void mlp_tile_layer(const uint16_t* act, int lda, int K, const uint16_t* __restrict__ Wt,
                                               int n0, floatx4 (&acc)[4][NC], int li, int lg) {
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int ct = 0; ct < NC; ++ct) acc[rt][ct] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 32) {
    bf16x8 a[4], bw[NC];
#pragma unroll
    for (int ct = 0; ct < NC; ++ct)
      bw[ct] = *reinterpret_cast<const bf16x8*>(Wt + (size_t)(n0 + ct * 16 + li) * K + k0 + lg * 8);
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) a[rt] = *reinterpret_cast<const bf16x8*>(act + (rt * 16 + li) * lda + k0 + lg * 8);
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int ct = 0; ct < NC; ++ct) acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[rt], bw[ct], acc[rt][ct], 0, 0, 0);
  }
}

template <int NC3>
__global__ __launch_bounds__(256) void mlp_start_fused_kernel(const uint16_t* __restrict__ X, int M, int F,
                                                              const uint16_t* __restrict__ W1t,
                                                              const float* __restrict__ b1,
                                                              const uint16_t* __restrict__ W2t,
                                                              const float* __restrict__ b2,
                                                              const uint16_t* __restrict__ W3t,
                                                              const float* __restrict__ b3, int N, float inv_temp,
                                                              uint32_t* __restrict__ thr) {
  constexpr int HID = 256, TM = 64;
  extern __shared__ __align__(16) uint16_t act[];
  const int lda = (F > HID ? F : HID) + 8;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, li = lane & 15, lg = lane >> 4;
  const int m0 = blockIdx.x * TM;
  const int cpr = F / 8;
  for (int x = tid; x < TM * cpr; x += 256) {
    const int r = x / cpr, c = x - r * cpr;
    uint4 v = {0u, 0u, 0u, 0u};
    if (m0 + r < M) v = *reinterpret_cast<const uint4*>(X + (size_t)(m0 + r) * F + 8 * c);
    *reinterpret_cast<uint4*>(act + r * lda + 8 * c) = v;
  }
  __syncthreads();
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    floatx4 acc[4][4];
    const int n0 = wv * 64;
    mlp_tile_layer<4>(act, lda, layer == 0 ? F : HID, layer == 0 ? W1t : W2t, n0, acc, li, lg);
    const float* bb = layer == 0 ? b1 : b2;
    __syncthreads();  // every wave has read the layer's input
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int ct = 0; ct < 4; ++ct) {
        const int col = n0 + ct * 16 + li;
        const float bn = bb[col];
#pragma unroll
        for (int v = 0; v < 4; ++v) act[(rt * 16 + lg * 4 + v) * lda + col] = f2bf(gelu_f(acc[rt][ct][v] + bn));
      }
    __syncthreads();
  }
  floatx4 acc[4][NC3];
  const int n0 = wv * NC3 * 16;
  mlp_tile_layer<NC3>(act, lda, HID, W3t, n0, acc, li, lg);
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int ct = 0; ct < NC3; ++ct) {
      const int col = n0 + ct * 16 + li;
      const float bn = b3[col];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + rt * 16 + lg * 4 + v;
        if (m >= M) continue;
        const double p = 1.0 / (1.0 + exp(-(double)(acc[rt][ct][v] + bn) * (double)inv_temp));
        double t = rint(p * 4294967296.0);
        t = t < 1.0 ? 1.0 : (t > 4294967295.0 ? 4294967295.0 : t);
        thr[(size_t)m * N + col] = (uint32_t)t;
      }
    }
}

hipError_t launch_mlp_start_fused(const uint16_t* X, int M, int F, const uint16_t* W1t, const float* b1,
                                  const uint16_t* W2t, const float* b2, const uint16_t* W3t, const float* b3, int N,
                                  float inv_temp, uint32_t* thr, hipStream_t st) {
  const int lda = (F > 256 ? F : 256) + 8;
  const size_t lds = sizeof(uint16_t) * 64 * (size_t)lda;   // the tile's activations: up to 128 KiB at F = 1016
  const dim3 grid((M + 63) / 64);
  if (lds > 65536) {
    hipFuncSetAttribute((const void*)mlp_start_fused_kernel<1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipFuncSetAttribute((const void*)mlp_start_fused_kernel<2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipFuncSetAttribute((const void*)mlp_start_fused_kernel<3>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipFuncSetAttribute((const void*)mlp_start_fused_kernel<4>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  }
  switch (N / 64) {
    case 1: hipLaunchKernelGGL((mlp_start_fused_kernel<1>), grid, dim3(256), lds, st, X, M, F, W1t, b1, W2t, b2, W3t, b3, N, inv_temp, thr); break;
    case 2: hipLaunchKernelGGL((mlp_start_fused_kernel<2>), grid, dim3(256), lds, st, X, M, F, W1t, b1, W2t, b2, W3t, b3, N, inv_temp, thr); break;
    case 3: hipLaunchKernelGGL((mlp_start_fused_kernel<3>), grid, dim3(256), lds, st, X, M, F, W1t, b1, W2t, b2, W3t, b3, N, inv_temp, thr); break;
    default: hipLaunchKernelGGL((mlp_start_fused_kernel<4>), grid, dim3(256), lds, st, X, M, F, W1t, b1, W2t, b2, W3t, b3, N, inv_temp, thr); break;
  }
  return hipGetLastError();
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
