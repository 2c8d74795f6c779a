// The return-map warm start: an 8 → 64 → 64 → 64 → 1 tanh network that maps a material
// point's (6 Voigt strain increments, ε̄ᵖ, T) to an estimate of its plastic increment Δε̄ᵖ,
// which seeds the safeguarded Newton solve of the Johnson-Cook return map.
//
//  * Training runs on the GPU in FP32 with these kernels: per-layer dense forward, the
//    backward pass (input gradients and per-chunk weight-gradient partials reduced in a
//    fixed order, so a training run is deterministic) and Adam.
//  * Inference runs as bf16 MFMA tiles (v_mfma_f32_16x16x16_bf16) with FP32 accumulation:
//    one wavefront pushes 16 points through the four layers, staging each hidden layer in
//    LDS between the MFMA output layout and the next layer's input layout.
//
// Flat parameter layout (FP32): W1 (64×8), b1 (64), W2 (64×64), b2 (64), W3 (64×64), b3 (64),
// W4 (1×64), b4 (1): 8961 values. Normalisation (FP32, 18 values): input mean[8], std[8],
// output mean, std; the network sees (x − mean)/std and its output maps back to
// Δε̄ᵖ = max(0, y·std + mean).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vpd {

constexpr int MLP_IN = 8, MLP_H = 64;
constexpr int MLP_OFF_W1 = 0, MLP_OFF_B1 = 512, MLP_OFF_W2 = 576, MLP_OFF_B2 = 4672,
              MLP_OFF_W3 = 4736, MLP_OFF_B3 = 8832, MLP_OFF_W4 = 8896, MLP_OFF_B4 = 8960;
constexpr int MLP_PARAMS = 8961;
constexpr int MLP_NORM = 18;

// ------------------------------------------------------------------ training (FP32)

// Y[b][n] = act(Σ_k X[b][k] W[n][k] + bias[n]); act 1 = tanh, 0 = identity.
__global__ void __launch_bounds__(256) k_mlp_dense(int B, int K, int N, const float *__restrict__ X,
                                                   const float *__restrict__ W,
                                                   const float *__restrict__ bias,
                                                   float *__restrict__ Y, int act) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)B * N)
    return;
  const int b = (int)(t / N), n = (int)(t % N);
  float acc = bias[n];
  for (int k = 0; k < K; ++k)
    acc += X[(int64_t)b * K + k] * W[n * K + k];
  Y[t] = act ? tanhf(acc) : acc;
}

// dy = 2 (y − target) / B and per-sample squared error (for the batch loss).
__global__ void __launch_bounds__(256) k_mlp_loss_grad(int B, const float *__restrict__ y,
                                                       const float *__restrict__ target,
                                                       float *__restrict__ dy,
                                                       float *__restrict__ sq) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B)
    return;
  const float r = y[b] - target[b];
  dy[b] = 2.0f * r / (float)B;
  sq[b] = r * r;
}

// dX[b][k] = (Σ_n dZ[b][n] W[n][k]) · (tanh' at H[b][k] when H is given).
__global__ void __launch_bounds__(256) k_mlp_back_input(int B, int K, int N,
                                                        const float *__restrict__ dZ,
                                                        const float *__restrict__ W,
                                                        const float *__restrict__ H,
                                                        float *__restrict__ dX) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)B * K)
    return;
  const int b = (int)(t / K), k = (int)(t % K);
  float acc = 0.f;
  for (int n = 0; n < N; ++n)
    acc += dZ[(int64_t)b * N + n] * W[n * K + k];
  if (H) {
    const float h = H[t];
    acc *= 1.0f - h * h;
  }
  dX[t] = acc;
}

// Weight-gradient partials over one chunk of samples: part[c][n][k] = Σ_{b∈chunk} dZ[b][n]
// X[b][k], with column k = K the bias (X = 1). Reduced afterwards in chunk order.
constexpr int MLP_CHUNK = 256;
__global__ void __launch_bounds__(256) k_mlp_wgrad_part(int B, int K, int N,
                                                        const float *__restrict__ dZ,
                                                        const float *__restrict__ X,
                                                        float *__restrict__ part) {
  const int cols = K + 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int chunk = blockIdx.y;
  if (t >= N * cols)
    return;
  const int n = t / cols, k = t % cols;
  const int b0 = chunk * MLP_CHUNK, b1 = min(B, b0 + MLP_CHUNK);
  float acc = 0.f;
  for (int b = b0; b < b1; ++b)
    acc += dZ[(int64_t)b * N + n] * (k < K ? X[(int64_t)b * K + k] : 1.0f);
  part[(int64_t)chunk * N * cols + t] = acc;
}

// grad[W: n*K + k] and grad[bias: n] from the partials, summed over chunks in order.
__global__ void __launch_bounds__(256) k_mlp_wgrad_reduce(int chunks, int K, int N,
                                                          const float *__restrict__ part,
                                                          float *__restrict__ gW,
                                                          float *__restrict__ gb) {
  const int cols = K + 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * cols)
    return;
  float acc = 0.f;
  for (int c = 0; c < chunks; ++c)
    acc += part[(int64_t)c * N * cols + t];
  const int n = t / cols, k = t % cols;
  if (k < K)
    gW[n * K + k] = acc;
  else
    gb[n] = acc;
}

// Adam (Kingma & Ba 2015) with bias correction.
__global__ void __launch_bounds__(256) k_mlp_adam(int P, float *__restrict__ theta,
                                                  const float *__restrict__ g,
                                                  float *__restrict__ m, float *__restrict__ v,
                                                  float lr, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P)
    return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  theta[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

// Copies the batch rows order[start .. start+B) (wrapping) out of the dataset.
__global__ void __launch_bounds__(256) k_mlp_gather(int B, int64_t n, int64_t start,
                                                    const int *__restrict__ order,
                                                    const float *__restrict__ xs,
                                                    const float *__restrict__ ys,
                                                    float *__restrict__ xb,
                                                    float *__restrict__ yb) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B)
    return;
  const int64_t src = order[(start + b) % n];
#pragma unroll
  for (int k = 0; k < MLP_IN; ++k)
    xb[(int64_t)b * MLP_IN + k] = xs[src * MLP_IN + k];
  yb[b] = ys[src];
}

// ------------------------------------------------------------------ inference (bf16 MFMA)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Round-to-nearest-even FP32 → bf16 bit pattern.
__device__ __forceinline__ short to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// One hidden or output layer for the wavefront's 16 points. Input activations h_in
// [16][K] (FP32, LDS), weights W [N][K] row-major (global), output → h_out [16][N] (LDS)
// after bias and activation. v_mfma_f32_16x16x16_bf16 operand layouts (lane l):
//   A[m][k]: m = l % 16, k = 4 (l / 16) + i    (i = 0..3 in the 4-vector)
//   B[k][n]: n = l % 16, k = 4 (l / 16) + i
//   D[m][n]: n = l % 16, m = 4 (l / 16) + i
template <int K, int N, bool TANH>
__device__ __forceinline__ void mlp_layer_mfma(const float *h_in, const float *__restrict__ W,
                                               const float *__restrict__ bias, float *h_out,
                                               int lane) {
  const int r = lane & 15, q = lane >> 4;
  constexpr int KT = (K + 15) / 16, NT = (N + 15) / 16;
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < KT; ++kt) {
      bf16x4 a, b;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k = kt * 16 + 4 * q + i;
        const int n = nt * 16 + r;
        a[i] = to_bf16(k < K ? h_in[r * K + k] : 0.f);
        b[i] = to_bf16(k < K && n < N ? W[n * K + k] : 0.f);
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const int n = nt * 16 + r;
    if (n < N) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float z = acc[i] + bias[n];
        h_out[(4 * q + i) * N + n] = TANH ? tanhf(z) : z;
      }
    }
  }
}

// (x − mean) / std in FP64, then rounded to FP32: a feature whose spread is far below its
// magnitude (the temperature moves by micro-kelvins at 293 K) keeps its information, which
// an FP32 subtraction would round away.
__device__ __forceinline__ void normalise_features(const double x[MLP_IN],
                                                   const float *__restrict__ norm, float *out) {
#pragma unroll
  for (int c = 0; c < MLP_IN; ++c)
    out[c] = (float)((x[c] - (double)norm[c]) / (double)norm[MLP_IN + c]);
}

// Network seeds for n points. Inputs per point: strain increments (6 planes, stride
// incStride), ε̄ᵖ and T (planes of `state`, stride stateStride: components 1 and 2 of the
// JC state). Writes seed[i] = max(0, y·std + mean). One wavefront per 16 points.
__global__ void __launch_bounds__(256) k_mlp_seed(int64_t n, const double *__restrict__ inc,
                                                  int64_t incStride,
                                                  const double *__restrict__ state,
                                                  int64_t stateStride,
                                                  const float *__restrict__ theta,
                                                  const float *__restrict__ norm,
                                                  double *__restrict__ seed) {
  __shared__ float buf[4][2][16 * MLP_H];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t base = ((int64_t)blockIdx.x * 4 + wave) * 16;
  // Every wave runs every layer (the MFMAs need the full wavefront and the barriers the
  // whole workgroup); waves past the end work on zeros and store nothing.
  float *h0 = buf[wave][0], *h1 = buf[wave][1];
  if (lane < 16) {
    const int64_t i = base + lane;
    double x[MLP_IN];
#pragma unroll
    for (int c = 0; c < MLP_IN; ++c)
      x[c] = norm[c]; // padding rows normalise to 0
    if (i < n) {
#pragma unroll
      for (int c = 0; c < 6; ++c)
        x[c] = inc[c * incStride + i];
      x[6] = state[1 * stateStride + i];
      x[7] = state[2 * stateStride + i];
    }
    normalise_features(x, norm, h0 + lane * MLP_IN);
  }
  __syncthreads();
  mlp_layer_mfma<MLP_IN, MLP_H, true>(h0, theta + MLP_OFF_W1, theta + MLP_OFF_B1, h1, lane);
  __syncthreads();
  mlp_layer_mfma<MLP_H, MLP_H, true>(h1, theta + MLP_OFF_W2, theta + MLP_OFF_B2, h0, lane);
  __syncthreads();
  mlp_layer_mfma<MLP_H, MLP_H, true>(h0, theta + MLP_OFF_W3, theta + MLP_OFF_B3, h1, lane);
  __syncthreads();
  mlp_layer_mfma<MLP_H, 1, false>(h1, theta + MLP_OFF_W4, theta + MLP_OFF_B4, h0, lane);
  __syncthreads();
  if (lane < 16 && base + lane < n) {
    const float y = h0[lane] * norm[2 * MLP_IN + 1] + norm[2 * MLP_IN];
    seed[base + lane] = y > 0.f ? (double)y : 0.0;
  }
}

// k_mlp_seed over a compacted point list: slot s of the list holds particle list[s], its
// strain increments at inc[c·incStride + s]; the state planes are indexed by particle. The
// list length is read on the device; blocks past it exit whole.
__global__ void __launch_bounds__(256) k_mlp_seed_list(const int *__restrict__ count,
                                                       const int *__restrict__ list,
                                                       const double *__restrict__ inc,
                                                       int64_t incStride,
                                                       const double *__restrict__ state,
                                                       int64_t stateStride,
                                                       const float *__restrict__ theta,
                                                       const float *__restrict__ norm,
                                                       double *__restrict__ seed) {
  const int64_t n = *count;
  if ((int64_t)blockIdx.x * 64 >= n) // the same for every thread of the block
    return;
  __shared__ float buf[4][2][16 * MLP_H];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t base = ((int64_t)blockIdx.x * 4 + wave) * 16;
  float *h0 = buf[wave][0], *h1 = buf[wave][1];
  if (lane < 16) {
    const int64_t s = base + lane;
    double x[MLP_IN];
#pragma unroll
    for (int c = 0; c < MLP_IN; ++c)
      x[c] = norm[c]; // padding rows normalise to 0
    if (s < n) {
      const int64_t i = list[s];
#pragma unroll
      for (int c = 0; c < 6; ++c)
        x[c] = inc[c * incStride + s];
      x[6] = state[1 * stateStride + i];
      x[7] = state[2 * stateStride + i];
    }
    normalise_features(x, norm, h0 + lane * MLP_IN);
  }
  __syncthreads();
  mlp_layer_mfma<MLP_IN, MLP_H, true>(h0, theta + MLP_OFF_W1, theta + MLP_OFF_B1, h1, lane);
  __syncthreads();
  mlp_layer_mfma<MLP_H, MLP_H, true>(h1, theta + MLP_OFF_W2, theta + MLP_OFF_B2, h0, lane);
  __syncthreads();
  mlp_layer_mfma<MLP_H, MLP_H, true>(h0, theta + MLP_OFF_W3, theta + MLP_OFF_B3, h1, lane);
  __syncthreads();
  mlp_layer_mfma<MLP_H, 1, false>(h1, theta + MLP_OFF_W4, theta + MLP_OFF_B4, h0, lane);
  __syncthreads();
  if (lane < 16 && base + lane < n) {
    const float y = h0[lane] * norm[2 * MLP_IN + 1] + norm[2 * MLP_IN];
    seed[base + lane] = y > 0.f ? (double)y : 0.0;
  }
}

// Raw network outputs (normalised units) for a batch of already-normalised inputs
// x[b][8]: the inference kernel's arithmetic, exposed for checking against the FP64 net.
__global__ void __launch_bounds__(256) k_mlp_infer_raw(int64_t n, const float *__restrict__ x,
                                                       const float *__restrict__ theta,
                                                       float *__restrict__ y) {
  __shared__ float buf[4][2][16 * MLP_H];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t base = ((int64_t)blockIdx.x * 4 + wave) * 16;
  float *h0 = buf[wave][0], *h1 = buf[wave][1];
  if (lane < 16) {
#pragma unroll
    for (int c = 0; c < MLP_IN; ++c)
      h0[lane * MLP_IN + c] = base + lane < n ? x[(base + lane) * MLP_IN + c] : 0.f;
  }
  __syncthreads();
  mlp_layer_mfma<MLP_IN, MLP_H, true>(h0, theta + MLP_OFF_W1, theta + MLP_OFF_B1, h1, lane);
  __syncthreads();
  mlp_layer_mfma<MLP_H, MLP_H, true>(h1, theta + MLP_OFF_W2, theta + MLP_OFF_B2, h0, lane);
  __syncthreads();
  mlp_layer_mfma<MLP_H, MLP_H, true>(h0, theta + MLP_OFF_W3, theta + MLP_OFF_B3, h1, lane);
  __syncthreads();
  mlp_layer_mfma<MLP_H, 1, false>(h1, theta + MLP_OFF_W4, theta + MLP_OFF_B4, h0, lane);
  __syncthreads();
  if (lane < 16 && base + lane < n)
    y[base + lane] = h0[lane];
}

} // namespace vpd
