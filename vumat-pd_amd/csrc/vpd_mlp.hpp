// The return-map warm start: an 8 → 64 → 64 → 64 → 1 ReLU network that maps a material
// point's (6 Voigt strain increments, ε̄ᵖ, T) to an estimate of its plastic increment Δε̄ᵖ,
// which seeds the safeguarded Newton solve of the Johnson-Cook return map.
//
//  * Training runs on the GPU in FP32 with these kernels: per-layer dense forward, the
//    backward pass (input gradients and per-chunk weight-gradient partials reduced in a
//    fixed order, so a training run is deterministic) and Adam.
//  * Inference runs as bf16 MFMA tiles (v_mfma_f32_16x16x16_bf16) with FP32 accumulation,
//    the whole network resident in each wavefront's registers (see MlpWeights). ReLU keeps
//    the per-element work between the MFMAs to a max and a packed bf16 conversion, so the
//    inference is bound by the matrix cores rather than by transcendental VALU work.
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

// Y[b][n] = act(Σ_k X[b][k] W[n][k] + bias[n]); act 1 = ReLU, 0 = identity.
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
  Y[t] = act ? fmaxf(acc, 0.f) : acc;
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

// dX[b][k] = (Σ_n dZ[b][n] W[n][k]) · (ReLU' at H[b][k] when H is given: 1 where H > 0).
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
  if (H && !(H[t] > 0.f))
    acc = 0.f;
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

// Inference keeps the whole network in the wavefront's registers and computes each layer as
// D[m][n] = Σ_k W[m][k] X[n][k] with the output feature as the MFMA M dimension and the point
// as N. v_mfma_f32_16x16x16_bf16 operand layouts (lane l, i = 0..3 in the 4-vector):
//   A[m][k]: m = l % 16, k = 4 (l / 16) + i    (a 16×16 weight tile)
//   B[k][n]: n = l % 16, k = 4 (l / 16) + i    (16 points' activations)
//   D[m][n]: n = l % 16, m = 4 (l / 16) + i
// A layer's D tile for features 16t..16t+15 is then exactly the B operand of the next layer's
// K-slice t, so activations stay in registers from input to output, and a wavefront loads
// the weights once for all the 16-point tiles it processes.
struct MlpWeights {
  bf16x4 w1[4];    // [M-tile]: features 16t.., inputs 0..15 (8 real)
  bf16x4 w2[4][4]; // [M-tile][K-slice]
  bf16x4 w3[4][4];
  bf16x4 w4[4];    // output 0 (tile row 0) × K-slices
  f32x4 b1[4], b2[4], b3[4]; // biases in the D layout: feature 16t + 4 (l / 16) + i
  float b4;
};

// FP32 → bf16 (round to nearest even) with the hardware's packed conversion.
// Converted in pairs (one v_cvt_pk_bf16_f32 each): a 4-wide conversion is lowered to
// per-element conversions and permutes.
typedef __bf16 bf16v2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef short s16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ s16x2 bf16_pair(float a, float b) {
  return __builtin_bit_cast(s16x2, __builtin_convertvector(f32x2{a, b}, bf16v2));
}
__device__ __forceinline__ bf16x4 bf16_of(f32x4 v) {
  const s16x2 lo = bf16_pair(v[0], v[1]), hi = bf16_pair(v[2], v[3]);
  return bf16x4{lo[0], lo[1], hi[0], hi[1]};
}

__device__ __forceinline__ void mlp_load_weights(const float *__restrict__ theta, int lane,
                                                 MlpWeights &w) {
  const int r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f32x4 v = {0.f, 0.f, 0.f, 0.f};
    if (q < 2)
      v = *(const f32x4 *)(theta + MLP_OFF_W1 + (16 * t + r) * MLP_IN + 4 * q);
    w.w1[t] = bf16_of(v);
  }
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int off = (16 * t + r) * MLP_H + 16 * s + 4 * q;
      w.w2[t][s] = bf16_of(*(const f32x4 *)(theta + MLP_OFF_W2 + off));
      w.w3[t][s] = bf16_of(*(const f32x4 *)(theta + MLP_OFF_W3 + off));
    }
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    f32x4 v = {0.f, 0.f, 0.f, 0.f};
    if (r == 0)
      v = *(const f32x4 *)(theta + MLP_OFF_W4 + 16 * s + 4 * q);
    w.w4[s] = bf16_of(v);
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    w.b1[t] = *(const f32x4 *)(theta + MLP_OFF_B1 + 16 * t + 4 * q);
    w.b2[t] = *(const f32x4 *)(theta + MLP_OFF_B2 + 16 * t + 4 * q);
    w.b3[t] = *(const f32x4 *)(theta + MLP_OFF_B3 + 16 * t + 4 * q);
  }
  w.b4 = theta[MLP_OFF_B4];
}

// ReLU after the conversion, on the packed bf16 pairs as signed 16-bit integers: a bf16
// with the sign bit set is a negative int16 and becomes 0, a positive one is kept (the bit
// patterns of non-negative bf16 values order like their values). Same result as rounding
// max(v, 0), at a quarter of the instructions of a per-element IEEE max.
__device__ __forceinline__ bf16x4 relu_bf16(f32x4 v) {
  const s16x2 z = {0, 0};
  const s16x2 lo = __builtin_elementwise_max(bf16_pair(v[0], v[1]), z);
  const s16x2 hi = __builtin_elementwise_max(bf16_pair(v[2], v[3]), z);
  return bf16x4{lo[0], lo[1], hi[0], hi[1]};
}

// The forward pass for one tile of 16 points. x: the input tile in the B layout (lanes
// 0-15 hold inputs 0-3 of point l % 16, lanes 16-31 inputs 4-7, the rest zero). Returns the
// raw output of point l in lanes l < 16.
// The K-slices of a layer run outermost so that the 4 M-tiles' accumulation chains are
// independent at every step and the MFMA pipeline stays full.
__device__ __forceinline__ void mlp_hidden(const bf16x4 (&w)[4][4], const f32x4 (&b)[4],
                                           const bf16x4 (&in)[4], bf16x4 (&out)[4]) {
  f32x4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
    acc[t] = b[t];
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int t = 0; t < 4; ++t)
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w[t][s], in[s], acc[t], 0, 0, 0);
#pragma unroll
  for (int t = 0; t < 4; ++t)
    out[t] = relu_bf16(acc[t]);
}

__device__ __forceinline__ float mlp_tile(const MlpWeights &w, bf16x4 x) {
  bf16x4 h[4], g[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w.w1[t], x, w.b1[t], 0, 0, 0);
    h[t] = relu_bf16(acc);
  }
  mlp_hidden(w.w2, w.b2, h, g);
  mlp_hidden(w.w3, w.b3, g, h);
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < 4; ++s)
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w.w4[s], h[s], acc, 0, 0, 0);
  return acc[0] + w.b4;
}

// (x − mean) · (1/std) in FP64, then bf16: a feature whose spread is far below its magnitude
// (the temperature moves by micro-kelvins at 293 K) keeps its information, which an FP32
// subtraction would round away.
__device__ __forceinline__ bf16x4 normalised_bf16(const double x[4], const double mean[4],
                                                  const double invStd[4]) {
  f32x4 v;
#pragma unroll
  for (int c = 0; c < 4; ++c)
    v[c] = (float)((x[c] - mean[c]) * invStd[c]);
  return bf16_of(v);
}

// The network's estimate for the lanes of one wavefront that ask for it (`need`), inside a
// kernel that holds their inputs in registers: 6 strain increments, ε̄ᵖ and T, normalised in
// FP64 and rounded to FP32 exactly as k_mlp_seed does. Those lanes' inputs go to the
// wavefront's LDS (`xs`: 64 × 8 floats, `ys`: 64 doubles), then 16-point tiles of them run
// through mlp_tile (up to 4 tiles) with every lane active, and each asking lane gets back
// max(0, y·std + mean), the value k_mlp_seed writes for the same point (one MFMA column per
// point: the tile a point shares does not change its bits). Every lane of the wavefront must
// call (the branch on the ballot is uniform); a wavefront with no asking lane returns at once.
__device__ __forceinline__ double wave_network_seed(bool need, const double inc[6], double ep,
                                                    double temp,
                                                    const float *__restrict__ theta,
                                                    const float *__restrict__ norm, float *xs,
                                                    double *ys) {
  const uint64_t mask = __ballot(need);
  if (mask == 0)
    return 0.0;
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
  const int rank = __popcll(mask & ((1ull << lane) - 1ull));
  if (need) {
    const double x[MLP_IN] = {inc[0], inc[1], inc[2], inc[3], inc[4], inc[5], ep, temp};
#pragma unroll
    for (int c = 0; c < MLP_IN; ++c)
      xs[MLP_IN * rank + c] =
          (float)((x[c] - (double)norm[c]) * (1.0 / (double)norm[MLP_IN + c]));
  }
  MlpWeights w;
  mlp_load_weights(theta, lane, w);
  const double mean = norm[2 * MLP_IN], sdev = norm[2 * MLP_IN + 1];
  const int count = __popcll(mask);
  __builtin_amdgcn_s_waitcnt(0xC07F); // lgkmcnt(0): this wavefront's LDS stores are done
  __builtin_amdgcn_wave_barrier();
  for (int t = 0; 16 * t < count; ++t) {
    const int pt = 16 * t + r;
    f32x4 v = {0.f, 0.f, 0.f, 0.f}; // a missing point normalises to 0, as in k_mlp_seed
    if (q < 2 && pt < count)
      v = *(const f32x4 *)(xs + MLP_IN * pt + 4 * q);
    bf16x4 x = {0, 0, 0, 0};
    if (q < 2)
      x = bf16_of(v);
    const float y = mlp_tile(w, x);
    if (lane < 16 && pt < count) {
      const double est = (double)y * sdev + mean;
      ys[pt] = est > 0.0 ? est : 0.0;
    }
  }
  __builtin_amdgcn_s_waitcnt(0xC07F);
  __builtin_amdgcn_wave_barrier();
  return need ? ys[rank] : 0.0;
}

// Return-map seeds for n points (particle s, increments at inc[c·incStride + s]). A point
// whose previous plastic increment (state 5, component 4 of the JC state planes, stride
// stateStride) is positive keeps it as its seed: on a continuing plastic path it is closer to
// the solution than the network's estimate (about 2% median error). A point with state 5 = 0
// (first yield, or reloading after an elastic increment), where that seed says nothing, takes
// the network's estimate from the 6 strain increments, ε̄ᵖ and T (components 1 and 2):
// seed = max(0, y·std + mean). Each wavefront walks 16-point tiles with a grid stride,
// reading state 5 first; a tile whose points all carry a state-5 seed skips the network and
// its inputs, and a wavefront loads the weights at its first tile that needs them.
// Wavefronts with no tile exit at once, and the tile loop (and the skip) is uniform per
// wavefront, so every MFMA runs with all lanes active. A point's estimate depends on its own
// inputs only (one MFMA column), not on the tile it shares.
__global__ void __launch_bounds__(256) k_mlp_seed(int64_t n, const double *__restrict__ inc,
                                                  int64_t incStride,
                                                  const double *__restrict__ state,
                                                  int64_t stateStride,
                                                  const float *__restrict__ theta,
                                                  const float *__restrict__ norm,
                                                  double *__restrict__ seed) {
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
  const int64_t tiles = (n + 15) / 16;
  const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int64_t waves = (int64_t)gridDim.x * 4;
  if (wave >= tiles)
    return;
  MlpWeights w;
  bool loaded = false;
  double mean = 0.0, sdev = 0.0, mu[4], inv[4];
  for (int64_t tile = wave; tile < tiles; tile += waves) {
    const int64_t s = tile * 16 + r;
    const double prev = q == 0 && s < n ? state[4 * stateStride + s] : 1.0; // state 5
    const bool needNet = q == 0 && s < n && !(prev > 0.0);
    float y = 0.f;
    if (__ballot(needNet) != 0) { // uniform: the tile has a point without a state-5 seed
      if (!loaded) {
        mlp_load_weights(theta, lane, w);
        mean = norm[2 * MLP_IN];
        sdev = norm[2 * MLP_IN + 1];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int f = q < 2 ? 4 * q + c : c;
          mu[c] = norm[f];
          inv[c] = 1.0 / (double)norm[MLP_IN + f];
        }
        loaded = true;
      }
      // This lane's 4 raw inputs (lanes 0-15: increments 0-3; lanes 16-31: increments 4-5,
      // ε̄ᵖ, T); missing points normalise to 0.
      double v[4];
#pragma unroll
      for (int c = 0; c < 4; ++c)
        v[c] = mu[c];
      if (q < 2 && s < n) {
        if (q == 0) {
#pragma unroll
          for (int c = 0; c < 4; ++c)
            v[c] = inc[c * incStride + s];
        } else {
          v[0] = inc[4 * incStride + s];
          v[1] = inc[5 * incStride + s];
          v[2] = state[1 * stateStride + s];
          v[3] = state[2 * stateStride + s];
        }
      }
      bf16x4 x = {0, 0, 0, 0};
      if (q < 2)
        x = normalised_bf16(v, mu, inv);
      y = mlp_tile(w, x);
    }
    if (q == 0 && s < n) {
      const double e = (double)y * sdev + mean;
      seed[s] = needNet ? (e > 0.0 ? e : 0.0) : prev;
    }
  }
}

// Raw network outputs (normalised units) for already-normalised FP32 inputs x[b][8]: the
// inference arithmetic of k_mlp_seed, exposed for checking against the FP64 network.
__global__ void __launch_bounds__(256) k_mlp_infer_raw(int64_t n, const float *__restrict__ x,
                                                       const float *__restrict__ theta,
                                                       float *__restrict__ y) {
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
  const int64_t tiles = (n + 15) / 16;
  const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int64_t waves = (int64_t)gridDim.x * 4;
  if (wave >= tiles)
    return;
  MlpWeights w;
  mlp_load_weights(theta, lane, w);
  for (int64_t tile = wave; tile < tiles; tile += waves) {
    const int64_t s = tile * 16 + r;
    bf16x4 xt = {0, 0, 0, 0};
    if (q < 2 && s < n)
      xt = bf16_of(*(const f32x4 *)(x + s * MLP_IN + 4 * q));
    const float out = mlp_tile(w, xt);
    if (q == 0 && s < n)
      y[s] = out;
  }
}

// ------------------------------------------------------------------ training (bf16 MFMA)
// The training step as three kernels, every product on the matrix cores
// (v_mfma_f32_16x16x16_bf16: bf16 operands, FP32 accumulation), in the operand layouts of
// the inference above:
//  1. k_mlp_fwd_bwd: per 16-point tile and wavefront, the forward pass (as mlp_tile), the
//     loss gradient dy = 2(y − t)/B and the backward pass down to the first layer's
//     pre-activation gradient: dz3 = W4ᵀ dy ⊙ [h3 > 0] elementwise, then
//     dz2 = (W3ᵀ dz3) ⊙ [h2 > 0] and dz1 = (W2ᵀ dz2) ⊙ [h1 > 0] as MFMAs with Wᵀ tiles as
//     the A operand and the previous gradient's accumulator tile as the B operand (the same
//     chaining as the forward pass). The layer inputs (x, h1, h2, h3) and gradients (dz1-3)
//     go to memory as bf16, feature-major ([feature][point]), dy as FP32.
//  2. k_mlp_wgrad: the weight and bias gradients as GEMMs over the batch, gW = dz·hᵀ with the
//     points as the MFMA K dimension (feature-major storage makes each operand a 4-point,
//     8-byte load), the bias gradient as the product with a ones column; one wavefront per
//     64-point chunk writes its partial gradient.
//  3. k_mlp_reduce_adam: the partials summed per parameter in chunk order, then (optionally)
//     the Adam update of k_mlp_adam. The sums run in a fixed order: a training run is
//     deterministic.
// Weights reach the wavefronts through LDS: each workgroup converts θ once into the bf16
// operand tiles (72 tiles of 64 lanes × 4) and the FP32 bias/W4 vectors in the D layout.
constexpr int MLP_WG_TILES = 72; // w1 4, w2 16, w3 16, w4 4, w2ᵀ 16, w3ᵀ 16
constexpr int MLP_WG_T_W1 = 0, MLP_WG_T_W2 = 4, MLP_WG_T_W3 = 20, MLP_WG_T_W4 = 36,
              MLP_WG_T_W2T = 40, MLP_WG_T_W3T = 56;
constexpr int MLP_WG_VECS = 16; // b1 4, b2 4, b3 4, W4 4 (f32x4 per lane, D layout)
constexpr int MLP_GRAD_CHUNK = 64; // points per wavefront of k_mlp_wgrad

// The bf16 operand tile `tile` for lane `lane` (see the A-layout above), from θ.
__device__ __forceinline__ bf16x4 mlp_operand(const float *__restrict__ theta, int tile,
                                              int lane) {
  const int r = lane & 15, q = lane >> 4;
  f32x4 v = {0.f, 0.f, 0.f, 0.f};
  if (tile < MLP_WG_T_W2) { // w1[t]: W1[16t + r][4q + i], inputs 8..15 zero
    if (q < 2)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        v[i] = theta[MLP_OFF_W1 + (16 * tile + r) * MLP_IN + 4 * q + i];
  } else if (tile < MLP_WG_T_W4) { // w2/w3[t][s]: W[16t + r][16s + 4q + i]
    const int off = tile < MLP_WG_T_W3 ? MLP_OFF_W2 : MLP_OFF_W3;
    const int ts = tile - (tile < MLP_WG_T_W3 ? MLP_WG_T_W2 : MLP_WG_T_W3);
    const int t = ts >> 2, sl = ts & 3;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      v[i] = theta[off + (16 * t + r) * MLP_H + 16 * sl + 4 * q + i];
  } else if (tile < MLP_WG_T_W2T) { // w4[s]: row 0 only, W4[16s + 4q + i]
    if (r == 0)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        v[i] = theta[MLP_OFF_W4 + 16 * (tile - MLP_WG_T_W4) + 4 * q + i];
  } else { // w2ᵀ/w3ᵀ[t][s]: A[16t + r][16s + 4q + i] = W[16s + 4q + i][16t + r]
    const int off = tile < MLP_WG_T_W3T ? MLP_OFF_W2 : MLP_OFF_W3;
    const int ts = tile - (tile < MLP_WG_T_W3T ? MLP_WG_T_W2T : MLP_WG_T_W3T);
    const int t = ts >> 2, sl = ts & 3;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      v[i] = theta[off + (16 * sl + 4 * q + i) * MLP_H + 16 * t + r];
  }
  return bf16_of(v);
}

__device__ __forceinline__ bool bf16_positive(short h) { return h > 0; }

// Pre-activation gradient tile: g ⊙ [h > 0] (h the layer's bf16 output), as bf16.
__device__ __forceinline__ bf16x4 masked_bf16(f32x4 g, bf16x4 h) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
    g[i] = bf16_positive(h[i]) ? g[i] : 0.f;
  return bf16_of(g);
}

// Stores a D-layout bf16 tile group (features 16t + 4q + i of point n = r) feature-major.
__device__ __forceinline__ void store_feature_major(short *__restrict__ out, int64_t ld,
                                                    int64_t n, int q, const bf16x4 (&v)[4]) {
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      out[(int64_t)(16 * t + 4 * q + i) * ld + n] = v[t][i];
}

struct MlpTrainBufs {
  int B;      // batch points (the loss is the mean over these)
  int64_t ld; // padded batch: a multiple of MLP_GRAD_CHUNK (points beyond B are zero-gradient)
  const float *xb, *yb; // normalised inputs [B][8], targets [B]
  short *xT, *h1T, *h2T, *h3T, *dz1T, *dz2T, *dz3T; // bf16 feature-major [features][ld]
  float *dy;  // [ld]
  float *sq;  // per-point squared error [B]
};

__global__ void __launch_bounds__(256) k_mlp_fwd_bwd(MlpTrainBufs b,
                                                     const float *__restrict__ theta) {
  __shared__ bf16x4 wl[MLP_WG_TILES][64];
  __shared__ f32x4 vl[MLP_WG_VECS][64];
  for (int e = threadIdx.x; e < MLP_WG_TILES * 64; e += 256)
    wl[e >> 6][e & 63] = mlp_operand(theta, e >> 6, e & 63);
  for (int e = threadIdx.x; e < MLP_WG_VECS * 64; e += 256) {
    const int v = e >> 6, l = e & 63, t = v & 3;
    const int off = v < 4 ? MLP_OFF_B1 : v < 8 ? MLP_OFF_B2 : v < 12 ? MLP_OFF_B3 : MLP_OFF_W4;
    vl[v][l] = *(const f32x4 *)(theta + off + 16 * t + 4 * (l >> 4));
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
  const int64_t tiles = b.ld / 16;
  const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wave >= tiles)
    return;
  const float b4 = theta[MLP_OFF_B4];
  const float invB = 2.0f / (float)b.B;
  for (int64_t tile = wave; tile < tiles; tile += (int64_t)gridDim.x * 4) {
    const int64_t n = tile * 16 + r;
    const bool real = n < b.B;
    // Forward (the arithmetic of mlp_tile, weights from LDS).
    bf16x4 x = {0, 0, 0, 0};
    if (q < 2 && real)
      x = bf16_of(*(const f32x4 *)(b.xb + n * MLP_IN + 4 * q));
    bf16x4 h1[4], h2[4], h3[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
      h1[t] = relu_bf16(__builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wl[MLP_WG_T_W1 + t][lane], x,
                                                                   vl[t][lane], 0, 0, 0));
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
      acc[t] = vl[4 + t][lane];
#pragma unroll
    for (int sl = 0; sl < 4; ++sl)
#pragma unroll
      for (int t = 0; t < 4; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wl[MLP_WG_T_W2 + 4 * t + sl][lane],
                                                           h1[sl], acc[t], 0, 0, 0);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      h2[t] = relu_bf16(acc[t]);
      acc[t] = vl[8 + t][lane];
    }
#pragma unroll
    for (int sl = 0; sl < 4; ++sl)
#pragma unroll
      for (int t = 0; t < 4; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wl[MLP_WG_T_W3 + 4 * t + sl][lane],
                                                           h2[sl], acc[t], 0, 0, 0);
#pragma unroll
    for (int t = 0; t < 4; ++t)
      h3[t] = relu_bf16(acc[t]);
    f32x4 out = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int sl = 0; sl < 4; ++sl)
      out = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wl[MLP_WG_T_W4 + sl][lane], h3[sl], out,
                                                      0, 0, 0);
    // Loss gradient of point r (lanes 0-15), then to every lane of that point.
    float dyr = 0.f;
    if (q == 0 && real) {
      const float e = (out[0] + b4) - b.yb[n];
      dyr = invB * e;
      b.sq[n] = e * e;
    }
    if (q == 0)
      b.dy[n] = dyr;
    dyr = __shfl(dyr, r, 64);
    // Backward.
    bf16x4 d3[4], d2[4], d1[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const f32x4 w4 = vl[12 + t][lane];
      f32x4 g;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        g[i] = w4[i] * dyr;
      d3[t] = masked_bf16(g, h3[t]);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
      acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int sl = 0; sl < 4; ++sl)
#pragma unroll
      for (int t = 0; t < 4; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wl[MLP_WG_T_W3T + 4 * t + sl][lane],
                                                           d3[sl], acc[t], 0, 0, 0);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      d2[t] = masked_bf16(acc[t], h2[t]);
      acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int sl = 0; sl < 4; ++sl)
#pragma unroll
      for (int t = 0; t < 4; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wl[MLP_WG_T_W2T + 4 * t + sl][lane],
                                                           d2[sl], acc[t], 0, 0, 0);
#pragma unroll
    for (int t = 0; t < 4; ++t)
      d1[t] = masked_bf16(acc[t], h1[t]);
    // Layer inputs and gradients, feature-major, for the weight-gradient GEMMs.
    if (q < 2)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        b.xT[(int64_t)(4 * q + i) * b.ld + n] = x[i];
    store_feature_major(b.h1T, b.ld, n, q, h1);
    store_feature_major(b.h2T, b.ld, n, q, h2);
    store_feature_major(b.h3T, b.ld, n, q, h3);
    store_feature_major(b.dz1T, b.ld, n, q, d1);
    store_feature_major(b.dz2T, b.ld, n, q, d2);
    store_feature_major(b.dz3T, b.ld, n, q, d3);
  }
}

// A K-slice operand of the gradient GEMMs: features 16·tile + r, points p0 + 4q .. + 3 of a
// feature-major bf16 array (8 bytes, aligned).
__device__ __forceinline__ bf16x4 fm_operand(const short *__restrict__ a, int64_t ld, int tile,
                                             int r, int q, int64_t p0) {
  return *(const bf16x4 *)(a + (int64_t)(16 * tile + r) * ld + p0 + 4 * q);
}

// Partial gradients of one 64-point chunk per wavefront; blockIdx.y picks the parameters:
// 0 = W3, b3; 1 = W2, b2; 2 = W1, b1, W4, b4. part[chunk][MLP_PARAMS].
__global__ void __launch_bounds__(256) k_mlp_wgrad(MlpTrainBufs b, float *__restrict__ part) {
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
  const int64_t chunk = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (chunk * MLP_GRAD_CHUNK >= b.ld)
    return;
  const int group = blockIdx.y;
  const bf16x4 ones = r == 0 ? bf16x4{0x3F80, 0x3F80, 0x3F80, 0x3F80} : bf16x4{0, 0, 0, 0};
  float *out = part + chunk * MLP_PARAMS;
  f32x4 gw[4][4], gb[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    gb[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int u = 0; u < 4; ++u)
      gw[t][u] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  if (group < 2) { // gW[m][k] = Σ_p dz[m][p] h[k][p]; gb[m] = Σ_p dz[m][p]
    const short *dz = group == 0 ? b.dz3T : b.dz2T;
    const short *h = group == 0 ? b.h2T : b.h1T;
    for (int k = 0; k < MLP_GRAD_CHUNK / 16; ++k) {
      const int64_t p0 = chunk * MLP_GRAD_CHUNK + 16 * k;
      bf16x4 a[4], bb[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        a[t] = fm_operand(dz, b.ld, t, r, q, p0);
        bb[t] = fm_operand(h, b.ld, t, r, q, p0);
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
          gw[t][u] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t], bb[u], gw[t][u], 0, 0, 0);
        gb[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t], ones, gb[t], 0, 0, 0);
      }
    }
    const int offW = group == 0 ? MLP_OFF_W3 : MLP_OFF_W2;
    const int offB = group == 0 ? MLP_OFF_B3 : MLP_OFF_B2;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = 16 * t + 4 * q + i;
#pragma unroll
        for (int u = 0; u < 4; ++u)
          out[offW + m * MLP_H + 16 * u + r] = gw[t][u][i];
        if (r == 0)
          out[offB + m] = gb[t][i];
      }
    return;
  }
  // gW1[m][k] = Σ_p dz1[m][p] x[k][p] (k < 8), gb1; gW4[k] = Σ_p dy[p] h3[k][p], gb4.
  f32x4 g4[4], gb4 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int u = 0; u < 4; ++u)
    g4[u] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k = 0; k < MLP_GRAD_CHUNK / 16; ++k) {
    const int64_t p0 = chunk * MLP_GRAD_CHUNK + 16 * k;
    const bf16x4 xs = r < MLP_IN ? fm_operand(b.xT, b.ld, 0, r, q, p0) : bf16x4{0, 0, 0, 0};
    f32x4 dv = {0.f, 0.f, 0.f, 0.f};
    if (r == 0)
      dv = *(const f32x4 *)(b.dy + p0 + 4 * q);
    const bf16x4 dyA = bf16_of(dv); // row 0 of the A operand: dy of the K-slice's points
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const bf16x4 a = fm_operand(b.dz1T, b.ld, t, r, q, p0);
      gw[t][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, xs, gw[t][0], 0, 0, 0);
      gb[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, ones, gb[t], 0, 0, 0);
      g4[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(dyA, fm_operand(b.h3T, b.ld, t, r, q, p0),
                                                        g4[t], 0, 0, 0);
    }
    gb4 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(dyA, ones, gb4, 0, 0, 0);
  }
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = 16 * t + 4 * q + i;
      if (r < MLP_IN)
        out[MLP_OFF_W1 + m * MLP_IN + r] = gw[t][0][i];
      if (r == 0)
        out[MLP_OFF_B1 + m] = gb[t][i];
    }
  if (q == 0) { // row m = 0 of the W4 and b4 products: lanes 0-15, component 0
#pragma unroll
    for (int t = 0; t < 4; ++t)
      out[MLP_OFF_W4 + 16 * t + r] = g4[t][0];
    if (r == 0)
      out[MLP_OFF_B4] = gb4[0];
  }
}

// grad[p] = Σ_chunks part[c][p] in chunk order; with `adam`, the Adam update of k_mlp_adam.
__global__ void __launch_bounds__(256) k_mlp_reduce_adam(int chunks,
                                                         const float *__restrict__ part,
                                                         float *__restrict__ grad,
                                                         float *__restrict__ theta,
                                                         float *__restrict__ m,
                                                         float *__restrict__ v, int adam,
                                                         float lr, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_PARAMS)
    return;
  float g = 0.f;
  for (int c = 0; c < chunks; ++c)
    g += part[(int64_t)c * MLP_PARAMS + p];
  grad[p] = g;
  if (!adam)
    return;
  const float mi = 0.9f * m[p] + 0.1f * g;
  const float vi = 0.999f * v[p] + 0.001f * g * g;
  m[p] = mi;
  v[p] = vi;
  theta[p] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

// Grid of the inference kernels: 2 four-wavefront workgroups per CU (the kernels' register
// use allows 2 wavefronts per SIMD), each wavefront walking many 16-point tiles so that the
// weights are loaded once per wavefront; at most one wavefront per tile.
inline int mlp_grid(int64_t points) {
  const int64_t tiles = (points + 15) / 16;
  const int64_t blocks = (tiles + 3) / 4;
  return (int)(blocks < 512 ? (blocks > 0 ? blocks : 1) : 512);
}

} // namespace vpd
