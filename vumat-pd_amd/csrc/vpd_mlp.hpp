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

// Return-map seeds for n points (SEED_SLOTS: for the `*count` points of a compacted list;
// slot s holds particle list[s] and its strain increments at inc[c·incStride + s]). A point
// whose previous plastic increment (state 5, component 4 of the JC state planes, stride
// stateStride) is positive keeps it as its seed: on a continuing plastic path it is closer to
// the solution than the network's estimate (about 2% median error). A point with state 5 = 0
// (first yield, or reloading after an elastic increment), where that seed says nothing, takes
// the network's estimate from the 6 strain increments, ε̄ᵖ and T (components 1 and 2):
// seed[s] = max(0, y·std + mean).
// Each wavefront walks 16-point tiles with a grid stride, reading state 5 first; a tile whose
// points all carry a state-5 seed skips the network and its inputs, and a wavefront loads the
// weights at its first tile that needs them. Wavefronts with no tile exit at once, and the
// tile loop (and the skip) is uniform per wavefront, so every MFMA runs with all lanes active.
enum SeedForm { SEED_ALL = 0, SEED_SLOTS = 1 };
template <int FORM>
__global__ void __launch_bounds__(256) k_mlp_seed(int64_t nPoints, const int *__restrict__ count,
                                                  const int *__restrict__ list,
                                                  const double *__restrict__ inc,
                                                  int64_t incStride,
                                                  const double *__restrict__ state,
                                                  int64_t stateStride,
                                                  const float *__restrict__ theta,
                                                  const float *__restrict__ norm,
                                                  double *__restrict__ seed) {
  constexpr bool LISTED = FORM == SEED_SLOTS;
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
  const int64_t n = LISTED ? (int64_t)*count : nPoints;
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
    const int64_t i = s < n ? (LISTED ? (int64_t)list[s] : s) : 0;
    const double prev = q == 0 && s < n ? state[4 * stateStride + i] : 1.0; // state 5
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
          v[2] = state[1 * stateStride + i];
          v[3] = state[2 * stateStride + i];
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

// Grid of the inference kernels: 2 four-wavefront workgroups per CU (the kernels' register
// use allows 2 wavefronts per SIMD), each wavefront walking many 16-point tiles so that the
// weights are loaded once per wavefront; at most one wavefront per tile.
inline int mlp_grid(int64_t points) {
  const int64_t tiles = (points + 15) / 16;
  const int64_t blocks = (tiles + 3) / 4;
  return (int)(blocks < 512 ? (blocks > 0 ? blocks : 1) : 512);
}

} // namespace vpd
