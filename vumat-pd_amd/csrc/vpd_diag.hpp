// Per-step run diagnostics of BoundaryValueCaseRunner.cpp:231-270, fused into the material and
// force kernels. Each kernel computes its per-particle terms in registers (the particle's
// stress, strain, state, force and velocity are already there), reduces them over its
// 256-particle chunk in a fixed tree order and writes one partial row per chunk; a one-block
// kernel then sums the rows in chunk order into the step's record. No extra pass over the
// particle state, and the result does not depend on the launch tuning.
#pragma once

#include "vpd_math.hpp"

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vpd {

constexpr int DIAG_STRIDE = 16;

// Record slots. The order matches enum vpd_diag_slot of the C API.
enum DiagSlot {
  D_INTERNAL_ENERGY = 0, // Σ contraction(½(σ̄_n + σ_{n+1}), E_{n+1} − E_n)·V   (:231-237)
  D_OBS_STRESS = 1,      // Σ σ_{n+1}[axis,axis]·w over the observation window  (:250-253)
  D_OBS_STRAIN = 2,      // Σ E_{n+1}[axis,axis]·w
  D_MAX_EQPS = 3,        // max(0, statev[eqps])                                (:256-258)
  D_MAX_INDICATOR = 4,   // max(0, internal indicator)                          (:259-260)
  D_NONFINITE_MATERIAL = 5, // 1 if a stress, strain or state value is not finite (:261-266)
  D_NET_X = 6,           // Σ f·V                                               (:248-249)
  D_NET_Y = 7,
  D_NET_Z = 8,
  D_REACTION = 9,        // Σ over the load band of (ρ a_ramp − f[axis])·V      (:268-271)
  D_KINETIC = 10,        // Σ ½ρV|v|² after the step's correction               (:276, :75-83)
  D_LATERAL_NEG = 11,    // Σ u[lateral] over the negative lateral face         (:283-285)
  D_LATERAL_POS = 12,    // Σ u[lateral] over the positive lateral face
  D_NONFINITE_FORCE = 13, // 1 if a force component is not finite
  D_COUNT = 14
};

__host__ __device__ constexpr bool diag_is_max(int s) {
  return s == D_MAX_EQPS || s == D_MAX_INDICATOR || s == D_NONFINITE_MATERIAL ||
         s == D_NONFINITE_FORCE;
}

// std::max's selection rule, (a < b) ? b : a: a NaN right operand never wins.
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }

struct DiagArgs {
  double *partials;     // [chunks][DIAG_STRIDE]; null: diagnostics off
  // Batched specimens: chunk c covers particles [chunkFirst[c], chunkEnd[c]), all of one
  // specimen and laid out from its first particle as that specimen alone lays out its chunks
  // (null: chunk c covers particles 256c .. 256c+255 of the block).
  const int64_t *chunkFirst, *chunkEnd;
  int64_t chunks;
  Planes dstrain;       // E_{n+1} − E_n (6 planes) between the two halves of a split update
  const double *vol;    // particle volumes
  const double *obsW;   // observation weights (0 outside the window)
  const uint8_t *latSet; // bit 0: negative lateral face, bit 1: positive lateral face
  int axis, lateral, eqpsIndex, nstatev;
  int indCount, indMin, indOneMinus;
  int ind[8];
};

// The particle this thread reduces for chunk `chunk` (valid: it exists).
__device__ __forceinline__ int64_t diag_particle(const DiagArgs &d, int64_t chunk, int64_t n,
                                                 bool &valid) {
  if (d.chunkFirst) {
    const int64_t i = d.chunkFirst[chunk] + threadIdx.x;
    valid = i < d.chunkEnd[chunk];
    return i;
  }
  const int64_t i = chunk * 256 + threadIdx.x;
  valid = i < n;
  return i;
}

// Green-Lagrange strain ½(FᵀF − I) in the operation order of CorotationalKinematics.cpp:141-149.
__device__ __forceinline__ T3 gl_strain(const T3 &f) {
  T3 e = t3_mul(t3_transpose(f), f);
  e.m[0] -= 1.0;
  e.m[4] -= 1.0;
  e.m[8] -= 1.0;
#pragma unroll
  for (int c = 0; c < 9; ++c)
    e.m[c] = e.m[c] * 0.5;
  return e;
}

// The strain-increment components the energy contraction reads: tensor entries 0, 4, 8, 1, 2, 5.
__device__ __forceinline__ void strain_increment(const T3 &e1, const T3 &e0, double d[6]) {
  const int idx[6] = {0, 4, 8, 1, 2, 5};
#pragma unroll
  for (int c = 0; c < 6; ++c)
    d[c] = e1.m[idx[c]] - e0.m[idx[c]];
}

// internalIndicator (BoundaryValueCaseRunner.cpp:95-110): min or max over the chosen state
// entries, optionally 1 − value; 0 with no entries.
__device__ __forceinline__ double internal_indicator(const DiagArgs &d, const double *st) {
  if (d.indCount == 0)
    return 0.0;
  double value = d.indMin ? __builtin_inf() : -__builtin_inf();
  for (int k = 0; k < d.indCount; ++k) {
    const double item = st[d.ind[k]];
    value = d.indMin ? ((item < value) ? item : value) : dmax(value, item);
  }
  if (d.indOneMinus)
    value = 1.0 - value;
  return value;
}

// Material-side terms of one particle into v[0..6) (slots D_INTERNAL_ENERGY..
// D_NONFINITE_MATERIAL). sOld: the stored corotational σ̄_n (Voigt) — the reference averages
// it with the spatial σ_{n+1}, and so does this; dE: E_{n+1} − E_n as strain_increment gives.
__device__ __forceinline__ void material_diag(const DiagArgs &d, int64_t i, const double sOld[6],
                                              const T3 &sigma, const T3 &e1, const double dE[6],
                                              const double *st, double v[6]) {
  const T3 sp = from_voigt(sOld);
  const int idx[6] = {0, 4, 8, 1, 2, 5};
  double a[6];
#pragma unroll
  for (int c = 0; c < 6; ++c)
    a[c] = 0.5 * (sp.m[idx[c]] + sigma.m[idx[c]]);
  const double contraction =
      a[0] * dE[0] + a[1] * dE[1] + a[2] * dE[2] + 2.0 * (a[3] * dE[3] + a[4] * dE[4] + a[5] * dE[5]);
  v[D_INTERNAL_ENERGY] = contraction * d.vol[i];
  const double w = d.obsW ? d.obsW[i] : 0.0;
  v[D_OBS_STRESS] = sigma.m[4 * d.axis] * w;
  v[D_OBS_STRAIN] = e1.m[4 * d.axis] * w;
  v[D_MAX_EQPS] = d.eqpsIndex >= 0 ? dmax(0.0, st[d.eqpsIndex]) : 0.0;
  v[D_MAX_INDICATOR] = dmax(0.0, internal_indicator(d, st));
  bool finite = true;
#pragma unroll
  for (int c = 0; c < 9; ++c)
    finite = finite && __builtin_isfinite(sigma.m[c]) && __builtin_isfinite(e1.m[c]);
  for (int c = 0; c < d.nstatev; ++c)
    finite = finite && __builtin_isfinite(st[c]);
  v[D_NONFINITE_MATERIAL] = finite ? 0.0 : 1.0;
}

// Block-wide reduction of NS per-thread values in a fixed tree order (sum, or dmax for the max
// slots); thread s < NS writes slot FIRST + s to out[s]. Every thread of the block calls it.
template <int NS, int FIRST>
__device__ __forceinline__ void diag_block_store(const double (&v)[NS], double *out) {
  __shared__ double red[NS][256];
  const int t = threadIdx.x;
#pragma unroll
  for (int s = 0; s < NS; ++s)
    red[s][t] = v[s];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int s = 0; s < NS; ++s)
        red[s][t] = diag_is_max(FIRST + s) ? dmax(red[s][t], red[s][t + w])
                                           : red[s][t] + red[s][t + w];
    }
    __syncthreads();
  }
  if (t < NS)
    out[t] = red[t][0];
  __syncthreads(); // red is reused by the next sweep of the calling kernel
}

// Sums the chunk rows in chunk order (each thread a strided run, then a fixed tree) into one
// step record. One block of 256 threads.
__device__ __forceinline__ void diag_reduce_rows(const double *__restrict__ partials,
                                                 int64_t chunks, double *record);
__global__ void __launch_bounds__(256) k_diag_reduce(const double *__restrict__ partials,
                                                     int64_t chunks, double *record) {
  diag_reduce_rows(partials, chunks, record);
}

// One record per specimen: block s sums its chunks [specChunk[s], specChunk[s+1]) as
// k_diag_reduce sums a single specimen's, into records[s].
__global__ void __launch_bounds__(256) k_diag_reduce_specimens(
    const double *__restrict__ partials, const int64_t *__restrict__ specChunk,
    double *records) {
  const int64_t s = blockIdx.x, c0 = specChunk[s];
  diag_reduce_rows(partials + c0 * DIAG_STRIDE, specChunk[s + 1] - c0,
                   records + s * DIAG_STRIDE);
}

__device__ __forceinline__ void diag_reduce_rows(const double *__restrict__ partials,
                                                 int64_t chunks, double *record) {
  __shared__ double red[D_COUNT][256];
  const int t = threadIdx.x;
  double acc[D_COUNT];
#pragma unroll
  for (int s = 0; s < D_COUNT; ++s)
    acc[s] = 0.0;
  for (int64_t c = t; c < chunks; c += 256) {
#pragma unroll
    for (int s = 0; s < D_COUNT; ++s) {
      const double x = partials[c * DIAG_STRIDE + s];
      acc[s] = diag_is_max(s) ? dmax(acc[s], x) : acc[s] + x;
    }
  }
#pragma unroll
  for (int s = 0; s < D_COUNT; ++s)
    red[s][t] = acc[s];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int s = 0; s < D_COUNT; ++s)
        red[s][t] = diag_is_max(s) ? dmax(red[s][t], red[s][t + w]) : red[s][t] + red[s][t + w];
    }
    __syncthreads();
  }
  if (t < DIAG_STRIDE)
    record[t] = t < D_COUNT ? red[t][0] : 0.0;
}

} // namespace vpd
