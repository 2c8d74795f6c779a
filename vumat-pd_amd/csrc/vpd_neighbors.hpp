// Cell-list neighbour search on the GPU, emitting CSR rows sorted ascending.
//
// Result contract (PdCorrespondenceSkeleton.cpp:30-39, 95-129): row i holds every j != i
// with sqrt((dx² + dy²) + dz²) <= δ + tol, dx = x_j − x_i, tol = 64 ε max(δ, max|x|),
// ascending. The reference bins by floor(x / (δ + 2 tol)) and scans the 27 surrounding
// cells; any binning whose cell width is at least the search radius finds the same set,
// so the GPU bins on the same width but in a dense grid offset by the lowest cell.
//
//   k_bounds      : max|x|, lowest/highest cell, non-finite flag (one pass, atomics)
//   k_cell_count  : particles per cell
//   scan          : cell starts (exclusive)
//   k_cell_fill   : particle ids into cell slots (unordered within a cell)
//   k_row_count   : neighbours per particle            -> scan -> rowptr
//   k_row_fill    : neighbour ids into each row (unordered)
//   k_row_sort    : one single-wave workgroup per row, bitonic sort in LDS (rows <= 1024),
//                   insertion sort by one lane for longer rows
#pragma once

#include "vpd_math.hpp"

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vpd {

struct Bounds {
  unsigned long long maxAbsBits; // max |x| as the bit pattern of a non-negative double
  long long cmin[3];
  long long cmax[3];
  int nonfinite;
};

__device__ __forceinline__ long long cell_of(double x, double cw) {
  return (long long)floor(x / cw);
}

__global__ void __launch_bounds__(256) k_abs_max(int64_t n, const double *__restrict__ x,
                                                 Bounds *b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double m = 0.0;
  int bad = 0;
  if (i < n) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double v = x[d * n + i];
      if (!isfinite(v))
        bad = 1;
      else
        m = fmax(m, fabs(v));
    }
  }
  // Non-negative doubles order as their bit patterns.
  for (int off = 32; off > 0; off >>= 1) {
    m = fmax(m, __shfl_xor(m, off));
    bad |= __shfl_xor(bad, off);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&b->maxAbsBits, (unsigned long long)__double_as_longlong(m));
    if (bad)
      atomicOr(&b->nonfinite, 1);
  }
}

__global__ void __launch_bounds__(256) k_cell_bounds(int64_t n, const double *__restrict__ x,
                                                     double cw, Bounds *b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const long long c = cell_of(x[d * n + i], cw);
    atomicMin(&b->cmin[d], c);
    atomicMax(&b->cmax[d], c);
  }
}

struct Grid {
  double cw, radius;
  long long c0[3];
  long long dim[3];
};

__device__ __forceinline__ long long cell_index(const Grid &g, long long cx, long long cy,
                                                long long cz) {
  return ((cx - g.c0[0]) * g.dim[1] + (cy - g.c0[1])) * g.dim[2] + (cz - g.c0[2]);
}

__global__ void __launch_bounds__(256) k_cell_count(int64_t n, const double *__restrict__ x,
                                                    Grid g, int *__restrict__ cellOf,
                                                    int *__restrict__ cellCount) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const long long c = cell_index(g, cell_of(x[i], g.cw), cell_of(x[n + i], g.cw),
                                 cell_of(x[2 * n + i], g.cw));
  cellOf[i] = (int)c;
  atomicAdd(&cellCount[c], 1);
}

__global__ void __launch_bounds__(256) k_cell_fill(int64_t n, const int *__restrict__ cellOf,
                                                   const int *__restrict__ cellStart,
                                                   int *__restrict__ cursor,
                                                   int *__restrict__ cellItems) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const int c = cellOf[i];
  const int slot = atomicAdd(&cursor[c], 1);
  cellItems[cellStart[c] + slot] = (int)i;
}

// The exact accept test of PdCorrespondenceSkeleton.cpp:122 (norm of x_j − x_i).
__device__ __forceinline__ bool within(const double *x, int64_t n, int64_t i, int64_t j,
                                       double xi0, double xi1, double xi2, double radius) {
  const double d0 = x[j] - xi0, d1 = x[n + j] - xi1, d2 = x[2 * n + j] - xi2;
  return sqrt(d0 * d0 + d1 * d1 + d2 * d2) <= radius;
}

template <bool FILL>
__global__ void __launch_bounds__(256)
    k_row_scan(int64_t n, const double *__restrict__ x, Grid g, const int *__restrict__ cellStart,
               const int *__restrict__ cellCount, const int *__restrict__ cellItems,
               int *__restrict__ rowCount, const int *__restrict__ rowptr, int *__restrict__ col) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const double xi0 = x[i], xi1 = x[n + i], xi2 = x[2 * n + i];
  const long long cx = cell_of(xi0, g.cw), cy = cell_of(xi1, g.cw), cz = cell_of(xi2, g.cw);
  int count = 0;
  int out = FILL ? rowptr[i] : 0;
  for (long long dx = -1; dx <= 1; ++dx) {
    const long long ax = cx + dx;
    if (ax < g.c0[0] || ax >= g.c0[0] + g.dim[0])
      continue;
    for (long long dy = -1; dy <= 1; ++dy) {
      const long long ay = cy + dy;
      if (ay < g.c0[1] || ay >= g.c0[1] + g.dim[1])
        continue;
      for (long long dz = -1; dz <= 1; ++dz) {
        const long long az = cz + dz;
        if (az < g.c0[2] || az >= g.c0[2] + g.dim[2])
          continue;
        const long long c = cell_index(g, ax, ay, az);
        const int s = cellStart[c], e = s + cellCount[c];
        for (int q = s; q < e; ++q) {
          const int j = cellItems[q];
          if (j != i && within(x, n, i, j, xi0, xi1, xi2, g.radius)) {
            if (FILL)
              col[out++] = j;
            else
              ++count;
          }
        }
      }
    }
  }
  if (!FILL)
    rowCount[i] = count;
}

// One single-wavefront workgroup sorts one row: bitonic network over the row padded with
// INT_MAX to a power of two, held in LDS. Rows longer than SORT_CAP fall back to an
// insertion sort by lane 0 directly in global memory.
constexpr int SORT_CAP = 1024;

__global__ void __launch_bounds__(64) k_row_sort(int64_t n, const int *__restrict__ rowptr,
                                                 int *__restrict__ col) {
  __shared__ int s[SORT_CAP];
  const int lane = threadIdx.x;
  const int64_t row = blockIdx.x;
  if (row >= n)
    return;
  const int b = rowptr[row], e = rowptr[row + 1], m = e - b;
  if (m <= 1)
    return;
  if (m > SORT_CAP) {
    if (lane == 0)
      for (int p = b + 1; p < e; ++p) {
        const int key = col[p];
        int q = p - 1;
        while (q >= b && col[q] > key) {
          col[q + 1] = col[q];
          --q;
        }
        col[q + 1] = key;
      }
    return;
  }
  int size = 1;
  while (size < m)
    size <<= 1;
  for (int t = lane; t < size; t += 64)
    s[t] = t < m ? col[b + t] : 0x7fffffff;
  __syncthreads();
  for (int k = 2; k <= size; k <<= 1)
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int t = lane; t < size; t += 64) {
        const int partner = t ^ jj;
        if (partner > t) {
          const int a = s[t], c = s[partner];
          const bool up = (t & k) == 0;
          if ((a > c) == up) {
            s[t] = c;
            s[partner] = a;
          }
        }
      }
      __syncthreads();
    }
  for (int t = lane; t < m; t += 64)
    col[b + t] = s[t];
}

// Exclusive scan of int counts: per-block scan (1024 items), block totals scanned by one
// block in a serial sweep, then the block offsets added. out may alias in.
constexpr int SCAN_ITEMS = 1024;

__global__ void __launch_bounds__(256) k_scan_block(int64_t n, const int *__restrict__ in,
                                                    int64_t *__restrict__ out,
                                                    int64_t *__restrict__ blockTotal) {
  __shared__ int64_t s[256];
  const int64_t base = (int64_t)blockIdx.x * SCAN_ITEMS + threadIdx.x * 4;
  int64_t v[4], sum = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = base + k < n ? in[base + k] : 0;
    sum += v[k];
  }
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int64_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  int64_t run = s[threadIdx.x] - sum;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < n)
      out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 255)
    blockTotal[blockIdx.x] = s[255];
}

__global__ void __launch_bounds__(256) k_scan_totals(int64_t nblocks, int64_t *blockTotal,
                                                     int64_t *grand) {
  __shared__ int64_t s[256];
  __shared__ int64_t carry;
  if (threadIdx.x == 0)
    carry = 0;
  __syncthreads();
  for (int64_t start = 0; start < nblocks; start += 256) {
    const int64_t i = start + threadIdx.x;
    const int64_t v = i < nblocks ? blockTotal[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      const int64_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nblocks)
      blockTotal[i] = carry + s[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 255)
      carry += s[255];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    *grand = carry;
}

// Adds block offsets and narrows to int32 (callers check the grand total fits first).
__global__ void __launch_bounds__(256) k_scan_add(int64_t n, const int64_t *__restrict__ part,
                                                  const int64_t *__restrict__ blockTotal,
                                                  int *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  out[i] = (int)(part[i] + blockTotal[i / SCAN_ITEMS]);
}

} // namespace vpd
