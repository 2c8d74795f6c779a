// gfx950 kernels of the explicit correspondence-PD increment. One lane owns one particle
// (or one material point); particle fields are SoA planes of FP64 ([component][N]), so a
// wavefront of 64 consecutive particles reads each plane with unit stride, and the
// neighbour gathers of a lattice-ordered block (row p+1 ≈ row p shifted by one index)
// stay close to unit stride as well. Bond sums run over each CSR row in order, one lane
// per row, which keeps the reference's summation order and needs no atomics.
#pragma once

#include "vpd_diag.hpp"
#include "vpd_math.hpp"
#include "vpd_mlp.hpp"

#include <hip/hip_runtime.h>
#include <climits>
#include <stdint.h>

namespace vpd {

__device__ __forceinline__ void raise_error(int *err, int bit) { atomicOr(err, bit); }

// Bond lists for the per-step passes in a 64-wide sliced layout grouped by 4 bonds: bond k
// of particle i = 64 s + L sits at base[s] + 256 (k / 4) + 4 L + k % 4. One 16-byte load gives
// a lane the column indices of 4 consecutive bonds (the lanes of a wavefront read 1 KiB
// contiguous), so a lane issues the data loads of 4 bonds at once; the flag bits of a 4-bond
// group are contiguous, and a flag word holds the groups of 8 lanes. A slice is padded to a
// multiple of 4 of its longest row with index 0 (a valid particle whose data is loaded and
// not used). Within a row the bonds keep their CSR (ascending) order.
struct Sell {
  const int64_t *base;
  const int *col;
  const int *rowptr;
};
__host__ __device__ __forceinline__ int64_t sell_slot(int64_t sliceBase, int64_t i, int64_t k) {
  return sliceBase + 256 * (k >> 2) + 4 * (i & 63) + (k & 3);
}
__device__ __forceinline__ int64_t sell_pos(const Sell &s, int64_t i, int k) {
  return sell_slot(s.base[i >> 6], i, k);
}
typedef int i4v __attribute__((ext_vector_type(4)));

// Fills the sliced layout from CSR (one lane per particle).
__global__ void __launch_bounds__(256) k_sell_fill(int64_t n, const int *__restrict__ rowptr,
                                                   const int *__restrict__ col,
                                                   const int64_t *__restrict__ base,
                                                   int *__restrict__ colS) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const int b = rowptr[i], e = rowptr[i + 1];
  for (int k = 0; k < e - b; ++k)
    colS[sell_slot(base[i >> 6], i, k)] = col[b + k];
}

// The static neighbour data of both bond passes as component pairs: (X_x, X_y), (X_z, V).
__global__ void __launch_bounds__(256) k_pack_xv(int64_t n, Planes x,
                                                 const double *__restrict__ vol, PairPlanes xv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  d2v a, b;
  a.x = x(0, i);
  a.y = x(1, i);
  b.x = x(2, i);
  b.y = vol[i];
  xv(0, i) = a;
  xv(1, i) = b;
}

// Workgroups are handed out to the 8 XCDs round-robin by workgroup id. Remapping the id so
// that each XCD walks one contiguous eighth of the particle range keeps every XCD's
// neighbour working set (a few lattice planes) inside its own L2.
__device__ __forceinline__ int64_t xcd_block(int64_t nblocks) {
  const int64_t b = blockIdx.x;
  const int64_t per = (nblocks + 7) / 8;
  const int64_t xcd = b % 8, slot = b / 8;
  const int64_t mapped = xcd * per + slot;
  return mapped < nblocks ? mapped : -1;
}

// The bond passes walk the particle range in `rounds` consecutive sweeps. Within a sweep
// every XCD works on one contiguous stretch, so the neighbour data an XCD touches at once is
// that stretch plus one horizon either side: with enough rounds it fits the XCD's 4 MB L2,
// while every sweep still fills the chip. The autotuner picks the number of rounds.
struct Rounds {
  int64_t n;
  int64_t blocksPerRound; // 256-particle workgroups per sweep
  int rounds;
  __device__ __forceinline__ int64_t particle(int r, int64_t blk) const {
    return ((int64_t)r * blocksPerRound + blk) * 256 + threadIdx.x;
  }
};

// Windowed neighbour gathers (the `_w` bond-pass kernels). A wavefront's 64 consecutive
// particles have neighbour rows that run over the same index ranges, shifted by one per
// lane, so most of the neighbour data a wavefront gathers lies in a few short runs of
// consecutive particles. The windowed kernels stage one such run — W consecutive particles
// from the smallest neighbour index any lane still needs — in the wavefront's LDS with one
// coalesced load per lane and component pair, let every lane consume its next bonds whose
// neighbour lies in the run, and move on. Each lane still walks its own row in CSR order
// with the arithmetic of the direct kernels, so results are identical bit for bit; only
// where the neighbour data comes from changes. A lane takes its next bonds only while their
// neighbour lies inside the window, from its start to its end: rows need not be sorted (a
// bond whose neighbour lies below the window start waits for a later window). The lane
// holding the smallest index always advances, so the loop ends.
//
// A lane may also stop after `take` bonds of a window and finish the rest in a later one
// (take ≥ 1; the lane holding the window's first index still advances). With rows made of
// runs of consecutive indices (a lattice numbered line by line) a window can hold one run of
// some lanes and one and a half of others, and the lanes with fewer bonds idle while the
// others finish; a cap near one run evens the lanes out at the price of more windows.
__device__ __forceinline__ int window_stop(int k, int take) {
  return k > INT_MAX - take ? INT_MAX : k + take;
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    v = min(v, __shfl_xor(v, off));
  return __builtin_amdgcn_readfirstlane(v); // the same in every lane: a scalar
}

// The wavefront's window stores (LDS-direct loads, vmcnt) and LDS accesses (lgkmcnt) are
// complete before it goes on, and the compiler moves no LDS access across this point.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_s_waitcnt(0x0070); // vmcnt(0), lgkmcnt(0); expcnt unconstrained
  __builtin_amdgcn_wave_barrier();
}
// The wavefront's LDS accesses are complete (its memory loads may still be in flight).
__device__ __forceinline__ void wave_lds_wait() {
  __builtin_amdgcn_s_waitcnt(0xC07F); // lgkmcnt(0); vmcnt and expcnt unconstrained
  __builtin_amdgcn_wave_barrier();
}

typedef __attribute__((address_space(3))) void lds_void;

// Window staging: each lane loads the 16-byte pair c of particle w0 + q0 + lane straight into
// LDS (no registers), at window slot q0 + lane of plane c. `planes` gives the pair planes. w0
// and `win` are the same for the whole wavefront, so the plane addresses are a scalar base
// plus the lane's offset and the LDS destinations are scalar.
// A pointer the whole wavefront holds, as a scalar (so that a load from it plus a lane
// offset takes the scalar-base form and no per-plane vector address is kept).
__device__ __forceinline__ const d2v *scalar_ptr(const d2v *p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return reinterpret_cast<const d2v *>(((uint64_t)hi << 32) | lo);
}

template <int PAIRS, int W, class Planeset>
__device__ __forceinline__ void stage_window(const Planeset &planes, int64_t n, int w0, d2v *win) {
  const int lane = threadIdx.x & 63;
  const int64_t left = n - (int64_t)w0; // particles from w0 to the end (≥ 1)
  const int count = left < W ? (int)left : W;
#pragma unroll
  for (int q0 = 0; q0 < W; q0 += 64) {
    if (q0 + lane < count) {
#pragma unroll
      for (int c = 0; c < PAIRS; ++c)
        __builtin_amdgcn_global_load_lds(
            (const void *)(scalar_ptr(planes.ptr(c, (int64_t)w0 + q0)) + lane),
            (lds_void *)(win + c * W + q0), 16, 0, 0);
    }
  }
}

// The wavefront's slot among the 4 of a 256-thread workgroup, as a scalar.
__device__ __forceinline__ int wave_slot() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// The pair planes of the two windowed passes, in window order.
struct DefgradPlanes {
  PairPlanes xv;
  UPairs u;
  __device__ __forceinline__ const d2v *ptr(int c, int64_t p) const {
    return c < 2 ? &xv(c, p) : reinterpret_cast<const d2v *>(u.p) + (c - 2) * u.n + p;
  }
};
struct ForcePlanes {
  PairPlanes xv, amap;
  __device__ __forceinline__ const d2v *ptr(int c, int64_t p) const {
    return c < 2 ? &xv(c, p) : &amap(c - 2, p);
  }
};

// Cursor over one lane's bonds in CSR order. The 16-byte column group (and flag word) of the
// next 4 bonds is loaded one group ahead, so that its latency is spent on the current group.
// Inside a group the columns and flag bits shift down one place per bond, so bond k's column
// is always `cols.x` and its flag bit 0 of `broken4`. `next` is bond k's column, or INT_MAX
// when the row is done.
struct BondCursor {
  int64_t g0;
  int k, len, next;
  i4v cols, colsAhead;
  uint32_t flagsAhead;
  uint32_t broken4; // flag bits of bond k and the rest of its group, bond k at bit 0
  __device__ __forceinline__ int64_t group_slot() const { return g0 + 64 * (int64_t)(k & ~3); }
  __device__ __forceinline__ void fetch_ahead(const Sell &b, const uint32_t *broken, bool damage,
                                              int groupStart) {
    if (groupStart < len) {
      const int64_t pg = g0 + 64 * (int64_t)groupStart;
      colsAhead = *(const i4v *)(b.col + pg);
      flagsAhead = damage ? broken[pg >> 5] : 0u;
    }
  }
  __device__ __forceinline__ void start(const Sell &b, const uint32_t *broken, bool damage,
                                        int64_t i) {
    k = 0;
    len = i >= 0 ? b.rowptr[i + 1] - b.rowptr[i] : 0;
    g0 = i >= 0 ? sell_pos(b, i, 0) : 0;
    next = INT_MAX;
    broken4 = 0u;
    if (len > 0) {
      cols = *(const i4v *)(b.col + g0);
      const uint32_t flags = damage ? broken[g0 >> 5] : 0u;
      broken4 = (flags >> (g0 & 31)) & 0xFu;
      next = cols.x;
      fetch_ahead(b, broken, damage, 4);
    }
  }
  __device__ __forceinline__ bool bond_broken() const { return broken4 & 1u; }
  // Bond k's bit in its flag word (a group's 4 slots are consecutive within one word).
  __device__ __forceinline__ uint32_t bit() const { return 1u << ((g0 + (k & 3)) & 31); }
  // Moves to bond k + 1; entering a new group takes the group loaded ahead and fetches the
  // one after it.
  __device__ __forceinline__ void advance(const Sell &b, const uint32_t *broken, bool damage) {
    ++k;
    if ((k & 3) == 0) {
      cols = colsAhead;
      broken4 = (flagsAhead >> (g0 & 31)) & 0xFu;
      fetch_ahead(b, broken, damage, k + 4);
    } else {
      cols.x = cols.y;
      cols.y = cols.z;
      cols.z = cols.w;
      broken4 >>= 1;
    }
    next = k < len ? cols.x : INT_MAX;
  }
};

// Per-component constraint codes: 2 bits per axis, 0 free, 1 held at zero, 2 follows the
// ramp (BoundaryConditions.cpp:81-115 collapsed to what its sequence of writes leaves).
struct Ramp {
  double u, v, a;
};

__device__ __forceinline__ void constrain(uint8_t code, const Ramp &r, double u[3], double v[3],
                                          double a[3]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int c = (code >> (2 * d)) & 3;
    if (c == 1) {
      u[d] = 0.0;
      v[d] = 0.0;
      a[d] = 0.0;
    } else if (c == 2) {
      u[d] = r.u;
      v[d] = r.v;
      a[d] = r.a;
    }
  }
}


// Specimens batched in one particle block (vpd_set_specimens): particle i belongs to specimen
// id[i], which has its own time step dt[s], its own end-face velocity v[s] (the velocity band
// u = v·t, v, a = 0 at t = step·dt[s], the expressions the single-specimen host uses) and
// stops after lastStep[s] steps. With `id` null the block is one specimen and the kernels
// take their scalar dt and ramps.
struct Batch {
  const uint16_t *id;
  const double *dt, *velocity;
  const int *lastStep;
  int step; // the step being taken (1-based; 0 for the initial constraints)
  // When set, the step is read from this device counter instead, so that one captured HIP
  // graph serves every step (the graph's force passes move the counters on).
  const int *stepDev;
  // Prescribed band kinematics (vpd_set_specimen_bands): specimen s with
  // bandOff[s] < bandOff[s+1] takes (u, v, a) of step k from band[3·(bandOff[s] + k)], the
  // values its host computed (e.g. the half-cosine ramp), instead of the velocity band.
  const double *band;
  const int64_t *bandOff;
  __device__ __forceinline__ int current() const { return stepDev ? *stepDev : step; }
};

struct PointStep {
  double dt;
  Ramp now, next;
  bool active, last; // last: the specimen's final step, which is not followed by a prediction
};

__device__ __forceinline__ bool point_active(const Batch &b, int64_t i) {
  return !b.id || b.current() <= b.lastStep[b.id[i]];
}

__device__ __forceinline__ PointStep point_step(const Batch &b, int64_t i, double dt,
                                                const Ramp &now, const Ramp &next) {
  PointStep p{dt, now, next, true, false};
  if (b.id) {
    const int s = b.id[i];
    p.dt = b.dt[s];
    const int step = b.current();
    p.active = step <= b.lastStep[s];
    p.last = step >= b.lastStep[s];
    const int64_t o = b.band ? b.bandOff[s] : 0;
    if (b.band && b.bandOff[s + 1] > o) {
      const int64_t k = o + (p.active ? step : b.lastStep[s]);
      const int64_t k1 = p.last ? k : k + 1;
      p.now = Ramp{b.band[3 * k], b.band[3 * k + 1], b.band[3 * k + 2]};
      p.next = Ramp{b.band[3 * k1], b.band[3 * k1 + 1], b.band[3 * k1 + 2]};
    } else {
      const double v = b.velocity[s];
      const double t = (double)step * p.dt, t1 = (double)(step + 1) * p.dt;
      p.now = Ramp{v * t, v, 0.0};
      p.next = Ramp{v * t1, v, 0.0};
    }
  }
  return p;
}

__global__ void k_set_step(int *stepDev, int step) {
  if (threadIdx.x == 0)
    *stepDev = step;
}


// Critical-stretch test of PdSimpleDamage.cpp:72-77 with a cheap exact filter.
// The reference decides stretch = (‖y‖ − ‖ξ‖)/‖ξ‖ > s_c (or < −f), with two square roots and a
// division per bond. With a = ‖ξ‖², b = ‖y‖² (the same sums the reference takes the roots
// of), that test lands on the same side as b ⋛ a(1+s_c)² whenever b/a is more than 1e-9
// away from the threshold ratio: the rounded stretch differs from the exact one by about
// 1e-15, far inside that margin (≥ 5e-10 in stretch). The filter may therefore round a and b
// its own way (fused multiply-adds, a few ulps from the exact sums); only bonds within the
// margin, and the degenerate a = 0, take the reference's exact expression.
struct StretchTest {
  double tLo, tHi; // (1+s_c)² (1 ∓ 1e-9)
  double cLo, cHi; // (1−f)² (1 ∓ 1e-9), used when the compression test is armed and f < 1
  double initiation, failure;
  int compression;
};

__host__ __device__ inline StretchTest make_stretch_test(double initiation, double failure,
                                                         int useCompression) {
  StretchTest t;
  const double up = (1.0 + initiation) * (1.0 + initiation);
  t.tLo = up * (1.0 - 1e-9);
  t.tHi = up * (1.0 + 1e-9);
  const double dn = (1.0 - failure) * (1.0 - failure);
  t.cLo = dn * (1.0 - 1e-9);
  t.cHi = dn * (1.0 + 1e-9);
  t.initiation = initiation;
  t.failure = failure;
  t.compression = useCompression ? (failure < 1.0 ? 1 : 2) : 0; // 2: armed but unreachable
  return t;
}

__device__ __forceinline__ bool stretch_breaks(const StretchTest &t, const double xi[3],
                                               const double yij[3]) {
  const double a = fma(xi[0], xi[0], fma(xi[1], xi[1], xi[2] * xi[2]));
  const double b = fma(yij[0], yij[0], fma(yij[1], yij[1], yij[2] * yij[2]));
  if (a > 0.0) {
    if (t.compression != 1) { // the tension test alone
      if (b <= a * t.tLo)
        return false;
      if (b >= a * t.tHi)
        return true;
    } else {
      const bool tensionDecided = b <= a * t.tLo || b >= a * t.tHi;
      const bool compressionDecided = b >= a * t.cHi || b <= a * t.cLo;
      if (tensionDecided && compressionDecided)
        return b >= a * t.tHi || b <= a * t.cLo;
    }
  }
  const double initial = sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]);
  const double stretch = (sqrt(yij[0] * yij[0] + yij[1] * yij[1] + yij[2] * yij[2]) - initial) /
                         initial;
  return stretch > t.initiation || (t.compression != 0 && stretch < -t.failure);
}

// ExplicitTimeIntegrator.cpp:23-34 then the constraints at t (BoundaryConditions.cpp:81).
__global__ void __launch_bounds__(256) k_predict(int64_t n, UPairs u, Planes v, Planes a,
                                                 const uint8_t *__restrict__ codes, double dtIn,
                                                 Ramp rIn, Batch batch) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const PointStep ps = point_step(batch, i, dtIn, rIn, rIn);
  if (!ps.active)
    return;
  const double dt = ps.dt;
  const Ramp &r = ps.now;
  double uu[3], vv[3], aa[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    vv[d] = v(d, i);
    aa[d] = a(d, i);
    uu[d] = u(d, i) + (dt * vv[d] + 0.5 * dt * dt * aa[d]);
  }
  constrain(codes[i], r, uu, vv, aa);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    u(d, i) = uu[d];
    v(d, i) = vv[d];
    a(d, i) = aa[d];
  }
}

__global__ void __launch_bounds__(256) k_constrain(int64_t n, UPairs u, Planes v, Planes a,
                                                   const uint8_t *__restrict__ codes, Ramp rIn,
                                                   Batch batch) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const uint8_t c = codes[i];
  if (c == 0)
    return;
  const PointStep ps = point_step(batch, i, 0.0, rIn, rIn);
  if (!ps.active)
    return;
  const Ramp &r = ps.now;
  double uu[3], vv[3], aa[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    uu[d] = u(d, i);
    vv[d] = v(d, i);
    aa[d] = a(d, i);
  }
  constrain(c, r, uu, vv, aa);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    u(d, i) = uu[d];
    v(d, i) = vv[d];
    a(d, i) = aa[d];
  }
}

// PdCorrespondenceSkeleton.cpp:149-170: K_i = Σ_j V_j ξ⊗ξ, stored as K_i⁻¹.
__global__ void __launch_bounds__(256) k_shape(int64_t n, Planes x, const double *__restrict__ vol,
                                               const int *__restrict__ rowptr,
                                               const int *__restrict__ col, Planes kinv,
                                               int *err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const double x0 = x(0, i), x1 = x(1, i), x2 = x(2, i);
  T3 k;
#pragma unroll
  for (int c = 0; c < 9; ++c)
    k.m[c] = 0.0;
  for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
    const int j = col[p];
    if (j < 0 || j >= n) {
      raise_error(err, ERR_NEIGHBOR_RANGE);
      continue;
    }
    const double xi[3] = {x(0, j) - x0, x(1, j) - x1, x(2, j) - x2};
    const double vj = vol[j];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
        k.m[3 * a + b] += vj * xi[a] * xi[b];
  }
  T3 ki;
  if (!t3_inverse_skeleton(k, ki)) {
    raise_error(err, ERR_SINGULAR_SHAPE);
#pragma unroll
    for (int c = 0; c < 9; ++c)
      ki.m[c] = 0.0;
  }
#pragma unroll
  for (int c = 0; c < 9; ++c)
    kinv(c, i) = ki.m[c];
}

// Bond pass 1. PdCorrespondenceSkeleton.cpp:172-197: F_i = (Σ_j V_j (y_j − y_i)⊗ξ) K_i⁻¹,
// and, with DAMAGE, PdSimpleDamage.cpp:63-84 on the same bond and the same predicted
// positions: a bond breaks (irreversibly) when its stretch exceeds `initiation` or, with a
// compression threshold armed, falls below −failure. Damage does not enter F (:181-196).
// Bonds run in groups of 4 (one 16-byte column load); each bond's X_j and V_j come as two
// component pairs, u_j as two more. A group's new breaks are set with one atomic OR on
// the flag word the lane shares with 7 others, and the particle is marked in anyBroken.
template <bool DAMAGE>
__device__ __forceinline__ void defgrad_one(int64_t i, Planes x, PairPlanes xv, UPairs u,
                                            Sell bonds, Planes kinv, Planes fout,
                                            uint32_t *broken, uint8_t *anyBroken,
                                            const StretchTest &test) {
  const double x0 = x(0, i), x1 = x(1, i), x2 = x(2, i);
  const double yi0 = x0 + u(0, i), yi1 = x1 + u(1, i), yi2 = x2 + u(2, i);
  double num[9];
#pragma unroll
  for (int c = 0; c < 9; ++c)
    num[c] = 0.0;
  const int len = bonds.rowptr[i + 1] - bonds.rowptr[i];
  const int64_t g0 = sell_pos(bonds, i, 0);
  for (int k0 = 0; k0 < len; k0 += 4) {
    const int64_t pg = g0 + 64 * (int64_t)k0; // slot of bond k0 (k0 is a multiple of 4)
    const i4v jj = *(const i4v *)(bonds.col + pg);
    const uint32_t flags = DAMAGE ? broken[pg >> 5] : 0u;
    uint32_t fresh = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int j = jj[q];
      const d2v a = xv(0, j), b = xv(1, j);
      const d2v ua = u.pair(0, j), ub = u.pair(1, j);
      const double uj0 = ua.x, uj1 = ua.y, uj2 = ub.x;
      if (k0 + q < len) {
        const double xi[3] = {a.x - x0, a.y - x1, b.x - x2};
        const double yij[3] = {(a.x + uj0) - yi0, (a.y + uj1) - yi1, (b.x + uj2) - yi2};
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c)
            num[3 * r + c] += b.y * yij[r] * xi[c];
        if (DAMAGE) {
          const uint32_t bit = 1u << ((pg + q) & 31);
          if (!(flags & bit) && stretch_breaks(test, xi, yij))
            fresh |= bit;
        }
      }
    }
    if (DAMAGE && fresh) {
      atomicOr(&broken[pg >> 5], fresh);
      anyBroken[i] = 1;
    }
  }
  T3 nm, ki;
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    nm.m[c] = num[c];
    ki.m[c] = kinv(c, i);
  }
  const T3 f = t3_mul(nm, ki);
#pragma unroll
  for (int c = 0; c < 9; ++c)
    fout(c, i) = f.m[c];
}

template <bool DAMAGE>
__global__ void __launch_bounds__(256)
    k_defgrad(Rounds rd, Planes x, PairPlanes xv, UPairs u, Sell bonds, Planes kinv, Planes fout,
              uint32_t *broken, uint8_t *anyBroken, double initiation, double failure,
              int useCompression, Batch batch) {
  const int64_t blk = xcd_block(rd.blocksPerRound);
  if (blk < 0)
    return;
  const StretchTest test = make_stretch_test(initiation, failure, useCompression);
  for (int r = 0; r < rd.rounds; ++r) {
    const int64_t i = rd.particle(r, blk);
    if (i >= rd.n)
      break;
    if (!point_active(batch, i))
      continue; // a finished specimen keeps its state
    defgrad_one<DAMAGE>(i, x, xv, u, bonds, kinv, fout, broken, anyBroken, test);
  }
}

// A = P K⁻¹ of particle i into its 5 component pairs.
__device__ __forceinline__ void store_amap(PairPlanes amap, int64_t i, const T3 &A) {
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    d2v v;
    v.x = A.m[2 * c];
    v.y = A.m[2 * c + 1];
    amap(c, i) = v;
  }
  d2v last;
  last.x = A.m[8];
  last.y = 0.0;
  amap(4, i) = last;
}

// Bond pass 1 with windowed neighbour gathers (see BondCursor): X_j, V_j and u_j of W
// consecutive particles per window, as 4 component pairs in the wavefront's LDS. A lane takes
// at most `take` bonds from one window (see window_stop). New breaks of a group are collected
// and set with one atomic OR when the lane leaves the group.
template <bool DAMAGE, int W>
__device__ __forceinline__ void defgrad_window(int64_t n, int64_t i, PairPlanes xv, UPairs u,
                                               Sell bonds, uint32_t *broken, uint8_t *anyBroken,
                                               const StretchTest &test, d2v *win, int take,
                                               double x0, double x1, double x2, double yi0,
                                               double yi1, double yi2, double num[9]) {
  BondCursor cur;
  cur.start(bonds, broken, DAMAGE, i);
  uint32_t fresh = 0;
  while (true) {
    const int w0 = wave_min(cur.next);
    if (w0 == INT_MAX)
      break;
    stage_window<4, W>(DefgradPlanes{xv, u}, n, w0, win);
    wave_lds_sync();
    const int end = w0 > INT_MAX - W ? INT_MAX : w0 + W; // (a finished lane's next is INT_MAX)
    const int stop = window_stop(cur.k, take);
    while (cur.next >= w0 && cur.next < end && cur.k < stop) {
      const int o = cur.next - w0;
      const d2v a = win[o], b = win[W + o], ua = win[2 * W + o], ub = win[3 * W + o];
      const double xi[3] = {a.x - x0, a.y - x1, b.x - x2};
      const double yij[3] = {(a.x + ua.x) - yi0, (a.y + ua.y) - yi1, (b.x + ub.x) - yi2};
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
          num[3 * r + c] += b.y * yij[r] * xi[c];
      if (DAMAGE) {
        const uint32_t bit = cur.bit();
        if (!cur.bond_broken() && stretch_breaks(test, xi, yij))
          fresh |= bit;
        if (((cur.k + 1) & 3) == 0 || cur.k + 1 >= cur.len) { // leaving the group
          if (fresh) {
            atomicOr(&broken[cur.group_slot() >> 5], fresh);
            anyBroken[i] = 1;
          }
          fresh = 0;
        }
      }
      cur.advance(bonds, broken, DAMAGE);
    }
    wave_lds_sync();
  }
}

template <bool DAMAGE, int W>
__global__ void __launch_bounds__(256)
    k_defgrad_w(Rounds rd, Planes x, PairPlanes xv, UPairs u, Sell bonds, Planes kinv,
                Planes fout, uint32_t *broken, uint8_t *anyBroken, double initiation,
                double failure, int useCompression, Batch batch, int take) {
  extern __shared__ d2v wlds[];
  d2v *win = wlds + wave_slot() * (4 * W);
  const int64_t blk = xcd_block(rd.blocksPerRound);
  if (blk < 0)
    return;
  const StretchTest test = make_stretch_test(initiation, failure, useCompression);
  for (int r = 0; r < rd.rounds; ++r) {
    const int64_t i = rd.particle(r, blk);
    if (i - (threadIdx.x & 63) >= rd.n) // the wavefront's first particle: a uniform exit
      break;
    const bool mine = i < rd.n && point_active(batch, i);
    double x0 = 0.0, x1 = 0.0, x2 = 0.0, yi0 = 0.0, yi1 = 0.0, yi2 = 0.0;
    if (mine) {
      x0 = x(0, i);
      x1 = x(1, i);
      x2 = x(2, i);
      yi0 = x0 + u(0, i);
      yi1 = x1 + u(1, i);
      yi2 = x2 + u(2, i);
    }
    double num[9];
#pragma unroll
    for (int c = 0; c < 9; ++c)
      num[c] = 0.0;
    defgrad_window<DAMAGE, W>(rd.n, mine ? i : -1, xv, u, bonds, broken, anyBroken, test, win,
                              take, x0, x1, x2, yi0, yi1, yi2, num);
    if (!mine)
      continue;
    T3 nm, ki;
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      nm.m[c] = num[c];
      ki.m[c] = kinv(c, i);
    }
    const T3 f = t3_mul(nm, ki);
#pragma unroll
    for (int c = 0; c < 9; ++c)
      fout(c, i) = f.m[c];
  }
}

// ---- Planned windows --------------------------------------------------------------------
// The window walk of the windowed kernels depends only on the bond lists, the window W and the
// bond limit, so it can be done once (k_window_plan) and replayed every step: per wavefront the
// first particle of each of its windows, and per window and lane one 8-byte record holding the
// LDS offsets (j − window start) of the bonds the lane takes from that window, in CSR order,
// lowest byte first, 0xFF after the last. The planned kernels (`_p`) then read no column
// indices, take no wave-wide minimum and keep no cursor: a lane's bonds in a window are the
// next ones of its row and the record says where their neighbours sit. With PREFETCH the next
// window is staged into a second LDS buffer while the current one is summed. Every bond is
// summed in the same order with the same arithmetic as before, so the results equal the
// direct gathers bit for bit.
struct WindowPlan {
  const int *first;     // per wavefront: index of its first window; first[wavefronts] = total
  const int *start;     // per window: the first particle it stages
  const uint64_t *rec;  // per window and lane: the bond offsets, see above
};
constexpr int kPlanTake = 8; // the most bonds per lane and window a record holds

// Wavefront w plans particles 64w..64w+63 (the wavefronts of every bond-pass launch). FILL
// false counts each wavefront's windows into `count`; FILL true writes them from first[w] on.
// take ≤ kPlanTake.
template <int W, bool FILL>
__global__ void __launch_bounds__(256)
    k_window_plan(int64_t n, Sell bonds, int take, int *count, const int *first, int *start,
                  uint64_t *rec) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int64_t wave = i >> 6;
  if (wave * 64 >= n) // the same for the whole wavefront
    return;
  BondCursor cur;
  cur.start(bonds, nullptr, false, i < n ? i : -1);
  int64_t t = FILL ? first[wave] : 0;
  while (true) {
    const int w0 = wave_min(cur.next);
    if (w0 == INT_MAX)
      break;
    const int end = w0 > INT_MAX - W ? INT_MAX : w0 + W;
    const int stop = window_stop(cur.k, take);
    uint64_t r = ~0ull;
    int q = 0;
    while (cur.next >= w0 && cur.next < end && cur.k < stop) {
      r &= ~(0xFFull << (8 * q));
      r |= (uint64_t)(cur.next - w0) << (8 * q);
      ++q;
      cur.advance(bonds, nullptr, false);
    }
    if (FILL) {
      rec[t * 64 + lane] = r;
      if (lane == 0)
        start[t] = w0;
    }
    ++t;
  }
  if (!FILL && lane == 0)
    count[wave] = (int)t;
}

// Bonds in a record (its offsets form a prefix; an offset is below W ≤ 255).
__device__ __forceinline__ int record_bonds(uint64_t r) {
  const uint64_t v = ~r;
  return v ? (71 - __builtin_clzll(v)) >> 3 : 0;
}

// Flag bits of bonds k, k+1, ... of a lane's row (bond k at bit 0, up to 12 bonds reaching
// into the next two groups): the flag words of the three groups from bond k & ~3 on that the
// row has (plan_words), then their 4-bit fields combined (plan_bits). g0 is the slot of the
// lane's bond 0; a group's 4 slots sit in one word, groups 8 words apart.
struct FlagWords {
  uint32_t w[3];
};
__device__ __forceinline__ FlagWords plan_words(const uint32_t *broken, int64_t g0, int k,
                                                int len) {
  const int64_t w = (g0 >> 5) + 8 * (int64_t)(k >> 2);
  FlagWords f;
#pragma unroll
  for (int q = 0; q < 3; ++q)
    f.w[q] = (k & ~3) + 4 * q < len ? broken[w + 8 * q] : 0u;
  return f;
}
__device__ __forceinline__ uint32_t plan_bits(const FlagWords &f, int64_t g0, int k) {
  const int sh = (int)(g0 & 31);
  uint32_t m = 0;
#pragma unroll
  for (int q = 0; q < 3; ++q)
    m |= ((f.w[q] >> sh) & 0xFu) << (4 * q);
  return m >> (k & 3);
}

// Sets new breaks of bonds k, k+1, ... (bond k at bit 0 of `fresh`) in the lane's flag words.
__device__ __forceinline__ void plan_flush(uint32_t *broken, int64_t g0, int k, uint32_t fresh) {
  const uint32_t m = fresh << (k & 3);
  const int64_t w = (g0 >> 5) + 8 * (int64_t)(k >> 2);
  const int sh = (int)(g0 & 31);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const uint32_t bits = (m >> (4 * q)) & 0xFu;
    if (bits)
      atomicOr(&broken[w + 8 * q], bits << sh);
  }
}

// The first double of a staged pair (an 8-byte LDS read where the second is padding).
__device__ __forceinline__ double lds_first(const d2v *p) {
  return *reinterpret_cast<const double *>(p);
}

// Uniform loads of the plan through the constant address space: scalar loads.
typedef __attribute__((address_space(4))) const int plan_int;
__device__ __forceinline__ int plan_load(const int *p, int64_t k) { return ((const plan_int *)p)[k]; }

// The window loop of a planned bond pass for the lane's particle (mine false: a lane with no
// particle, which still stages its share). body(buf, r, brk, k, g0) sums the lane's bonds of
// one window: `buf` holds the window, `r` the record, `brk` the flag bits from the lane's
// bond k on (with FLAGS, for a lane whose particle is `flagged`; 0 otherwise), g0 the slot of
// its bond 0. Without PREFETCH the windows go through the
// wavefront's part of the dynamic LDS, `win`; with it, through two static buffers, the next
// window being staged into one while the current one is summed from the other. (The buffers
// are separate LDS objects and the loop alternates them explicitly, so the compiler can tell
// the window being staged from the one being read and does not wait for the staging.)
template <int PAIRS, int W, bool PREFETCH, bool FLAGS, class Planeset, class Body>
__device__ __forceinline__ void planned_windows(const Planeset &planes, int64_t n,
                                                const WindowPlan &plan, const Sell &bonds,
                                                const uint32_t *broken, int64_t i, bool mine,
                                                bool flagged, d2v *win, Body &&body) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)((i - lane) >> 6));
  int t = plan_load(plan.first, wave);
  const int t1 = plan_load(plan.first, wave + 1);
  const int len = mine ? bonds.rowptr[i + 1] - bonds.rowptr[i] : 0;
  const int64_t g0 = mine ? sell_pos(bonds, i, 0) : 0;
  int k = 0;
  int wNext = t < t1 ? plan_load(plan.start, t) : 0; // first particle of window t
  if constexpr (!PREFETCH) {
    // The record and flag words of window t + 1 are loaded once window t is staged, so that
    // their memory latency passes while window t is summed (the records are read once per
    // pass and come from HBM); the wait at the end of a window is for its LDS reads only.
    uint64_t r = mine && t < t1 ? plan.rec[(int64_t)t * 64 + lane] : ~0ull;
    FlagWords fw{};
    if (FLAGS && flagged && t < t1)
      fw = plan_words(broken, g0, k, len);
    for (; t < t1; ++t) {
      const int w0 = wNext;
      if (t + 1 < t1)
        wNext = plan_load(plan.start, t + 1);
      stage_window<PAIRS, W>(planes, n, w0, win);
      wave_lds_sync(); // window t is staged (and its record and flag words are in)
      const uint64_t rc = r;
      const uint32_t brk = FLAGS ? plan_bits(fw, g0, k) : 0u;
      const int kNext = k + record_bonds(rc);
      r = ~0ull;
      if (t + 1 < t1 && mine) {
        if (FLAGS && flagged)
          fw = plan_words(broken, g0, kNext, len);
        r = plan.rec[(int64_t)(t + 1) * 64 + lane];
      }
      body(win, rc, brk, k, g0);
      k = kNext;
      wave_lds_wait(); // window t's LDS reads are done before it is restaged
    }
  } else {
    __shared__ d2v bufA[4 * PAIRS * W], bufB[4 * PAIRS * W];
    d2v *a = bufA + wave_slot() * (PAIRS * W);
    d2v *b = bufB + wave_slot() * (PAIRS * W);
    // The record and flag words of the window to be summed next are loaded with its data
    // and used only after the wait for it, so that no wait falls inside a window's sum.
    uint64_t r = ~0ull;
    FlagWords fw{};
    if (t < t1) {
      stage_window<PAIRS, W>(planes, n, wNext, a);
      r = mine ? plan.rec[(int64_t)t * 64 + lane] : ~0ull;
      if (FLAGS && flagged)
        fw = plan_words(broken, g0, k, len);
      if (t + 1 < t1)
        wNext = plan_load(plan.start, t + 1);
    }
    // Window t from `cur`; window t + 1 staged into `next` meanwhile.
    auto one = [&](const d2v *cur, d2v *next) {
      wave_lds_sync(); // window t is staged; the reads of window t − 1 (in `next`) are done
      const uint32_t brk = FLAGS ? plan_bits(fw, g0, k) : 0u;
      const uint64_t rc = r;
      const int kNext = k + record_bonds(rc);
      r = ~0ull;
      if (t + 1 < t1) {
        stage_window<PAIRS, W>(planes, n, wNext, next);
        r = mine ? plan.rec[(int64_t)(t + 1) * 64 + lane] : ~0ull;
        if (FLAGS && flagged)
          fw = plan_words(broken, g0, kNext, len);
        if (t + 2 < t1)
          wNext = plan_load(plan.start, t + 2);
      }
      body(cur, rc, brk, k, g0);
      k = kNext;
      ++t;
    };
    while (t < t1) {
      one(a, b);
      if (t >= t1)
        break;
      one(b, a);
    }
    wave_lds_sync(); // the last window's reads are done before the buffers are reused
  }
}

// Bond pass 1 with planned windows (4 component pairs: X_j, V_j, u_j).
// Without PREFETCH the pass fits 5 wavefronts per SIMD in registers (95 VGPRs) when asked to;
// with it the two window buffers already limit the workgroups per CU.
template <bool DAMAGE, int W, bool PREFETCH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PREFETCH ? 1 : 5)))
    k_defgrad_p(Rounds rd, Planes x, PairPlanes xv, UPairs u, Sell bonds, Planes kinv,
                Planes fout, uint32_t *broken, uint8_t *anyBroken, double initiation,
                double failure, int useCompression, Batch batch, WindowPlan plan) {
  extern __shared__ d2v wlds[];
  d2v *win = wlds + wave_slot() * (4 * W); // (unused with PREFETCH)
  const int64_t blk = xcd_block(rd.blocksPerRound);
  if (blk < 0)
    return;
  const StretchTest test = make_stretch_test(initiation, failure, useCompression);
  for (int r = 0; r < rd.rounds; ++r) {
    const int64_t i = rd.particle(r, blk);
    if (i - (threadIdx.x & 63) >= rd.n) // the wavefront's first particle: a uniform exit
      break;
    const bool mine = i < rd.n && point_active(batch, i);
    double x0 = 0.0, x1 = 0.0, x2 = 0.0, yi0 = 0.0, yi1 = 0.0, yi2 = 0.0;
    if (mine) {
      x0 = x(0, i);
      x1 = x(1, i);
      x2 = x(2, i);
      yi0 = x0 + u(0, i);
      yi1 = x1 + u(1, i);
      yi2 = x2 + u(2, i);
    }
    double num[9];
#pragma unroll
    for (int c = 0; c < 9; ++c)
      num[c] = 0.0;
    // The stretch test runs on every bond, broken or not: setting a bit that is set changes
    // nothing, so the flag words are read only in the windows where a bond breaks, to keep
    // the atomics to the new breaks.
    planned_windows<4, W, PREFETCH, false>(
        DefgradPlanes{xv, u}, rd.n, plan, bonds, broken, i, mine, false, win,
        [&](const d2v *buf, uint64_t rec, uint32_t, int k, int64_t g0) {
          uint32_t fresh = 0;
          for (int q = 0; q < kPlanTake; ++q) {
            const int o = (int)(rec & 0xFFu);
            if (o == 0xFF)
              break;
            const d2v a = buf[o], b = buf[W + o], ua = buf[2 * W + o];
            const double uz = lds_first(buf + 3 * W + o); // (u_z, 0): 8 bytes suffice
            const double xi[3] = {a.x - x0, a.y - x1, b.x - x2};
            const double yij[3] = {(a.x + ua.x) - yi0, (a.y + ua.y) - yi1, (b.x + uz) - yi2};
#pragma unroll
            for (int rr = 0; rr < 3; ++rr)
#pragma unroll
              for (int c = 0; c < 3; ++c)
                num[3 * rr + c] += b.y * yij[rr] * xi[c];
            if (DAMAGE && stretch_breaks(test, xi, yij))
              fresh |= 1u << q;
            rec >>= 8;
          }
          if (DAMAGE && fresh) { // (only a lane with a particle has bonds)
            const int len = bonds.rowptr[i + 1] - bonds.rowptr[i];
            fresh &= ~plan_bits(plan_words(broken, g0, k, len), g0, k);
            if (fresh) {
              plan_flush(broken, g0, k, fresh);
              anyBroken[i] = 1;
            }
          }
        });
    if (!mine)
      continue;
    T3 nm, ki;
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      nm.m[c] = num[c];
      ki.m[c] = kinv(c, i);
    }
    const T3 f = t3_mul(nm, ki);
#pragma unroll
    for (int c = 0; c < 9; ++c)
      fout(c, i) = f.m[c];
  }
}

// ---- Group windows ----------------------------------------------------------------------
// The windows above serve one wavefront each. With the lattice numbered line by line a
// wavefront's 64 particles cover about one line, and every wavefront stages the neighbour
// lines around it again. The 256 consecutive particles of a workgroup share most of their
// neighbours, so here the whole workgroup stages one window of W consecutive particles (from
// the smallest neighbour index any of its lanes still needs) and every lane takes all its
// next bonds whose neighbour lies in it. On the 200×50×50 lattice (δ = 3.015Δx, W = 512) a
// workgroup's ~8 windows stage about 11 neighbour particles per particle, against 28–40 for
// the wavefront windows, and its lanes take near-equal numbers of bonds per window (one
// neighbour plane each). The plan (k_group_plan) holds per window its first particle, the
// particles it stages, its first entry block and the entries each wavefront reads; per entry
// block and thread one 8-byte entry with the LDS offsets (j − window start) of four of the
// lane's bonds in CSR order, 16 bits each, lowest first, 0xFFFF after the last. Each lane
// sums its bonds in CSR order with the arithmetic of the direct kernels: results equal them
// bit for bit.
struct GroupPlan {
  const int *first;     // per workgroup: index of its first window; first[groups] = total
  const int *meta;      // per window, 8 ints: first particle, particles staged, first entry
                        // block, 0, then the entries wavefronts 0..3 read
  const uint64_t *off;  // per entry block and thread (block·256 + thread): see above
};
constexpr int kGroupMeta = 8;

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    v = max(v, __shfl_xor(v, off));
  return __builtin_amdgcn_readfirstlane(v);
}
// Minimum / maximum over the 256 threads of the workgroup (every thread calls; `red` holds
// 4 ints of LDS).
template <bool MIN> __device__ __forceinline__ int group_reduce(int v, int *red) {
  v = MIN ? wave_min(v) : wave_max(v);
  __syncthreads(); // earlier readers of `red` are done
  if ((threadIdx.x & 63) == 0)
    red[threadIdx.x >> 6] = v;
  __syncthreads();
  return MIN ? min(min(red[0], red[1]), min(red[2], red[3]))
             : max(max(red[0], red[1]), max(red[2], red[3]));
}

// Workgroup g plans particles 256g..256g+255 (the workgroups of every bond-pass launch: the
// grid has one workgroup per 256 particles). FILL false counts each workgroup's windows and
// entry blocks into count[2g], count[2g + 1]; FILL true writes them from firstWin[g] and
// firstBlock[g] on.
template <int W, bool FILL>
__global__ void __launch_bounds__(256)
    k_group_plan(int64_t n, Sell bonds, int *count, const int *firstWin, const int *firstBlock,
                 int *meta, uint64_t *off) {
  __shared__ int red[4];
  const int64_t g = blockIdx.x;
  const int64_t i = g * 256 + threadIdx.x;
  BondCursor cur;
  cur.start(bonds, nullptr, false, i < n ? i : -1);
  int64_t t = FILL ? firstWin[g] : 0;
  int64_t blk = FILL ? firstBlock[g] : 0;
  while (true) {
    const int w0 = group_reduce<true>(cur.next, red);
    if (w0 == INT_MAX)
      break;
    const int end = w0 > INT_MAX - W ? INT_MAX : w0 + W;
    int q = 0, reach = 0; // bonds taken, and one past the largest offset among them
    uint64_t e = ~0ull;
    while (cur.next >= w0 && cur.next < end) {
      const int o = cur.next - w0;
      reach = max(reach, o + 1);
      const int sh = 16 * (q & 3);
      e = (e & ~(0xFFFFull << sh)) | ((uint64_t)o << sh);
      ++q;
      if ((q & 3) == 0) {
        if (FILL)
          off[(blk + (q >> 2) - 1) * 256 + threadIdx.x] = e;
        e = ~0ull;
      }
      cur.advance(bonds, nullptr, false);
    }
    const int entries = (q + 3) >> 2;
    if (FILL && (q & 3))
      off[(blk + entries - 1) * 256 + threadIdx.x] = e;
    const int ext = group_reduce<false>(reach, red);
    const int waveEntries = wave_max(entries);
    const int blocks = group_reduce<false>(waveEntries, red);
    if (FILL) {
      for (int x = entries; x < blocks; ++x)
        off[(blk + x) * 256 + threadIdx.x] = ~0ull;
      int *m = meta + kGroupMeta * t;
      if (threadIdx.x == 0) {
        m[0] = w0;
        m[1] = ext;
        m[2] = (int)blk;
        m[3] = 0;
      }
      if ((threadIdx.x & 63) == 0)
        m[4 + (threadIdx.x >> 6)] = waveEntries;
    }
    ++t;
    blk += blocks;
  }
  if (!FILL && threadIdx.x == 0) {
    count[2 * g] = (int)t;
    count[2 * g + 1] = blk < INT_MAX ? (int)blk : INT_MAX;
  }
}

// The window loop of a group-window bond pass for workgroup `group`; every thread of the
// workgroup runs it (one without a particle, mine false, stages its share and sums nothing).
// Wavefront q stages window slots q·64 + lane, q·64 + 256 + lane, ... into `win` (PAIRS
// planes of W pairs); after a workgroup barrier every lane reads its entries and calls
// body(win, entry, k) with k its index of the entry's first bond; body returns the bonds it
// summed. A second barrier keeps the window until every wavefront has read it.
template <int PAIRS, int W, class Planeset, class Body>
__device__ __forceinline__ void group_windows(const Planeset &planes, const GroupPlan &plan,
                                              int64_t group, bool mine, d2v *win, Body &&body) {
  const int lane = threadIdx.x & 63;
  const int wave = wave_slot();
  const int t1 = plan_load(plan.first, group + 1);
  int k = 0;
  for (int t = plan_load(plan.first, group); t < t1; ++t) {
    const int64_t m = kGroupMeta * (int64_t)t;
    const int w0 = plan_load(plan.meta, m);
    const int ext = plan_load(plan.meta, m + 1);
    const int64_t blk = plan_load(plan.meta, m + 2);
    const int entries = plan_load(plan.meta, m + 4 + wave);
    // The first two entries are loaded with the window's data (the wait below covers them);
    // later ones two entries ahead of their use.
    const uint64_t *offs = plan.off + blk * 256 + threadIdx.x;
    uint64_t e = mine && entries > 0 ? offs[0] : ~0ull;
    uint64_t eNext = mine && entries > 1 ? offs[256] : ~0ull;
#pragma unroll
    for (int q0 = 0; q0 < W; q0 += 256) {
      const int s = q0 + 64 * wave;
      if (s < ext && s + lane < ext) {
#pragma unroll
        for (int c = 0; c < PAIRS; ++c)
          __builtin_amdgcn_global_load_lds(
              (const void *)(scalar_ptr(planes.ptr(c, (int64_t)w0 + s)) + lane),
              (lds_void *)(win + c * W + s), 16, 0, 0);
      }
    }
    wave_lds_sync(); // this wavefront's staging is complete
    __syncthreads(); // and every other's
    for (int x = 0; x < entries; ++x) {
      const uint64_t cur = e;
      e = eNext;
      if (mine && x + 2 < entries)
        eNext = offs[256 * (x + 2)];
      k += body(win, cur, k);
    }
    wave_lds_wait();
    __syncthreads(); // every wavefront has read the window
  }
}

// Wavefronts per SIMD a group-window pass can hold: each workgroup (one wavefront per SIMD)
// takes PAIRS × W × 16 bytes of the CU's 160 KB of LDS; at most 5 (96 VGPRs each).
template <int PAIRS, int W> constexpr int group_waves() {
  return 163840 / (PAIRS * W * 16) < 5 ? 163840 / (PAIRS * W * 16) : 5;
}

// Bond pass 1 with group windows (4 component pairs: X_j, V_j, u_j).
template <bool DAMAGE, int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(group_waves<4, W>())))
    k_defgrad_g(Rounds rd, Planes x, PairPlanes xv, UPairs u, Sell bonds, Planes kinv,
                Planes fout, uint32_t *broken, uint8_t *anyBroken, double initiation,
                double failure, int useCompression, Batch batch, GroupPlan plan) {
  extern __shared__ d2v wlds[];
  const int64_t blk = xcd_block(rd.blocksPerRound);
  if (blk < 0)
    return;
  const StretchTest test = make_stretch_test(initiation, failure, useCompression);
  for (int r = 0; r < rd.rounds; ++r) {
    const int64_t i = rd.particle(r, blk);
    const int64_t group = i >> 8;
    if (group * 256 >= rd.n) // the same for the whole workgroup
      break;
    const bool mine = i < rd.n && point_active(batch, i);
    double x0 = 0.0, x1 = 0.0, x2 = 0.0, yi0 = 0.0, yi1 = 0.0, yi2 = 0.0;
    int64_t g0 = 0;
    if (mine) {
      x0 = x(0, i);
      x1 = x(1, i);
      x2 = x(2, i);
      yi0 = x0 + u(0, i);
      yi1 = x1 + u(1, i);
      yi2 = x2 + u(2, i);
      g0 = sell_pos(bonds, i, 0);
    }
    double num[9];
#pragma unroll
    for (int c = 0; c < 9; ++c)
      num[c] = 0.0;
    // As in k_defgrad_p, the stretch test runs on every bond and the flag words are read
    // only where a bond breaks.
    group_windows<4, W>(
        DefgradPlanes{xv, u}, plan, group, mine, wlds,
        [&](const d2v *buf, uint64_t e, int k) -> int {
          uint32_t fresh = 0;
          int q = 0;
          for (; q < 4; ++q) {
            const int o = (int)(e & 0xFFFFu);
            if (o == 0xFFFF)
              break;
            const d2v a = buf[o], b = buf[W + o], ua = buf[2 * W + o];
            const double uz = lds_first(buf + 3 * W + o); // (u_z, 0): 8 bytes suffice
            const double xi[3] = {a.x - x0, a.y - x1, b.x - x2};
            const double yij[3] = {(a.x + ua.x) - yi0, (a.y + ua.y) - yi1, (b.x + uz) - yi2};
#pragma unroll
            for (int rr = 0; rr < 3; ++rr)
#pragma unroll
              for (int c = 0; c < 3; ++c)
                num[3 * rr + c] += b.y * yij[rr] * xi[c];
            if (DAMAGE && stretch_breaks(test, xi, yij))
              fresh |= 1u << q;
            e >>= 16;
          }
          if (DAMAGE && fresh) { // (only a lane with a particle has bonds)
            const int len = bonds.rowptr[i + 1] - bonds.rowptr[i];
            fresh &= ~plan_bits(plan_words(broken, g0, k, len), g0, k);
            if (fresh) {
              plan_flush(broken, g0, k, fresh);
              anyBroken[i] = 1;
            }
          }
          return q;
        });
    if (!mine)
      continue;
    T3 nm, ki;
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      nm.m[c] = num[c];
      ki.m[c] = kinv(c, i);
    }
    const T3 f = t3_mul(nm, ki);
#pragma unroll
    for (int c = 0; c < 9; ++c)
      fout(c, i) = f.m[c];
  }
}

// Material point update. VumatPdCorrespondenceMaterial.cpp:110-173 for every point at
// once (nblock = N; the 256-row padding of the CPU host leaves results unchanged because
// the routine is point-wise): Flanagan-Taylor kinematics, strainInc = D̄·dt and stressOld
// packed as Voigt (VumatMaterialAdapter.cpp:290-330), the Johnson-Cook return map,
// σ = R σ̄ Rᵀ, history commit; then A_i = (J σ F⁻ᵀ) K_i⁻¹ for the force pass
// (PdCorrespondenceSkeleton.cpp:209-210). sigOut/strainOut (spatial σ, Green-Lagrange E)
// are written only when non-null (diagnostics and snapshots).
struct MaterialArgs {
  int64_t n;
  int64_t nblocks;
  Planes fnew, fN, vN, rN, sig, statev, kinv;
  // Compacted update: copyF false leaves F_{n+1} in fnew for the host to swap the two planes
  // (no copy into fN); fCur is the plane the list kernels read F_{n+1} from.
  Planes fCur;
  bool copyF;
  PairPlanes amap; // A = P K⁻¹ as 5 component pairs (the last one (A_8, 0))
  double *temp, *eint, *einel;
  Planes sigOut, strainOut, incOut;
  int *iterOut, *fallbackOut;
  const double *seed;
  JcParams jc;
  ElasticParams elastic;
  double dt, rtol;
  int *err;
  DiagArgs diag;
  Batch batch;
};

// The stages after the return map, shared by the fused and the split material kernels:
// history commit, σ = R σ̄ Rᵀ, optional field outputs, the run diagnostics' material terms
// (DIAG: into dv, from the old stress sOld and the strain increment dE), A = P K⁻¹.
template <bool DIAG, bool JC_STATE = true>
__device__ __forceinline__ void material_finish(const MaterialArgs &g, int64_t i, const T3 &f1,
                                                const T3 &r1, const double sNew[6],
                                                const double st[6], double eint, double einel,
                                                const ReturnMapStats &stats, const double sOld[6],
                                                const double dE[6], double *dv) {
  const T3 sFull = from_voigt(sNew);
  const T3 sigma = t3_mul(r1, t3_mul(sFull, t3_transpose(r1)));
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    g.sig(c, i) = sNew[c];
    if (JC_STATE)
      g.statev(c, i) = st[c];
  }
  if (JC_STATE) { // the elastic VUMAT leaves state, temperature and energies as they were
    g.temp[i] = st[2];
    g.eint[i] = eint;
    g.einel[i] = einel;
  }
  if (g.iterOut)
    g.iterOut[i] = stats.iterations;
  if (g.fallbackOut)
    g.fallbackOut[i] = stats.fallbacks;
  if (g.sigOut.p) {
#pragma unroll
    for (int c = 0; c < 9; ++c)
      g.sigOut(c, i) = sigma.m[c];
  }
  if (g.strainOut.p || DIAG) {
    const T3 e = gl_strain(f1);
    if (g.strainOut.p) {
#pragma unroll
      for (int c = 0; c < 9; ++c)
        g.strainOut(c, i) = e.m[c];
    }
    if (DIAG)
      material_diag(g.diag, i, sOld, sigma, e, dE, st, dv);
  }
  T3 P;
  if (!first_piola(sigma, f1, P)) {
    raise_error(g.err, ERR_PIOLA_DET);
    return;
  }
  T3 ki;
#pragma unroll
  for (int c = 0; c < 9; ++c)
    ki.m[c] = g.kinv(c, i);
  store_amap(g.amap, i, t3_mul(P, ki));
}

// Material point update, fused: VumatPdCorrespondenceMaterial.cpp:110-173 for every point at
// once (nblock = N; the 256-row padding of the CPU host leaves results unchanged because
// the routine is point-wise): Flanagan-Taylor kinematics, strainInc = D̄·dt and stressOld
// packed as Voigt (VumatMaterialAdapter.cpp:290-330), the Johnson-Cook return map,
// σ = R σ̄ Rᵀ, history commit; then A_i = (J σ F⁻ᵀ) K_i⁻¹ for the force pass
// (PdCorrespondenceSkeleton.cpp:209-210). With FT_ONLY the kernel stops after the
// kinematics: it commits F, V, R and writes the strain increments to incOut, for a
// return-map stage that needs them first (the learned seed).
// One material point of k_material (see there); DIAG adds the diagnostics terms to dv.
template <int MODE, bool FT_ONLY, bool DIAG>
__device__ __forceinline__ void material_one(const MaterialArgs &g, int64_t i, double *dv) {
  const Ramp none{0.0, 0.0, 0.0};
  const PointStep ps = point_step(g.batch, i, g.dt, none, none);
  if (!ps.active)
    return;
  const double dt = ps.dt;
  T3 f1, f0, v0, r0;
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    f1.m[c] = g.fnew(c, i);
    f0.m[c] = g.fN(c, i);
    v0.m[c] = g.vN(c, i);
    r0.m[c] = g.rN(c, i);
  }
  Corot k;
  if (!corotational(f0, f1, v0, r0, dt, k)) {
    raise_error(g.err, ERR_SINGULAR_COROT);
    return;
  }
  double inc[6];
  to_voigt(k.Dbar, inc);
#pragma unroll
  for (int c = 0; c < 6; ++c)
    inc[c] *= dt;
  if (g.incOut.p) {
#pragma unroll
    for (int c = 0; c < 6; ++c)
      g.incOut(c, i) = inc[c];
  }
  double dE[6];
  if (DIAG) // E_n of the committed F_n equals the stored strainN of the reference bit for bit
    strain_increment(gl_strain(f1), gl_strain(f0), dE);
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    g.fN(c, i) = f1.m[c];
    g.vN(c, i) = k.V1.m[c];
    g.rN(c, i) = k.R1.m[c];
  }
  if (FT_ONLY) {
    if (DIAG) {
#pragma unroll
      for (int c = 0; c < 6; ++c)
        g.diag.dstrain(c, i) = dE[c];
    }
    return;
  }
  double sOld[6], sNew[6], st[6];
#pragma unroll
  for (int c = 0; c < 6; ++c)
    sOld[c] = g.sig(c, i);
  if constexpr (MODE == RM_ELASTIC) {
    // (no diagnostics terms: the host refuses run diagnostics with this material)
    elastic_point(g.elastic, inc, sOld, sNew);
    ReturnMapStats none{};
#pragma unroll
    for (int c = 0; c < 6; ++c)
      st[c] = 0.0;
    material_finish<false, false>(g, i, f1, k.R1, sNew, st, 0.0, 0.0, none, sOld, dE, dv);
  } else {
#pragma unroll
    for (int c = 0; c < 6; ++c)
      st[c] = g.statev(c, i);
    double eint = g.eint[i], einel = g.einel[i];
    ReturnMapStats stats;
    jc_point<MODE>(g.jc, dt, inc, sOld, st, eint, einel, sNew, 0.0, g.rtol, &stats);
    material_finish<DIAG>(g, i, f1, k.R1, sNew, st, eint, einel, stats, sOld, dE, dv);
  }
}

// Material point update, fused: VumatPdCorrespondenceMaterial.cpp:110-173 for every point at
// once (nblock = N; the 256-row padding of the CPU host leaves results unchanged because
// the routine is point-wise): Flanagan-Taylor kinematics, strainInc = D̄·dt and stressOld
// packed as Voigt (VumatMaterialAdapter.cpp:290-330), the Johnson-Cook return map,
// σ = R σ̄ Rᵀ, history commit; then A_i = (J σ F⁻ᵀ) K_i⁻¹ for the force pass
// (PdCorrespondenceSkeleton.cpp:209-210). With FT_ONLY the kernel stops after the
// kinematics: it commits F, V, R and writes the strain increments to incOut, for a
// return-map stage that needs them first (the learned seed). With DIAG (and not FT_ONLY)
// each 256-point chunk also writes its row of the run diagnostics' material terms.
template <int MODE, bool FT_ONLY = false, bool DIAG = false>
__global__ void __launch_bounds__(256) k_material(MaterialArgs g) {
  const int64_t blk = xcd_block(g.nblocks);
  if (blk < 0)
    return;
  const int64_t i = blk * blockDim.x + threadIdx.x;
  if (!DIAG) {
    if (i < g.n)
      material_one<MODE, FT_ONLY, DIAG>(g, i, nullptr);
    return;
  }
  if (FT_ONLY) { // (the diagnostics launch walks the chunks)
    bool valid;
    const int64_t ip = diag_particle(g.diag, blk, g.n, valid);
    if (valid)
      material_one<MODE, FT_ONLY, DIAG>(g, ip, nullptr);
    return;
  }
  double dv[D_NET_X] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  bool valid;
  const int64_t ip = diag_particle(g.diag, blk, g.n, valid);
  if (valid)
    material_one<MODE, false, true>(g, ip, dv);
  diag_block_store<D_NET_X, 0>(dv, g.diag.partials + blk * DIAG_STRIDE);
}

// The return-map half of the split material update: reads the strain increments written by
// k_material<…, true>, the committed F and R, and the per-point seed.
template <int MODE, bool DIAG>
__device__ __forceinline__ void material_return_one(const MaterialArgs &g, int64_t i,
                                                    double *dv) {
  const Ramp none{0.0, 0.0, 0.0};
  const PointStep ps = point_step(g.batch, i, g.dt, none, none);
  if (!ps.active)
    return;
  double inc[6], sOld[6], sNew[6], st[6], dE[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    inc[c] = g.incOut(c, i);
    sOld[c] = g.sig(c, i);
    st[c] = g.statev(c, i);
    dE[c] = DIAG ? g.diag.dstrain(c, i) : 0.0;
  }
  double eint = g.eint[i], einel = g.einel[i];
  ReturnMapStats stats;
  jc_point<MODE>(g.jc, ps.dt, inc, sOld, st, eint, einel, sNew, g.seed ? g.seed[i] : 0.0, g.rtol,
                 &stats);
  T3 f1, r1;
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    f1.m[c] = g.fN(c, i);
    r1.m[c] = g.rN(c, i);
  }
  material_finish<DIAG>(g, i, f1, r1, sNew, st, eint, einel, stats, sOld, dE, dv);
}

template <int MODE, bool DIAG = false>
__global__ void __launch_bounds__(256) k_material_return(MaterialArgs g) {
  const int64_t blk = xcd_block(g.nblocks);
  if (blk < 0)
    return;
  const int64_t i = blk * blockDim.x + threadIdx.x;
  if (!DIAG) {
    if (i < g.n)
      material_return_one<MODE, false>(g, i, nullptr);
    return;
  }
  double dv[D_NET_X] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  bool valid;
  const int64_t ip = diag_particle(g.diag, blk, g.n, valid);
  if (valid)
    material_return_one<MODE, true>(g, ip, dv);
  diag_block_store<D_NET_X, 0>(dv, g.diag.partials + blk * DIAG_STRIDE);
}

// ---- the compacted material update -----------------------------------------------------
// Most points of a step are elastic, and the plastic ones sit scattered through the block
// (about one lane in five of a wavefront in the necking bar): in k_material every wavefront
// with one plastic lane runs the whole return-map loop. k_material_screen runs the fused
// update for every point whose return map ends at its first iteration (reusing that
// iteration's yield) and appends the others to a list (one atomic per workgroup); the return
// map then runs on that dense list only (k_material_return_list). In the learned mode a
// listed point whose state 5 (its previous plastic increment) is positive is seeded with it
// and goes to the front of the list; the others (first yield, or yield again after an
// elastic increment) go to its back; the list kernel visits the back first, so its first
// wavefronts hold only them and run the network on them (wave_network_seed) before their
// Newton solves; no other wavefront loads the network. Each point's arithmetic is that of the fused kernel, so the result is
// bit for bit the same and does not depend on the list order.
struct CompactArgs {
  int *count;   // count[0]: slots from the front, count[1]: slots from the back (one 8-byte
                // word, zeroed before the screen)
  int *list;    // particle index per slot; back slot b is list[capacity − 1 − b]
  Planes inc;   // per slot: the strain increments (planes 0-5) and the first iteration's
                // 1 − θᵐ and yield as the screen evaluated them (planes 6, 7; JcFirst)
  int64_t capacity;
  const float *theta, *norm; // learned mode: the network (vpd_mlp)
};

__global__ void k_zero_count(int *count) {
  count[0] = 0;
  count[1] = 0;
}

// Slot of list position t, or −1 past the end: the back's slots first (in the learned mode
// the network's points, whose wavefronts then start first and overlap the network's latency
// with the rest of the list), then the front's.
__device__ __forceinline__ int64_t list_slot(const CompactArgs &cl, int64_t t, int front,
                                             int back) {
  return t < back ? cl.capacity - 1 - t : t < (int64_t)front + back ? t - back : -1;
}

// Positions within this workgroup's share of two lists: every thread calls; returns the
// thread's rank among the workgroup's threads taking A (and B), and the workgroup's totals.
// `base` holds 8 ints of LDS.
struct GroupRanks {
  int a, b, totalA, totalB;
};
__device__ __forceinline__ GroupRanks group_ranks(bool takeA, bool takeB, int *base) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6; // ≤ 4
  const uint64_t ma = __ballot(takeA), mb = __ballot(takeB);
  if (lane == 0) {
    base[wv] = __popcll(ma);
    base[4 + wv] = __popcll(mb);
  }
  __syncthreads();
  GroupRanks r{0, 0, 0, 0};
  for (int w = 0; w < nw; ++w) {
    if (w < wv) {
      r.a += base[w];
      r.b += base[4 + w];
    }
    r.totalA += base[w];
    r.totalB += base[4 + w];
  }
  const uint64_t below = (1ull << lane) - 1ull;
  r.a += (int)__popcll(ma & below);
  r.b += (int)__popcll(mb & below);
  __syncthreads(); // `base` may be reused
  return r;
}

// One append per workgroup to two lists behind one 64-bit counter (count[0] in its low half,
// count[1] in its high half): returns this thread's positions for `takeA` and `takeB`. Every
// thread of the workgroup calls; `base` holds 10 ints of LDS. Appends from many workgroups to
// one address serialise in the memory system, so a kernel appends once per workgroup at most.
struct Appended {
  int a, b;
};
__device__ __forceinline__ Appended group_append2(bool takeA, bool takeB, int *count,
                                                  int *base) {
  const GroupRanks r = group_ranks(takeA, takeB, base);
  if (threadIdx.x == 0) {
    unsigned long long old = 0;
    if (r.totalA + r.totalB > 0)
      old = atomicAdd((unsigned long long *)count,
                      (unsigned long long)r.totalA + ((unsigned long long)r.totalB << 32));
    base[8] = (int)(old & 0xFFFFFFFFull);
    base[9] = (int)(old >> 32);
  }
  __syncthreads();
  const Appended pos{base[8] + r.a, base[9] + r.b};
  __syncthreads(); // `base` may be reused
  return pos;
}

// Inclusive prefix sum over the 256 threads of a workgroup (`red` holds 4 ints of LDS).
__device__ __forceinline__ int group_inclusive_sum(int v, int *red) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(v, off);
    if (lane >= off)
      v += t;
  }
  if (lane == 63)
    red[wv] = v;
  __syncthreads();
  for (int w = 0; w < wv; ++w)
    v += red[w];
  __syncthreads(); // `red` may be reused
  return v;
}

// Return-map seeds for a block of n points in one launch: workgroup g takes the 2048 points
// from 2048·g (8 per thread, their loads in flight together), writes state 5 (component 4 of
// the state planes, stride stateStride) where it is positive, lists the others in its LDS and
// runs the network on them in 16-point tiles, tile t on wavefront t mod 4: the seed
// k_mlp_seed gives each of those points (one MFMA column per point, inputs normalised
// alike), with no device-wide list and no atomic. allSeeds false: only the network's seeds
// are written (a reader takes state 5 itself where it is positive).
__global__ void __launch_bounds__(256) k_seed_block(int64_t n, const double *__restrict__ inc,
                                                    int64_t incStride,
                                                    const double *__restrict__ state,
                                                    int64_t stateStride,
                                                    const float *__restrict__ theta,
                                                    const float *__restrict__ norm,
                                                    double *__restrict__ seed, bool allSeeds) {
  __shared__ int ids[2048];
  __shared__ int red[4];
  __shared__ int total;
  const int64_t g0 = (int64_t)blockIdx.x * 2048;
  const int64_t p0 = g0 + 8 * (int64_t)threadIdx.x;
  double prev[8];
#pragma unroll
  for (int j = 0; j < 8; ++j)
    prev[j] = p0 + j < n ? state[4 * stateStride + p0 + j] : 1.0;
  uint32_t mine = 0; // bit j: point p0 + j takes the network's seed
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    if (prev[j] > 0.0) {
      if (allSeeds && p0 + j < n)
        seed[p0 + j] = prev[j];
    } else {
      mine |= 1u << j;
    }
  }
  const int cnt = __popc(mine);
  const int incl = group_inclusive_sum(cnt, red);
  int pos = incl - cnt;
  for (int j = 0; j < 8; ++j)
    if ((mine >> j) & 1u)
      ids[pos++] = 8 * (int)threadIdx.x + j;
  if (threadIdx.x == 255)
    total = incl;
  __syncthreads();
  const int count = total, tiles = (count + 15) / 16;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
  if (wv >= tiles)
    return; // (uniform per wavefront; no barrier follows)
  MlpWeights w;
  mlp_load_weights(theta, lane, w);
  const double mean = norm[2 * MLP_IN], sdev = norm[2 * MLP_IN + 1];
  double mu[4], inv[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int f = q < 2 ? 4 * q + c : c;
    mu[c] = norm[f];
    inv[c] = 1.0 / (double)norm[MLP_IN + f];
  }
  for (int tile = wv; tile < tiles; tile += 4) {
    const int e = 16 * tile + r;
    const int64_t s = e < count ? g0 + ids[e] : 0;
    // This lane's 4 raw inputs (lanes 0-15: increments 0-3; lanes 16-31: increments 4-5,
    // ε̄ᵖ, T); missing points normalise to 0.
    double v[4];
#pragma unroll
    for (int c = 0; c < 4; ++c)
      v[c] = mu[c];
    if (q < 2 && e < count) {
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
    const float y = mlp_tile(w, x);
    if (q == 0 && e < count) {
      const double est = (double)y * sdev + mean;
      seed[s] = est > 0.0 ? est : 0.0;
    }
  }
}

template <int MODE>
__global__ void __launch_bounds__(256) k_material_screen(MaterialArgs g, CompactArgs cl) {
  __shared__ int base[10];
  const int64_t blk = xcd_block(g.nblocks);
  if (blk < 0)
    return;
  const int64_t i = blk * blockDim.x + threadIdx.x;
  // Every thread reaches the workgroup's list append below (one atomic per workgroup: one
  // per wavefront, serialised on the counter, costs more than the screen's own work).
  bool live = i < g.n;
  PointStep ps{};
  if (live) {
    const Ramp none{0.0, 0.0, 0.0};
    ps = point_step(g.batch, i, g.dt, none, none);
    live = ps.active;
  }
  const double dt = ps.dt;
  T3 f1, r1;
  double inc[6], sOld[6], st[6];
  bool plastic = false;
  JcFirst first{};
  if (live) {
    T3 f0, v0, r0;
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      f1.m[c] = g.fnew(c, i);
      f0.m[c] = g.fN(c, i);
      v0.m[c] = g.vN(c, i);
      r0.m[c] = g.rN(c, i);
    }
    Corot k;
    if (!corotational(f0, f1, v0, r0, dt, k)) {
      raise_error(g.err, ERR_SINGULAR_COROT);
      live = false;
    } else {
      to_voigt(k.Dbar, inc);
#pragma unroll
      for (int c = 0; c < 6; ++c)
        inc[c] *= dt;
      if (g.incOut.p) {
#pragma unroll
        for (int c = 0; c < 6; ++c)
          g.incOut(c, i) = inc[c];
      }
      if (g.copyF) {
#pragma unroll
        for (int c = 0; c < 9; ++c)
          g.fN(c, i) = f1.m[c];
      }
#pragma unroll
      for (int c = 0; c < 9; ++c) {
        g.vN(c, i) = k.V1.m[c];
        g.rN(c, i) = k.R1.m[c];
      }
      r1 = k.R1;
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        sOld[c] = g.sig(c, i);
        st[c] = g.statev(c, i);
      }
      plastic = !jc_elastic<MODE == RM_BISECTION || MODE == RM_BISECTION_TIGHT,
                            MODE == RM_BISECTION_TIGHT>(g.jc, inc, sOld, st, g.rtol, &first);
    }
  }
  // The elastic points finish before the workgroup meets at the append, so that no
  // wavefront waits there with its update still to do.
  if (live && !plastic) {
    double sNew[6];
    double eint = g.eint[i], einel = g.einel[i];
    ReturnMapStats stats;
    jc_point<MODE, true>(g.jc, dt, inc, sOld, st, eint, einel, sNew, 0.0, g.rtol, &stats,
                         &first);
    material_finish<false>(g, i, f1, r1, sNew, st, eint, einel, stats, sOld, nullptr, nullptr);
  }
  // Learned mode: the points the network seeds (no positive state 5) go to the back.
  const bool back = MODE == RM_LEARNED && plastic && !(st[4] > 0.0);
  const Appended pos = group_append2(plastic && !back, back, cl.count, base);
  if (plastic) {
    const int64_t slot = back ? cl.capacity - 1 - pos.b : pos.a;
    cl.list[slot] = (int)i;
#pragma unroll
    for (int c = 0; c < 6; ++c)
      cl.inc(c, slot) = inc[c];
    cl.inc(6, slot) = first.soft;
    cl.inc(7, slot) = first.yield;
  }
}

template <int MODE>
__global__ void __launch_bounds__(256) k_material_return_list(MaterialArgs g, CompactArgs cl) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int front = cl.count[0], back = cl.count[1];
  if (t - (threadIdx.x & 63) >= (int64_t)front + back)
    return; // a wavefront past the list (the others keep every lane: the network needs them)
  const int64_t slot = list_slot(cl, t, front, back);
  const bool valid = slot >= 0;
  int64_t i = 0;
  double inc[6], sOld[6], sNew[6], st[6];
#pragma unroll
  for (int c = 0; c < 6; ++c)
    inc[c] = sOld[c] = st[c] = 0.0;
  if (valid) {
    i = cl.list[slot];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      inc[c] = cl.inc(c, slot);
      sOld[c] = g.sig(c, i);
      st[c] = g.statev(c, i);
    }
  }
  double seed = 0.0;
  if constexpr (MODE == RM_LEARNED) {
    // State 5 where positive (the front); the network on the back's points, which fill the
    // list's first wavefronts.
    __shared__ float xs[4][64 * MLP_IN];
    __shared__ double ys[4][64];
    const bool net = valid && t < back;
    const int wv = threadIdx.x >> 6;
    const double est = wave_network_seed(net, inc, st[1], st[2], cl.theta, cl.norm, xs[wv], ys[wv]);
    seed = net ? est : st[4];
  }
  if (!valid)
    return;
  const JcFirst first{cl.inc(6, slot), cl.inc(7, slot)};
  double eint = g.eint[i], einel = g.einel[i];
  ReturnMapStats stats;
  const Ramp none{0.0, 0.0, 0.0};
  const double dt = point_step(g.batch, i, g.dt, none, none).dt;
  jc_point<MODE>(g.jc, dt, inc, sOld, st, eint, einel, sNew, seed, g.rtol, &stats, &first);
  T3 f1, r1;
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    f1.m[c] = g.fCur(c, i);
    r1.m[c] = g.rN(c, i);
  }
  material_finish<false>(g, i, f1, r1, sNew, st, eint, einel, stats, sOld, nullptr, nullptr);
}

// The bisections' list on lane groups, for short lists (small blocks), where one lane per
// point leaves the GPU idle and the list waits on the longest chain of iterations: G lanes
// (8, 16, 32 or 64) share each listed point and run its bisection log2(G) iterations per round
// (jc_point's lane groups: the single lane's bits); the group's first lane writes it. A round
// costs a few single-lane iterations, so a list longer than kGroupedLanes / G points (more
// than about two wavefronts per SIMD) runs one lane per point instead; the grid covers the
// block's every point with G lanes.
constexpr int64_t kGroupedLanes = 131072;
template <int MODE, int G>
__device__ __forceinline__ void list_point_bisect(const MaterialArgs &g, const CompactArgs &cl,
                                                  int64_t t, int front, int back, bool lead) {
  const int64_t slot = list_slot(cl, t, front, back);
  const int64_t i = cl.list[slot];
  double inc[6], sOld[6], sNew[6], st[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    inc[c] = cl.inc(c, slot);
    sOld[c] = g.sig(c, i);
    st[c] = g.statev(c, i);
  }
  const JcFirst first{cl.inc(6, slot), cl.inc(7, slot)};
  double eint = g.eint[i], einel = g.einel[i];
  ReturnMapStats stats;
  const Ramp none{0.0, 0.0, 0.0};
  const double dt = point_step(g.batch, i, g.dt, none, none).dt;
  jc_point<MODE, false, G>(g.jc, dt, inc, sOld, st, eint, einel, sNew, 0.0, g.rtol, &stats,
                           &first);
  if (!lead)
    return;
  T3 f1, r1;
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    f1.m[c] = g.fCur(c, i);
    r1.m[c] = g.rN(c, i);
  }
  material_finish<false>(g, i, f1, r1, sNew, st, eint, einel, stats, sOld, nullptr, nullptr);
}
template <int MODE, int G>
__global__ void __launch_bounds__(256) k_material_return_list_g(MaterialArgs g, CompactArgs cl) {
  static_assert(MODE == RM_BISECTION || MODE == RM_BISECTION_TIGHT, "bisections only");
  const int front = cl.count[0], back = cl.count[1];
  const int64_t count = (int64_t)front + back;
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (count * G <= kGroupedLanes) {
    const int64_t t = gid / G;
    if (t >= count)
      return; // the whole group
    list_point_bisect<MODE, G>(g, cl, t, front, back, (threadIdx.x & (G - 1)) == 0);
  } else {
    if (gid >= count)
      return;
    list_point_bisect<MODE, 1>(g, cl, gid, front, back, true);
  }
}

// Bond pass 2. PdCorrespondenceSkeleton.cpp:213-227: f_i = Σ_k (1 − d_ik)(A_i ξ − A_j(−ξ)) V_j
// in CSR order (a broken bond contributes exactly zero, so it is skipped), then the
// velocity-Verlet corrector (ExplicitTimeIntegrator.cpp:36-52) and the constraints at t;
// with FUSE_NEXT also the next step's predictor and constraints at t + dt
// (ExplicitTimeIntegrator.cpp:23-34), which only touch the lane's own particle.
struct ForceArgs {
  int64_t n;
  int64_t nblocks;
  Rounds rd;
  Planes x, v, a, forceOut;
  UPairs u;
  PairPlanes xv, amap; // neighbour data as component pairs: (X_x, X_y), (X_z, V); A
  const double *vol;
  Sell bonds;
  const uint32_t *broken;
  // Per particle: 1 when any of its bonds may be broken (set with every new break, cleared
  // with the flags; a particle marked with no broken bond only costs the flag reads).
  const uint8_t *anyBroken;
  const uint8_t *codes;
  double dt, density;
  Ramp rampNow, rampNext;
  DiagArgs diag;
  Batch batch;
  int take; // windowed gathers: bonds a lane takes from one window at most (see window_stop)
  // When set, the pass clears the two list lengths of the compacted material update (the
  // step's material stage has run; the next one appends from zero).
  int *zeroCount = nullptr;
  // Captured steps: the counter slot the next step reads its number from (this step's plus
  // one; the step's own slot is another, read by every block of this pass).
  int *stepNext = nullptr;
};

// The force pass's one-thread bookkeeping for the next step: clear the material lists'
// lengths and, in a captured step, set the next step's number.
__device__ __forceinline__ void clear_list_length(const ForceArgs &g) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (g.zeroCount) {
      g.zeroCount[0] = 0;
      g.zeroCount[1] = 0;
    }
    if (g.stepNext)
      *g.stepNext = g.batch.current() + 1;
  }
}

// A of particle j from its 5 component pairs.
__device__ __forceinline__ T3 load_amap(PairPlanes amap, int64_t j) {
  T3 A;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const d2v v = amap(c, j);
    A.m[2 * c] = v.x;
    A.m[2 * c + 1] = v.y;
  }
  A.m[8] = amap(4, j).x;
  return A;
}

// One bond's term of PdCorrespondenceSkeleton.cpp:219-225 in the reference's operation
// order: t_i = A_i ξ, t_j = A_j (−ξ), term = (t_i − t_j) V_j (the caller adds it).
__device__ __forceinline__ void force_bond_term(const T3 &ai, const T3 &aj, const double xi3[3],
                                                double vj, double term[3]) {
  V3 xi;
  xi.v[0] = xi3[0];
  xi.v[1] = xi3[1];
  xi.v[2] = xi3[2];
  const V3 ti = t3_mulv(ai, xi);
  V3 nxi;
  nxi.v[0] = -xi.v[0];
  nxi.v[1] = -xi.v[1];
  nxi.v[2] = -xi.v[2];
  const V3 tj = t3_mulv(aj, nxi);
#pragma unroll
  for (int d = 0; d < 3; ++d)
    term[d] = (ti.v[d] - tj.v[d]) * vj;
}

// Bond pass 2, direct gathers: the bonds of particle i in groups of 4 (one 16-byte column
// load, one flag word), each bond's X_j, V_j and A_j as 7 component-pair loads; a broken
// bond contributes exactly zero, so it is skipped. The sum runs in CSR order.
template <bool DAMAGE>
__device__ __forceinline__ void force_sum_direct(const ForceArgs &g, int64_t i, double x0,
                                                 double x1, double x2, const T3 &ai,
                                                 double acc[3]) {
  const int len = g.bonds.rowptr[i + 1] - g.bonds.rowptr[i];
  const int64_t g0 = sell_pos(g.bonds, i, 0);
  for (int k0 = 0; k0 < len; k0 += 4) {
    const int64_t pg = g0 + 64 * (int64_t)k0; // slot of bond k0 (k0 is a multiple of 4)
    const i4v jj = *(const i4v *)(g.bonds.col + pg);
    const uint32_t flags = DAMAGE ? g.broken[pg >> 5] : 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int j = jj[q];
      const d2v a = g.xv(0, j), b = g.xv(1, j);
      const T3 aj = load_amap(g.amap, j);
      const bool live = k0 + q < len && !(DAMAGE && (flags & (1u << ((pg + q) & 31))));
      if (live) {
        const double xi[3] = {a.x - x0, a.y - x1, b.x - x2};
        double term[3];
        force_bond_term(ai, aj, xi, b.y, term);
#pragma unroll
        for (int d = 0; d < 3; ++d)
          acc[d] += term[d];
      }
    }
  }
}

// Bond pass 2 with windowed gathers (see BondCursor): X_j, V_j and A_j of W consecutive
// particles per window as 7 component pairs in the wavefront's LDS, at most g.take bonds per
// lane and window. i < 0: a lane with no particle, which still takes part in the wavefront's
// window steps.
template <bool DAMAGE, int W>
__device__ __forceinline__ void force_sum_window(const ForceArgs &g, int64_t i, d2v *win,
                                                 double x0, double x1, double x2, const T3 &ai,
                                                 double acc[3]) {
  BondCursor cur;
  cur.start(g.bonds, g.broken, DAMAGE, i);
  while (true) {
    const int w0 = wave_min(cur.next);
    if (w0 == INT_MAX)
      break;
    stage_window<7, W>(ForcePlanes{g.xv, g.amap}, g.n, w0, win);
    wave_lds_sync();
    const int end = w0 > INT_MAX - W ? INT_MAX : w0 + W; // (a finished lane's next is INT_MAX)
    const int stop = window_stop(cur.k, g.take);
    while (cur.next >= w0 && cur.next < end && cur.k < stop) {
      const int o = cur.next - w0;
      if (!(DAMAGE && cur.bond_broken())) {
        const d2v a = win[o], b = win[W + o];
        T3 aj;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const d2v v = win[(2 + c) * W + o];
          aj.m[2 * c] = v.x;
          aj.m[2 * c + 1] = v.y;
        }
        aj.m[8] = win[6 * W + o].x;
        const double xi[3] = {a.x - x0, a.y - x1, b.x - x2};
        double term[3];
        force_bond_term(ai, aj, xi, b.y, term);
#pragma unroll
        for (int d = 0; d < 3; ++d)
          acc[d] += term[d];
      }
      cur.advance(g.bonds, g.broken, DAMAGE);
    }
    wave_lds_sync();
  }
}

// The rest of the force pass for particle i after its bond sum: the force density output,
// velocity-Verlet correction and constraints, the diagnostics' force terms and the fused
// next prediction.
template <bool FUSE_NEXT, bool INTEGRATE, bool DIAG>
__device__ __forceinline__ void force_finish(const ForceArgs &g, int64_t i, const double acc[3],
                                             double *dv) {
  if (g.forceOut.p) {
#pragma unroll
    for (int d = 0; d < 3; ++d)
      g.forceOut(d, i) = acc[d];
  }
  if (!INTEGRATE)
    return;
  const PointStep ps = point_step(g.batch, i, g.dt, g.rampNow, g.rampNext);
  double uu[3], vv[3], aa[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double next = acc[d] / g.density;
    vv[d] = g.v(d, i) + 0.5 * ps.dt * (g.a(d, i) + next);
    aa[d] = next;
    uu[d] = g.u(d, i);
  }
  const uint8_t code = g.codes[i];
  constrain(code, ps.now, uu, vv, aa);
  if (DIAG) { // the force-side terms of BoundaryValueCaseRunner.cpp:244-285, in its expressions
    const double vi = g.vol[i];
    const int ax = g.diag.axis;
    dv[D_NET_X - D_NET_X] = acc[0] * vi;
    dv[D_NET_Y - D_NET_X] = acc[1] * vi;
    dv[D_NET_Z - D_NET_X] = acc[2] * vi;
    dv[D_REACTION - D_NET_X] =
        ((code >> (2 * ax)) & 3) == 2 /* load band */ ? (g.density * ps.now.a - acc[ax]) * vi : 0.0;
    dv[D_KINETIC - D_NET_X] =
        0.5 * g.density * vi * (vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2]);
    const uint8_t lat = g.diag.latSet ? g.diag.latSet[i] : 0;
    dv[D_LATERAL_NEG - D_NET_X] = (lat & 1) ? uu[g.diag.lateral] : 0.0;
    dv[D_LATERAL_POS - D_NET_X] = (lat & 2) ? uu[g.diag.lateral] : 0.0;
    dv[D_NONFINITE_FORCE - D_NET_X] =
        (__builtin_isfinite(acc[0]) && __builtin_isfinite(acc[1]) && __builtin_isfinite(acc[2]))
            ? 0.0
            : 1.0;
  }
  if (FUSE_NEXT && !ps.last) {
#pragma unroll
    for (int d = 0; d < 3; ++d)
      uu[d] = uu[d] + (ps.dt * vv[d] + 0.5 * ps.dt * ps.dt * aa[d]);
    constrain(code, ps.next, uu, vv, aa);
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    g.u(d, i) = uu[d];
    g.v(d, i) = vv[d];
    g.a(d, i) = aa[d];
  }
}

template <bool DAMAGE, bool FUSE_NEXT, bool INTEGRATE, bool DIAG>
__device__ __forceinline__ void force_one(const ForceArgs &g, int64_t i, double *dv) {
  if (!point_active(g.batch, i))
    return;
  const double x0 = g.x(0, i), x1 = g.x(1, i), x2 = g.x(2, i);
  const T3 ai = load_amap(g.amap, i);
  double acc[3] = {0.0, 0.0, 0.0};
  force_sum_direct<DAMAGE>(g, i, x0, x1, x2, ai, acc);
  force_finish<FUSE_NEXT, INTEGRATE, DIAG>(g, i, acc, dv);
}

// With DIAG (and INTEGRATE) each 256-particle chunk also writes its row of the run
// diagnostics' force terms; the chunk loop then runs block-uniformly so that every thread
// reaches the chunk's reduction.
template <bool DAMAGE, bool FUSE_NEXT, bool INTEGRATE = true, bool DIAG = false>
__global__ void __launch_bounds__(256) k_force(ForceArgs g) {
  clear_list_length(g);
  static_assert(!DIAG || INTEGRATE, "the diagnostics read the corrected velocity");
  const int64_t blk = xcd_block(g.rd.blocksPerRound);
  if (blk < 0)
    return;
  for (int r = 0; r < g.rd.rounds; ++r) {
    const int64_t i = g.rd.particle(r, blk);
    if (!DIAG) {
      if (i >= g.rd.n)
        break;
      force_one<DAMAGE, FUSE_NEXT, INTEGRATE, false>(g, i, nullptr);
      continue;
    }
    const int64_t chunk = (int64_t)r * g.rd.blocksPerRound + blk;
    if (g.diag.chunkFirst ? chunk >= g.diag.chunks : chunk * 256 >= g.rd.n) // block-uniform
      break;
    constexpr int NS = D_COUNT - D_NET_X;
    double dv[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s)
      dv[s] = 0.0;
    bool valid;
    const int64_t ip = diag_particle(g.diag, chunk, g.rd.n, valid);
    if (valid)
      force_one<DAMAGE, FUSE_NEXT, INTEGRATE, true>(g, ip, dv);
    diag_block_store<NS, D_NET_X>(dv, g.diag.partials + chunk * DIAG_STRIDE + D_NET_X);
  }
}

// The force pass with windowed gathers (no diagnostics form). Every wavefront keeps all its
// lanes in the window steps; the rounds loop exits on the wavefront's first particle.
template <bool DAMAGE, bool FUSE_NEXT, bool INTEGRATE, int W>
__global__ void __launch_bounds__(256) k_force_w(ForceArgs g) {
  clear_list_length(g);
  extern __shared__ d2v wlds[];
  d2v *win = wlds + wave_slot() * (7 * W);
  const int64_t blk = xcd_block(g.rd.blocksPerRound);
  if (blk < 0)
    return;
  for (int r = 0; r < g.rd.rounds; ++r) {
    const int64_t i = g.rd.particle(r, blk);
    if (i - (threadIdx.x & 63) >= g.rd.n)
      break;
    const bool mine = i < g.rd.n && point_active(g.batch, i);
    double x0 = 0.0, x1 = 0.0, x2 = 0.0;
    T3 ai;
#pragma unroll
    for (int c = 0; c < 9; ++c)
      ai.m[c] = 0.0;
    if (mine) {
      x0 = g.x(0, i);
      x1 = g.x(1, i);
      x2 = g.x(2, i);
      ai = load_amap(g.amap, i);
    }
    double acc[3] = {0.0, 0.0, 0.0};
    force_sum_window<DAMAGE, W>(g, mine ? i : -1, win, x0, x1, x2, ai, acc);
    if (mine)
      force_finish<FUSE_NEXT, INTEGRATE, false>(g, i, acc, nullptr);
  }
}

// The force pass with planned windows (7 component pairs: X_j, V_j, A_j; see WindowPlan).
template <bool DAMAGE, bool FUSE_NEXT, bool INTEGRATE, int W, bool PREFETCH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PREFETCH ? 1 : 5)))
    k_force_p(ForceArgs g, WindowPlan plan) {
  clear_list_length(g);
  extern __shared__ d2v wlds[];
  d2v *win = wlds + wave_slot() * (7 * W); // (unused with PREFETCH)
  const int64_t blk = xcd_block(g.rd.blocksPerRound);
  if (blk < 0)
    return;
  for (int r = 0; r < g.rd.rounds; ++r) {
    const int64_t i = g.rd.particle(r, blk);
    if (i - (threadIdx.x & 63) >= g.rd.n)
      break;
    const bool mine = i < g.rd.n && point_active(g.batch, i);
    double x0 = 0.0, x1 = 0.0, x2 = 0.0;
    T3 ai;
#pragma unroll
    for (int c = 0; c < 9; ++c)
      ai.m[c] = 0.0;
    if (mine) {
      x0 = g.x(0, i);
      x1 = g.x(1, i);
      x2 = g.x(2, i);
      ai = load_amap(g.amap, i);
    }
    double acc[3] = {0.0, 0.0, 0.0};
    // Only a particle marked in anyBroken can have broken bonds: the others skip the flags.
    planned_windows<7, W, PREFETCH, DAMAGE>(
        ForcePlanes{g.xv, g.amap}, g.n, plan, g.bonds, g.broken, i, mine,
        DAMAGE && mine && g.anyBroken[i] != 0, win,
        [&](const d2v *buf, uint64_t rec, uint32_t brk, int, int64_t) {
          for (int q = 0; q < kPlanTake; ++q) {
            const int o = (int)(rec & 0xFFu);
            if (o == 0xFF)
              break;
            if (!(DAMAGE && (brk & 1u))) {
              const d2v a = buf[o], b = buf[W + o];
              T3 aj;
#pragma unroll
              for (int c = 0; c < 4; ++c) {
                const d2v v = buf[(2 + c) * W + o];
                aj.m[2 * c] = v.x;
                aj.m[2 * c + 1] = v.y;
              }
              aj.m[8] = lds_first(buf + 6 * W + o); // (A_8, 0): 8 bytes suffice
              const double xi[3] = {a.x - x0, a.y - x1, b.x - x2};
              double term[3];
              force_bond_term(ai, aj, xi, b.y, term);
#pragma unroll
              for (int d = 0; d < 3; ++d)
                acc[d] += term[d];
            }
            rec >>= 8;
            brk >>= 1;
          }
        });
    if (mine)
      force_finish<FUSE_NEXT, INTEGRATE, false>(g, i, acc, nullptr);
  }
}

// The force pass with group windows (7 component pairs: X_j, V_j, A_j; see GroupPlan).
template <bool DAMAGE, bool FUSE_NEXT, bool INTEGRATE, int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(group_waves<7, W>())))
    k_force_g(ForceArgs g, GroupPlan plan) {
  clear_list_length(g);
  extern __shared__ d2v wlds[];
  const int64_t blk = xcd_block(g.rd.blocksPerRound);
  if (blk < 0)
    return;
  for (int r = 0; r < g.rd.rounds; ++r) {
    const int64_t i = g.rd.particle(r, blk);
    const int64_t group = i >> 8;
    if (group * 256 >= g.rd.n) // the same for the whole workgroup
      break;
    const bool mine = i < g.rd.n && point_active(g.batch, i);
    double x0 = 0.0, x1 = 0.0, x2 = 0.0;
    T3 ai;
#pragma unroll
    for (int c = 0; c < 9; ++c)
      ai.m[c] = 0.0;
    int64_t g0 = 0;
    int len = 0;
    bool flagged = false;
    if (mine) {
      x0 = g.x(0, i);
      x1 = g.x(1, i);
      x2 = g.x(2, i);
      ai = load_amap(g.amap, i);
      g0 = sell_pos(g.bonds, i, 0);
      len = g.bonds.rowptr[i + 1] - g.bonds.rowptr[i];
      // Only a particle marked in anyBroken can have broken bonds: the others skip the flags.
      flagged = DAMAGE && g.anyBroken[i] != 0;
    }
    double acc[3] = {0.0, 0.0, 0.0};
    group_windows<7, W>(
        ForcePlanes{g.xv, g.amap}, plan, group, mine, wlds,
        [&](const d2v *buf, uint64_t e, int k) -> int {
          uint32_t brk = 0;
          if (DAMAGE && flagged)
            brk = plan_bits(plan_words(g.broken, g0, k, len), g0, k);
          int q = 0;
          for (; q < 4; ++q) {
            const int o = (int)(e & 0xFFFFu);
            if (o == 0xFFFF)
              break;
            if (!(DAMAGE && (brk & 1u))) {
              const d2v a = buf[o], b = buf[W + o];
              T3 aj;
#pragma unroll
              for (int c = 0; c < 4; ++c) {
                const d2v v = buf[(2 + c) * W + o];
                aj.m[2 * c] = v.x;
                aj.m[2 * c + 1] = v.y;
              }
              aj.m[8] = lds_first(buf + 6 * W + o); // (A_8, 0): 8 bytes suffice
              const double xi[3] = {a.x - x0, a.y - x1, b.x - x2};
              double term[3];
              force_bond_term(ai, aj, xi, b.y, term);
#pragma unroll
              for (int d = 0; d < 3; ++d)
                acc[d] += term[d];
            }
            e >>= 16;
            brk >>= 1;
          }
          return q;
        });
    if (mine)
      force_finish<FUSE_NEXT, INTEGRATE, false>(g, i, acc, nullptr);
  }
}

// Shared rows (gather form 4), for small blocks, where one lane per particle leaves the GPU
// nearly idle and a step waits on each lane's chain of ~120 dependent bond gathers: a group
// of R lanes (16, 32 or 64: a quarter, half or whole wavefront) owns one particle. In each
// chunk of kRowChunks·R bonds every lane computes the terms of kRowChunks bonds (the
// expressions of the direct kernels; all their loads issued before any term, slots past the
// row's end loading its last bond and left unread) into the group's LDS; then lane c adds
// component c of the chunk's terms in CSR order (9 components for F, 3 for the force). Each
// component's sum is the direct kernel's sum term for term, so the result is the same bits.
// A group exits alone (a particle outside the block or a finished specimen); groups share
// only their wavefront, so no workgroup barrier is used.
constexpr int kRowChunks = 2;
// The group's LDS terms are written (read) by all its lanes before any lane goes on; the
// compiler keeps its memory accesses on their side of this point.
__device__ __forceinline__ void row_lds_sync() {
  asm volatile("" ::: "memory");
  wave_lds_wait();
  asm volatile("" ::: "memory");
}

template <bool DAMAGE, int R>
__global__ void __launch_bounds__(256)
    k_defgrad_m(int64_t n, Planes x, PairPlanes xv, UPairs u, Sell bonds, Planes kinv,
                Planes fout, uint32_t *broken, uint8_t *anyBroken, double initiation,
                double failure, int useCompression, Batch batch) {
  constexpr int CH = kRowChunks;
  __shared__ double terms[256 / R][CH * R][9];
  const int gp = threadIdx.x / R, l = threadIdx.x % R;
  const int64_t i = (int64_t)blockIdx.x * (256 / R) + gp;
  if (i >= n || !point_active(batch, i))
    return;
  const StretchTest test = make_stretch_test(initiation, failure, useCompression);
  const double x0 = x(0, i), x1 = x(1, i), x2 = x(2, i);
  const double yi0 = x0 + u(0, i), yi1 = x1 + u(1, i), yi2 = x2 + u(2, i);
  const int len = bonds.rowptr[i + 1] - bonds.rowptr[i];
  const int64_t sliceBase = bonds.base[i >> 6];
  double num = 0.0; // component l of Σ V_j (y ⊗ ξ), l < 9
  bool fresh = false;
  for (int k0 = 0; k0 < len; k0 += CH * R) {
    int64_t pos[CH];
    int j[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      pos[c] = sell_slot(sliceBase, i, min(k0 + c * R + l, len - 1));
      j[c] = bonds.col[pos[c]];
    }
    d2v a[CH], b[CH], ua[CH], ub[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      a[c] = xv(0, j[c]);
      b[c] = xv(1, j[c]);
      ua[c] = u.pair(0, j[c]);
      ub[c] = u.pair(1, j[c]);
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const double xi[3] = {a[c].x - x0, a[c].y - x1, b[c].x - x2};
      const double yij[3] = {(a[c].x + ua[c].x) - yi0, (a[c].y + ua[c].y) - yi1,
                             (b[c].x + ub[c].x) - yi2};
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int q = 0; q < 3; ++q)
          terms[gp][c * R + l][3 * r + q] = b[c].y * yij[r] * xi[q];
      if (DAMAGE && k0 + c * R + l < len) {
        const uint32_t bit = 1u << (pos[c] & 31);
        if (!(broken[pos[c] >> 5] & bit) && stretch_breaks(test, xi, yij)) {
          atomicOr(&broken[pos[c] >> 5], bit);
          fresh = true;
        }
      }
    }
    row_lds_sync(); // the group's terms are in LDS
    if (l < 9) {
      const int m1 = min(CH * R, len - k0);
      for (int m = 0; m < m1; ++m)
        num += terms[gp][m][l];
    }
    row_lds_sync(); // and read before the next chunk overwrites them
  }
  if (DAMAGE && fresh)
    anyBroken[i] = 1;
  if (l < 9)
    terms[gp][0][l] = num;
  row_lds_sync();
  if (l == 0) {
    T3 nm, ki;
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      nm.m[c] = terms[gp][0][c];
      ki.m[c] = kinv(c, i);
    }
    const T3 f = t3_mul(nm, ki);
#pragma unroll
    for (int c = 0; c < 9; ++c)
      fout(c, i) = f.m[c];
  }
}

template <bool DAMAGE, bool FUSE_NEXT, bool INTEGRATE, int R>
__global__ void __launch_bounds__(256) k_force_m(ForceArgs g) {
  clear_list_length(g);
  constexpr int CH = kRowChunks;
  __shared__ double terms[256 / R][CH * R][3];
  const int gp = threadIdx.x / R, l = threadIdx.x % R;
  const int64_t i = (int64_t)blockIdx.x * (256 / R) + gp;
  if (i >= g.n || !point_active(g.batch, i))
    return;
  const double x0 = g.x(0, i), x1 = g.x(1, i), x2 = g.x(2, i);
  const T3 ai = load_amap(g.amap, i);
  const int len = g.bonds.rowptr[i + 1] - g.bonds.rowptr[i];
  const int64_t sliceBase = g.bonds.base[i >> 6];
  const int shift = R * ((threadIdx.x & 63) / R); // the group's bits in a ballot
  const uint64_t mask = R == 64 ? ~0ull : (1ull << (R & 63)) - 1;
  double acc = 0.0; // component l of the force density, l < 3
  for (int k0 = 0; k0 < len; k0 += CH * R) {
    int64_t pos[CH];
    int j[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      pos[c] = sell_slot(sliceBase, i, min(k0 + c * R + l, len - 1));
      j[c] = g.bonds.col[pos[c]];
    }
    d2v a[CH], b[CH];
    T3 aj[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      a[c] = g.xv(0, j[c]);
      b[c] = g.xv(1, j[c]);
      aj[c] = load_amap(g.amap, j[c]);
    }
    uint64_t lives[CH]; // a broken bond contributes exactly zero: it is skipped
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const bool live = k0 + c * R + l < len &&
                        !(DAMAGE && (g.broken[pos[c] >> 5] & (1u << (pos[c] & 31))));
      lives[c] = (__ballot(live) >> shift) & mask;
      const double xi[3] = {a[c].x - x0, a[c].y - x1, b[c].x - x2};
      double term[3];
      force_bond_term(ai, aj[c], xi, b[c].y, term);
#pragma unroll
      for (int d = 0; d < 3; ++d)
        terms[gp][c * R + l][d] = term[d];
    }
    row_lds_sync();
    if (l < 3) {
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int mc = min(R, len - k0 - c * R);
        // A select, not a branch, so that the LDS reads need not wait on the sums; a broken
        // bond leaves the sum as it was (not + 0.0, which would turn −0.0 into +0.0).
        for (int m = 0; m < mc; ++m) {
          const double t = terms[gp][c * R + m][l];
          acc = ((lives[c] >> m) & 1ull) ? acc + t : acc;
        }
      }
    }
    row_lds_sync();
  }
  if (l < 3)
    terms[gp][0][l] = acc;
  row_lds_sync();
  if (l == 0) {
    const double f[3] = {terms[gp][0][0], terms[gp][0][1], terms[gp][0][2]};
    force_finish<FUSE_NEXT, INTEGRATE, false>(g, i, f, nullptr);
  }
}

// cauchyToFirstPiola · K⁻¹ for a caller-supplied spatial stress (PdCorrespondenceSkeleton.cpp:
// 83-91, 209-210): A_i = (J σ F⁻ᵀ) K_i⁻¹.
__global__ void __launch_bounds__(256) k_piola(int64_t n, Planes sigma, Planes f, Planes kinv,
                                               PairPlanes amap, int *err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  T3 s, ff, ki, P;
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    s.m[c] = sigma(c, i);
    ff.m[c] = f(c, i);
    ki.m[c] = kinv(c, i);
  }
  if (!first_piola(s, ff, P)) {
    raise_error(err, ERR_PIOLA_DET);
    return;
  }
  store_amap(amap, i, t3_mul(P, ki));
}

// The full Flanagan-Taylor update of CorotationalKinematics.cpp:61-135 with every output,
// including the velocity gradient and the relative spin increment (:125-134), for callers
// of the per-point API.
__global__ void __launch_bounds__(256) k_corotational_full(int64_t n, Planes fN, Planes fNP1,
                                                           Planes vN, Planes rN, double dt,
                                                           Planes L, Planes dbar, Planes spin,
                                                           Planes r1, Planes v1, int *err) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  T3 f0, f1, v0, r0;
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    f0.m[c] = fN(c, i);
    f1.m[c] = fNP1(c, i);
    v0.m[c] = vN(c, i);
    r0.m[c] = rN(c, i);
  }
  Corot k;
  T3 lv, w, om;
  if (!corotational_full(f0, f1, v0, r0, dt, k, lv, w, om)) {
    raise_error(err, ERR_SINGULAR_COROT);
    return;
  }
  T3 rs;
#pragma unroll
  for (int c = 0; c < 9; ++c)
    rs.m[c] = w.m[c] - om.m[c];
  T3 sp = t3_mul(t3_transpose(k.R1), t3_mul(rs, k.R1));
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    L(c, i) = lv.m[c];
    dbar(c, i) = k.Dbar.m[c];
    spin(c, i) = sp.m[c] * dt;
    r1(c, i) = k.R1.m[c];
    v1(c, i) = k.V1.m[c];
  }
}

// σ = R σ̄ Rᵀ (CorotationalKinematics.cpp:137-139) and E = ½(FᵀF − I) (:141-149).
__global__ void __launch_bounds__(256) k_rotate_gl(int64_t n, Planes s, Planes r, Planes sOut,
                                                   Planes f, Planes eOut) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  if (sOut.p) {
    T3 a, b;
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      a.m[c] = s(c, i);
      b.m[c] = r(c, i);
    }
    const T3 o = t3_mul(b, t3_mul(a, t3_transpose(b)));
#pragma unroll
    for (int c = 0; c < 9; ++c)
      sOut(c, i) = o.m[c];
  }
  if (eOut.p) {
    T3 ff;
#pragma unroll
    for (int c = 0; c < 9; ++c)
      ff.m[c] = f(c, i);
    T3 e = t3_mul(t3_transpose(ff), ff);
    e.m[0] -= 1.0;
    e.m[4] -= 1.0;
    e.m[8] -= 1.0;
#pragma unroll
    for (int c = 0; c < 9; ++c)
      eOut(c, i) = e.m[c] * 0.5;
  }
}


// ExplicitTimeIntegrator.cpp:36-52 for a caller-supplied force density.
__global__ void __launch_bounds__(256) k_correct(int64_t n, Planes v, Planes a, Planes force,
                                                 double dt, double density) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double next = force(d, i) / density;
    v(d, i) = v(d, i) + 0.5 * dt * (a(d, i) + next);
    a(d, i) = next;
  }
}

// The bond loop of PdSimpleDamage.cpp:63-84 on its own (critical stretch).
__global__ void __launch_bounds__(256) k_bond_damage(int64_t n, Planes x, UPairs u, Sell bonds,
                                                     uint32_t *broken, uint8_t *anyBroken,
                                                     double initiation, double failure,
                                                     int useCompression) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const StretchTest test = make_stretch_test(initiation, failure, useCompression);
  const double x0 = x(0, i), x1 = x(1, i), x2 = x(2, i);
  const double yi0 = x0 + u(0, i), yi1 = x1 + u(1, i), yi2 = x2 + u(2, i);
  const int len = bonds.rowptr[i + 1] - bonds.rowptr[i];
  for (int k = 0; k < len; ++k) {
    const int64_t p = sell_pos(bonds, i, k);
    const uint32_t bit = 1u << (p & 31);
    if (broken[p >> 5] & bit)
      continue;
    const int j = bonds.col[p];
    const double xj0 = x(0, j), xj1 = x(1, j), xj2 = x(2, j);
    const double xi[3] = {xj0 - x0, xj1 - x1, xj2 - x2};
    const double yij[3] = {(xj0 + u(0, j)) - yi0, (xj1 + u(1, j)) - yi1, (xj2 + u(2, j)) - yi2};
    if (stretch_breaks(test, xi, yij)) {
      atomicOr(&broken[p >> 5], bit);
      anyBroken[i] = 1;
    }
  }
}

// PdSimpleDamage.cpp:23-36: D_i = Σ d V_j / Σ V_j in CSR order.
__global__ void __launch_bounds__(256) k_damage_index(int64_t n, const double *__restrict__ vol,
                                                      Sell bonds,
                                                      const uint32_t *__restrict__ broken,
                                                      double *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  double damaged = 0.0, total = 0.0;
  const int len = bonds.rowptr[i + 1] - bonds.rowptr[i];
  for (int k = 0; k < len; ++k) {
    const int64_t p = sell_pos(bonds, i, k);
    const double vj = vol[bonds.col[p]];
    const double d = (broken[p >> 5] >> (p & 31)) & 1u ? 1.0 : 0.0;
    damaged += d * vj;
    total += vj;
  }
  out[i] = total > 0.0 ? damaged / total : 0.0;
}

// The Johnson-Cook VUMAT on a caller's column-major device block ([k + nblock*c]),
// JohnsonCook.f90:74-156, one lane per point. Unread VUMAT outputs are copied through as
// the routine does (stretch, defgrad, field). MODE selects the return-map solver; seed,
// iterations and fallbacks are optional per-point arrays.
struct VumatArgs {
  int64_t nblock;
  int nstatev, nfieldv;
  JcParams jc;
  double dt, rtol;
  const double *strainInc, *tempOld, *stretchOld, *defgradOld, *fieldOld, *stressOld, *stateOld,
      *enerInternOld, *enerInelasOld, *seed;
  double *tempNew, *stretchNew, *defgradNew, *fieldNew, *stressNew, *stateNew, *enerInternNew,
      *enerInelasNew;
  int *iterations, *fallbacks;
  // The learned mode's seeds hold only the network's (k_seed_block, allSeeds false): a point
  // with a positive state 5 is seeded with it.
  bool seedFromState5;
};

// One point of the device VUMAT: the Johnson-Cook update of point k with the given seed. G > 1:
// the G lanes of a lane group share the point (jc_point's lane groups); the first writes it.
template <int MODE, int G = 1>
__device__ __forceinline__ void vumat_jc_point(const VumatArgs &g, int64_t k, double seed) {
  const int64_t nb = g.nblock;
  const bool lead = G == 1 || (threadIdx.x & (G - 1)) == 0;
  double inc[6], sOld[6], sNew[6], st[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    inc[c] = g.strainInc[k + nb * c];
    sOld[c] = g.stressOld[k + nb * c];
    st[c] = g.stateOld[k + nb * c];
  }
  if (lead) {
    if (g.stretchNew)
      for (int c = 0; c < 6; ++c)
        g.stretchNew[k + nb * c] = g.stretchOld[k + nb * c];
    for (int c = 6; c < g.nstatev; ++c)
      g.stateNew[k + nb * c] = g.stateOld[k + nb * c];
    if (g.defgradNew)
      for (int c = 0; c < 9; ++c)
        g.defgradNew[k + nb * c] = g.defgradOld[k + nb * c];
    if (g.fieldNew)
      for (int c = 0; c < g.nfieldv; ++c)
        g.fieldNew[k + nb * c] = g.fieldOld[k + nb * c];
  }
  double eint = g.enerInternOld[k], einel = g.enerInelasOld[k];
  ReturnMapStats stats;
  jc_point<MODE, false, G>(g.jc, g.dt, inc, sOld, st, eint, einel, sNew, seed, g.rtol, &stats);
  if (!lead)
    return;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    g.stressNew[k + nb * c] = sNew[c];
    g.stateNew[k + nb * c] = st[c];
  }
  if (g.tempNew)
    g.tempNew[k] = st[2];
  g.enerInternNew[k] = eint;
  g.enerInelasNew[k] = einel;
  if (g.iterations)
    g.iterations[k] = stats.iterations;
  if (g.fallbacks)
    g.fallbacks[k] = stats.fallbacks;
}

template <int MODE>
__global__ void __launch_bounds__(256) k_vumat_jc(VumatArgs g) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g.nblock)
    return;
  double seed = 0.0;
  if (g.seed) {
    const double s5 = MODE == RM_LEARNED && g.seedFromState5 ? g.stateOld[k + g.nblock * 4] : 0.0;
    seed = s5 > 0.0 ? s5 : g.seed[k];
  }
  vumat_jc_point<MODE>(g, k, seed);
}

// The parity bisection of the device VUMAT on lane groups of G (8, 16, 32 or 64) per point.
template <int G> __global__ void __launch_bounds__(256) k_vumat_jc_g(VumatArgs g) {
  const int64_t k = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  if (k >= g.nblock)
    return; // the whole group
  vumat_jc_point<RM_BISECTION, G>(g, k, 0.0);
}

} // namespace vpd
