// ORACLE — test infrastructure only. Nothing in the product links, loads or calls this
// file; tests, the smoke check and the CPU-baseline leg of the benchmark use it as the
// checker.
//
// CPU restatement of the reference's explicit correspondence-PD increment with a VUMAT
// material: neighbour search, shape tensor, deformation gradient, Flanagan-Taylor
// corotational kinematics, VUMAT block packing, bond damage, force-state gather,
// velocity-Verlet integrator, lattice geometry, boundary sets and regional constraints,
// and the step loop that strings them together. Each function names the reference
// file:line it follows. Data are held as the reference holds them: points as
// row-major AoS (x[3*i+d], tensors t[9*i+3*r+c]); neighbours as CSR rows sorted
// ascending (the reference's vector<vector<int>>).
//
// Arithmetic keeps the reference's expression order; built with -ffp-contract=off so
// no multiply-add is fused, as in a default x86-64 build of the reference.
//
// The OpenMP placement is the reference's: parallel loops where it has
// `omp parallel for schedule(static)`, serial loops where it is serial (predictor,
// corrector, damage, constraints), so the timed CPU figure is the reference's shape.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" int orc_vumat_jc(int nblock, int ndir, int nshr, int nstatev, int nfieldv,
                            int nprops, const double *dtArray, const double *props,
                            const double *strainInc, const double *tempOld,
                            const double *stretchOld, const double *defgradOld,
                            const double *fieldOld, const double *stressOld,
                            const double *stateOld, const double *enerInternOld,
                            const double *enerInelasOld, double *tempNew, double *stretchNew,
                            double *defgradNew, double *fieldNew, double *stressNew,
                            double *stateNew, double *enerInternNew, double *enerInelasNew);

namespace {

using Vec3 = std::array<double, 3>;
using T3 = std::array<double, 9>;

thread_local std::string g_error;

int fail(const char *msg) {
  g_error = msg;
  return 1;
}

Vec3 load3(const double *p, long i) {
  return {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
}
T3 load9(const double *p, long i) {
  T3 t;
  for (int c = 0; c < 9; ++c)
    t[c] = p[9 * i + c];
  return t;
}
void store9(double *p, long i, const T3 &t) {
  for (int c = 0; c < 9; ++c)
    p[9 * i + c] = t[c];
}

// PdCorrespondenceSkeleton.cpp:18-29
Vec3 vadd(const Vec3 &a, const Vec3 &b) {
  return {a[0] + b[0], a[1] + b[1], a[2] + b[2]};
}
Vec3 vsub(const Vec3 &a, const Vec3 &b) {
  return {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
}
double vnorm(const Vec3 &a) {
  return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
}

// PdCorrespondenceSkeleton.cpp:47-54 and CorotationalKinematics.cpp:27-34: the
// product accumulates from zero in k order.
T3 mul(const T3 &a, const T3 &b) {
  T3 o{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k)
        o[3 * i + j] += a[3 * i + k] * b[3 * k + j];
  return o;
}
// PdCorrespondenceSkeleton.cpp:55-59
Vec3 mulv(const T3 &a, const Vec3 &x) {
  return {a[0] * x[0] + a[1] * x[1] + a[2] * x[2], a[3] * x[0] + a[4] * x[1] + a[5] * x[2],
          a[6] * x[0] + a[7] * x[1] + a[8] * x[2]};
}
T3 transp(const T3 &a) {
  return {a[0], a[3], a[6], a[1], a[4], a[7], a[2], a[5], a[8]};
}
// PdCorrespondenceSkeleton.cpp:63-66 (same form in CorotationalKinematics.cpp:36-39)
double det3(const T3 &a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}
// PdCorrespondenceSkeleton.cpp:67-76 (negated-cofactor form)
bool invSkeleton(const T3 &a, T3 &out) {
  const double d = det3(a);
  if (std::abs(d) < 1.0e-14)
    return false;
  out = {(a[4] * a[8] - a[5] * a[7]) / d, -(a[1] * a[8] - a[2] * a[7]) / d,
         (a[1] * a[5] - a[2] * a[4]) / d, -(a[3] * a[8] - a[5] * a[6]) / d,
         (a[0] * a[8] - a[2] * a[6]) / d, -(a[0] * a[5] - a[2] * a[3]) / d,
         (a[3] * a[7] - a[4] * a[6]) / d, -(a[0] * a[7] - a[1] * a[6]) / d,
         (a[0] * a[4] - a[1] * a[3]) / d};
  return true;
}
// CorotationalKinematics.cpp:41-50 (swapped-operand form)
bool invCorot(const T3 &a, T3 &out) {
  const double d = det3(a);
  if (std::abs(d) < 1.0e-14)
    return false;
  out = {(a[4] * a[8] - a[5] * a[7]) / d, (a[2] * a[7] - a[1] * a[8]) / d,
         (a[1] * a[5] - a[2] * a[4]) / d, (a[5] * a[6] - a[3] * a[8]) / d,
         (a[0] * a[8] - a[2] * a[6]) / d, (a[2] * a[3] - a[0] * a[5]) / d,
         (a[3] * a[7] - a[4] * a[6]) / d, (a[1] * a[6] - a[0] * a[7]) / d,
         (a[0] * a[4] - a[1] * a[3]) / d};
  return true;
}
T3 ident() {
  return {1, 0, 0, 0, 1, 0, 0, 0, 1};
}

// PdCorrespondenceSkeleton.cpp:30-39
bool horizonTolerance(const double *x, long n, double horizon, double &tol) {
  double scale = horizon;
  for (long i = 0; i < 3 * n; ++i) {
    if (!std::isfinite(x[i]))
      return false;
    scale = std::max(scale, std::abs(x[i]));
  }
  tol = 64.0 * std::numeric_limits<double>::epsilon() * scale;
  return true;
}

// PdCorrespondenceSkeleton.cpp:83-91
bool firstPiola(const T3 &stress, const T3 &f, T3 &p) {
  const double j = det3(f);
  if (std::abs(j) < 1.0e-14)
    return false;
  T3 fi;
  if (!invSkeleton(f, fi))
    return false;
  p = mul(stress, transp(fi));
  for (double &v : p)
    v *= j;
  return true;
}

// TensorVoigtMap.cpp:9-32
std::array<double, 6> toVoigt(const T3 &f) {
  return {f[0], f[4], f[8], 0.5 * (f[5] + f[7]), 0.5 * (f[2] + f[6]), 0.5 * (f[1] + f[3])};
}
T3 fromVoigt(const std::array<double, 6> &v) {
  return {v[0], v[5], v[4], v[5], v[1], v[3], v[4], v[3], v[2]};
}

struct Kinematics {
  T3 L, Dbar, relSpin, R1, V1;
};

// CorotationalKinematics.cpp:61-135 (Flanagan-Taylor rate update).
bool corotational(const T3 &fN, const T3 &fNP1, const T3 &vN, const T3 &rN, double dt,
                  Kinematics &out) {
  if (dt <= 0.0)
    return false;
  T3 fDot;
  for (int i = 0; i < 9; ++i)
    fDot[i] = (fNP1[i] - fN[i]) / dt;
  T3 fNi;
  if (!invCorot(fN, fNi))
    return false;
  out.L = mul(fDot, fNi);
  T3 d, w;
  const T3 lt = transp(out.L);
  for (int i = 0; i < 9; ++i) {
    d[i] = 0.5 * (out.L[i] + lt[i]);
    w[i] = 0.5 * (out.L[i] - lt[i]);
  }
  const double zX =
      -vN[2] * d[3] - vN[5] * d[4] - vN[8] * d[5] + vN[1] * d[6] + vN[4] * d[7] + vN[7] * d[8];
  const double zY =
      vN[2] * d[0] + vN[5] * d[1] + vN[8] * d[2] - vN[0] * d[6] - vN[3] * d[7] - vN[6] * d[8];
  const double zZ =
      -vN[1] * d[0] - vN[4] * d[1] - vN[7] * d[2] + vN[0] * d[3] + vN[3] * d[4] + vN[6] * d[5];
  const double wX = 0.5 * (w[7] - w[5]);
  const double wY = 0.5 * (w[2] - w[6]);
  const double wZ = 0.5 * (w[3] - w[1]);
  const double trV = vN[0] + vN[4] + vN[8];
  T3 m = ident();
  for (double &x : m)
    x *= trV;
  for (int i = 0; i < 9; ++i)
    m[i] -= vN[i];
  T3 c;
  if (!invCorot(m, c))
    return false;
  const double oX = wX + c[0] * zX + c[1] * zY + c[2] * zZ;
  const double oY = wY + c[3] * zX + c[4] * zY + c[5] * zZ;
  const double oZ = wZ + c[6] * zX + c[7] * zY + c[8] * zZ;
  const T3 omega = {0.0, -oZ, oY, oZ, 0.0, -oX, -oY, oX, 0.0};
  const double oSq = oX * oX + oY * oY + oZ * oZ;
  T3 q = ident();
  if (oSq > 1.0e-30) {
    const double mag = std::sqrt(oSq);
    const double s1 = std::sin(dt * mag) / mag;
    const double hs = std::sin(0.5 * dt * mag);
    const double s2 = 2.0 * hs * hs / oSq;
    const T3 o2 = mul(omega, omega);
    for (int i = 0; i < 9; ++i)
      q[i] += s1 * omega[i] + s2 * o2[i];
  }
  out.R1 = mul(q, rN);
  const T3 ra = mul(out.L, vN);
  const T3 rb = mul(vN, omega);
  for (int i = 0; i < 9; ++i)
    out.V1[i] = vN[i] + dt * (ra[i] - rb[i]);
  out.Dbar = mul(transp(out.R1), mul(d, out.R1));
  T3 rs;
  for (int i = 0; i < 9; ++i)
    rs[i] = w[i] - omega[i];
  out.relSpin = mul(transp(out.R1), mul(rs, out.R1));
  for (double &x : out.relSpin)
    x *= dt;
  return true;
}

// CorotationalKinematics.cpp:137-139
T3 rotateToSpatial(const T3 &s, const T3 &r) {
  return mul(r, mul(s, transp(r)));
}
// CorotationalKinematics.cpp:141-149
T3 greenLagrange(const T3 &f) {
  T3 e = mul(transp(f), f);
  e[0] -= 1.0;
  e[4] -= 1.0;
  e[8] -= 1.0;
  for (double &x : e)
    x *= 0.5;
  return e;
}

} // namespace

extern "C" {

const char *orc_last_error() {
  return g_error.c_str();
}

int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// ---------------------------------------------------------------- neighbours

// PdCorrespondenceSkeleton.cpp:95-129 — cell list keyed by floor(x / (δ + 2 tol)) in a
// std::map, candidate test ‖x_j − x_i‖ ≤ δ + tol, rows sorted ascending. Returns the
// bond count; rowptr (n+1) is always written, col only when cap >= count. -1 on error.
long orc_build_neighbors(const double *x, long n, double horizon, int *rowptr, int *col,
                         long cap) {
  if (horizon <= 0.0)
    return fail("PD horizon must be positive."), -1;
  double tol;
  if (!horizonTolerance(x, n, horizon, tol))
    return fail("Nonfinite PD coordinate."), -1;
  using Cell = std::tuple<long long, long long, long long>;
  const double radius = horizon + tol;
  const double cw = horizon + 2.0 * tol;
  std::map<Cell, std::vector<int>> cells;
  auto cellOf = [cw](const Vec3 &p) {
    return Cell{(long long)std::floor(p[0] / cw), (long long)std::floor(p[1] / cw),
                (long long)std::floor(p[2] / cw)};
  };
  for (long i = 0; i < n; ++i)
    cells[cellOf(load3(x, i))].push_back((int)i);
  std::vector<std::vector<int>> rows(n);
  for (long i = 0; i < n; ++i) {
    const Vec3 xi = load3(x, i);
    const auto [cx, cy, cz] = cellOf(xi);
    for (long long dx = -1; dx <= 1; ++dx)
      for (long long dy = -1; dy <= 1; ++dy)
        for (long long dz = -1; dz <= 1; ++dz) {
          auto it = cells.find(Cell{cx + dx, cy + dy, cz + dz});
          if (it == cells.end())
            continue;
          for (int j : it->second)
            if (j != i && vnorm(vsub(load3(x, j), xi)) <= radius)
              rows[i].push_back(j);
        }
    std::sort(rows[i].begin(), rows[i].end());
  }
  long nnz = 0;
  rowptr[0] = 0;
  for (long i = 0; i < n; ++i) {
    nnz += (long)rows[i].size();
    rowptr[i + 1] = (int)nnz;
  }
  if (cap >= nnz)
    for (long i = 0; i < n; ++i)
      std::copy(rows[i].begin(), rows[i].end(), col + rowptr[i]);
  return nnz;
}

// PdCorrespondenceSkeleton.cpp:131-142 — all pairs.
long orc_build_neighbors_brute(const double *x, long n, double horizon, int *rowptr, int *col,
                               long cap) {
  if (horizon <= 0.0)
    return fail("PD horizon must be positive."), -1;
  double tol;
  if (!horizonTolerance(x, n, horizon, tol))
    return fail("Nonfinite PD coordinate."), -1;
  const double radius = horizon + tol;
  long nnz = 0;
  rowptr[0] = 0;
  for (long i = 0; i < n; ++i) {
    const Vec3 xi = load3(x, i);
    for (long j = 0; j < n; ++j)
      if (i != j && vnorm(vsub(load3(x, j), xi)) <= radius) {
        if (cap >= 0 && nnz < cap && col)
          col[nnz] = (int)j;
        ++nnz;
      }
    rowptr[i + 1] = (int)nnz;
  }
  return nnz;
}

// ---------------------------------------------------------------- correspondence

// PdCorrespondenceSkeleton.cpp:149-170 — K_i = Σ_j V_j ξ⊗ξ, K⁻¹ by cofactors.
int orc_shape_inverse(const double *x, const double *vol, long n, const int *rowptr,
                      const int *col, double *kinv) {
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (long i = 0; i < n; ++i) {
    T3 k{};
    const Vec3 xi0 = load3(x, i);
    for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
      const long j = col[p];
      if (j < 0 || j >= n) {
        bad |= 2;
        continue;
      }
      const Vec3 xi = vsub(load3(x, j), xi0);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          k[3 * a + b] += vol[j] * xi[a] * xi[b];
    }
    T3 ki;
    if (!invSkeleton(k, ki))
      bad |= 1;
    else
      store9(kinv, i, ki);
  }
  if (bad & 2)
    return fail("PD neighbor index is out of range.");
  if (bad & 1)
    return fail("Singular tensor in PD shape tensor.");
  return 0;
}

// PdCorrespondenceSkeleton.cpp:172-197 — F_i = (Σ_j V_j (y_j − y_i)⊗ξ_ij) K_i⁻¹.
int orc_deformation_gradient(const double *x, const double *u, const double *vol, long n,
                             const int *rowptr, const int *col, const double *kinv, double *f) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    T3 num{};
    const Vec3 x0 = load3(x, i);
    const Vec3 yi = vadd(x0, load3(u, i));
    for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
      const long j = col[p];
      const Vec3 xj = load3(x, j);
      const Vec3 xi = vsub(xj, x0);
      const Vec3 yij = vsub(vadd(xj, load3(u, j)), yi);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          num[3 * a + b] += vol[j] * yij[a] * xi[b];
    }
    store9(f, i, mul(num, load9(kinv, i)));
  }
  return 0;
}

// PdCorrespondenceSkeleton.cpp:199-230 — A_i = (J σ F⁻ᵀ) K_i⁻¹, then the gather
// f_i = Σ_k (1 − clamp(d_ik)) (A_i ξ − A_j(−ξ)) V_j in CSR order. bondDamage is one
// value per CSR entry (the reference's vector<vector<double>>), may be null (intact).
int orc_force_density(const double *x, const double *vol, long n, const int *rowptr,
                      const int *col, const double *kinv, const double *f, const double *sigma,
                      const double *bondDamage, double *force, double *piolaK) {
  std::vector<double> map(9 * (size_t)n);
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (long i = 0; i < n; ++i) {
    T3 p;
    if (!firstPiola(load9(sigma, i), load9(f, i), p)) {
      bad = 1;
      continue;
    }
    store9(map.data(), i, mul(p, load9(kinv, i)));
  }
  if (bad)
    return fail("Invalid deformation gradient determinant.");
  if (piolaK)
    std::copy(map.begin(), map.end(), piolaK);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) {
    Vec3 acc{0.0, 0.0, 0.0};
    const Vec3 x0 = load3(x, i);
    const T3 ai = load9(map.data(), i);
    for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
      const long j = col[p];
      const Vec3 xi = vsub(load3(x, j), x0);
      const Vec3 ti = mulv(ai, xi);
      const Vec3 nxi{-xi[0], -xi[1], -xi[2]};
      const Vec3 tj = mulv(load9(map.data(), j), nxi);
      const double dmg = bondDamage ? bondDamage[p] : 0.0;
      const double intact = 1.0 - std::clamp(dmg, 0.0, 1.0);
      for (int d = 0; d < 3; ++d)
        acc[d] += intact * (ti[d] - tj[d]) * vol[j];
    }
    for (int d = 0; d < 3; ++d)
      force[3 * i + d] = acc[d];
  }
  return 0;
}

// ---------------------------------------------------------------- kinematics

// CorotationalKinematics.cpp:61-135 over n points. Any output pointer may be null.
int orc_corotational(const double *fN, const double *fNP1, const double *vN, const double *rN,
                     double dt, long n, double *L, double *Dbar, double *relSpin, double *R1,
                     double *V1) {
  for (long i = 0; i < n; ++i) {
    Kinematics k;
    if (!corotational(load9(fN, i), load9(fNP1, i), load9(vN, i), load9(rN, i), dt, k))
      return fail(dt <= 0.0 ? "Corotational update requires dt > 0."
                            : "Corotational update encountered a singular tensor.");
    if (L)
      store9(L, i, k.L);
    if (Dbar)
      store9(Dbar, i, k.Dbar);
    if (relSpin)
      store9(relSpin, i, k.relSpin);
    if (R1)
      store9(R1, i, k.R1);
    if (V1)
      store9(V1, i, k.V1);
  }
  return 0;
}

int orc_rotate_to_spatial(const double *s, const double *r, long n, double *out) {
  for (long i = 0; i < n; ++i)
    store9(out, i, rotateToSpatial(load9(s, i), load9(r, i)));
  return 0;
}

int orc_green_lagrange(const double *f, long n, double *out) {
  for (long i = 0; i < n; ++i)
    store9(out, i, greenLagrange(load9(f, i)));
  return 0;
}

// ---------------------------------------------------------------- material

// VumatPdCorrespondenceMaterial.cpp:48-70 — initial history for n points.
int orc_material_init(const double *f0, long n, int nstatev, const double *initialState,
                      double tref, double *fN, double *vN, double *rN, double *strainN,
                      double *stressN, double *statev, double *temp, double *eint,
                      double *einel) {
  if (n == 0)
    return fail("VUMAT-PD material bridge requires at least one point.");
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i)
    store9(strainN, i, greenLagrange(load9(f0, i)));
  for (long i = 0; i < n; ++i) {
    store9(fN, i, load9(f0, i));
    store9(vN, i, ident());
    store9(rN, i, ident());
    store9(stressN, i, T3{});
    for (int c = 0; c < nstatev; ++c)
      statev[(size_t)nstatev * i + c] = initialState ? initialState[c] : 0.0;
    temp[i] = tref;
    eint[i] = 0.0;
    einel[i] = 0.0;
  }
  return 0;
}

// VumatPdCorrespondenceMaterial.cpp:72-176 with the JC routine behind the packing of
// VumatMaterialAdapter.cpp:238-362: 256-point blocks, last row duplicated into the
// padding, FT kinematics per point, Voigt packing (stress [11,22,33,23,13,12] with
// symmetrised shear; strainInc = D̄·dt), the VUMAT call, σ = R σ̄ Rᵀ, history commit.
// State arrays are updated in place; sigmaSpatial (n×9) and strain (n×9) are outputs.
int orc_material_update(const double *fNew, long n, double dt, int nstatev, int nprops,
                        const double *props, double *fN, double *vN, double *rN,
                        double *strainN, double *stressN, double *statev, double *temp,
                        double *eint, double *einel, double *sigmaSpatial) {
  if (dt <= 0.0)
    return fail("VUMAT-PD material update requires dt > 0.");
  const long width = 256;
  const long batches = (n + width - 1) / width;
  const int ntens = 6, nfield = 1;
  int bad = 0;
#pragma omp parallel reduction(| : bad)
  {
    std::vector<double> sInc(width * ntens), sOld(width * ntens), stOld(width * nstatev);
    std::vector<double> tOld(width), eiOld(width), elOld(width);
    std::vector<double> zero6(width * ntens, 0.0), zero9(width * 9, 0.0), zf(width * nfield, 0.0);
    std::vector<double> tNew(width), s6New(width * ntens), d9New(width * 9), fNew1(width * nfield);
    std::vector<double> sNew(width * ntens), stNew(width * nstatev), eiNew(width), elNew(width);
    std::vector<T3> rot(width), str(width), stretch(width);
#pragma omp for schedule(static)
    for (long b = 0; b < batches; ++b) {
      const long begin = b * width, count = std::min(width, n - begin);
      for (long jj = 0; jj < width; ++jj) {
        const long j = jj < count ? jj : count - 1;
        const long i = begin + j;
        if (jj < count) {
          Kinematics k;
          if (!corotational(load9(fN, i), load9(fNew, i), load9(vN, i), load9(rN, i), dt, k)) {
            bad = 1;
            continue;
          }
          rot[jj] = k.R1;
          stretch[jj] = k.V1;
          str[jj] = greenLagrange(load9(fNew, i));
          auto e = toVoigt(k.Dbar);
          for (double &v : e)
            v *= dt;
          const auto s = toVoigt(load9(stressN, i));
          for (int c = 0; c < ntens; ++c) {
            sInc[jj + width * c] = e[c];
            sOld[jj + width * c] = s[c];
          }
          for (int c = 0; c < nstatev; ++c)
            stOld[jj + width * c] = statev[(size_t)nstatev * i + c];
          tOld[jj] = temp[i];
          eiOld[jj] = eint[i];
          elOld[jj] = einel[i];
        } else {
          for (int c = 0; c < ntens; ++c) {
            sInc[jj + width * c] = sInc[j + width * c];
            sOld[jj + width * c] = sOld[j + width * c];
          }
          for (int c = 0; c < nstatev; ++c)
            stOld[jj + width * c] = stOld[j + width * c];
          tOld[jj] = tOld[j];
          eiOld[jj] = eiOld[j];
          elOld[jj] = elOld[j];
        }
      }
      if (orc_vumat_jc((int)width, 3, 3, nstatev, nfield, nprops, &dt, props, sInc.data(),
                       tOld.data(), zero6.data(), zero9.data(), zf.data(), sOld.data(),
                       stOld.data(), eiOld.data(), elOld.data(), tNew.data(), s6New.data(),
                       d9New.data(), fNew1.data(), sNew.data(), stNew.data(), eiNew.data(),
                       elNew.data()) != 0) {
        bad = 2;
        continue;
      }
      for (long j = 0; j < count; ++j) {
        const long i = begin + j;
        std::array<double, 6> v;
        for (int c = 0; c < ntens; ++c)
          v[c] = sNew[j + width * c];
        const T3 sFull = fromVoigt(v);
        store9(sigmaSpatial, i, rotateToSpatial(sFull, rot[j]));
        store9(fN, i, load9(fNew, i));
        store9(vN, i, stretch[j]);
        store9(rN, i, rot[j]);
        store9(strainN, i, str[j]);
        store9(stressN, i, sFull);
        for (int c = 0; c < nstatev; ++c)
          statev[(size_t)nstatev * i + c] = stNew[j + width * c];
        temp[i] = tNew[j];
        eint[i] = eiNew[j];
        einel[i] = elNew[j];
      }
    }
  }
  if (bad == 1)
    return fail("Corotational update encountered a singular tensor.");
  if (bad)
    return fail("JohnsonCook VUMAT dimension error.");
  return 0;
}

// ---------------------------------------------------------------- damage

// PdSimpleDamage.cpp:51-96, critical_stretch mode, serial as in the reference. damage is
// one value (0/1) per CSR entry and is updated in place (irreversible). damageIndex may
// be null.
int orc_damage_update(const double *x, const double *u, const double *vol, long n,
                      const int *rowptr, const int *col, double initiation, double failure,
                      double *damage, double *damageIndex) {
  if (initiation <= 0.0)
    return fail("critical_stretch requires initiation > 0.");
  const bool useCompression = failure > 0.0 && std::isfinite(failure);
  for (long i = 0; i < n; ++i) {
    const Vec3 x0 = load3(x, i);
    const Vec3 yi = vadd(x0, load3(u, i));
    for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
      const long j = col[p];
      const Vec3 xj = load3(x, j);
      const Vec3 xi = vsub(xj, x0);
      const double initial = vnorm(xi);
      const double stretch = (vnorm(vsub(vadd(xj, load3(u, j)), yi)) - initial) / initial;
      if (stretch > initiation || (useCompression && stretch < -failure))
        damage[p] = 1.0;
    }
  }
  if (damageIndex)
    for (long i = 0; i < n; ++i) {
      double damaged = 0.0, total = 0.0;
      for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
        damaged += damage[p] * vol[col[p]];
        total += vol[col[p]];
      }
      damageIndex[i] = total > 0.0 ? damaged / total : 0.0;
    }
  return 0;
}

// ---------------------------------------------------------------- integrator

// ExplicitTimeIntegrator.cpp:23-34
void orc_predict(double *u, const double *v, const double *a, long n, double dt) {
  for (long i = 0; i < 3 * n; ++i)
    u[i] += dt * v[i] + 0.5 * dt * dt * a[i];
}
// ExplicitTimeIntegrator.cpp:36-52
void orc_correct(double *v, double *a, const double *force, long n, double dt, double density) {
  for (long i = 0; i < n; ++i) {
    double next[3];
    for (int d = 0; d < 3; ++d) {
      next[d] = force[3 * i + d] / density;
      v[3 * i + d] += 0.5 * dt * (a[3 * i + d] + next[d]);
    }
    for (int d = 0; d < 3; ++d)
      a[3 * i + d] = next[d];
  }
}

// ---------------------------------------------------------------- host geometry and sets

// Geometry.cpp:10-34 — lattice, index (i·ny + j)·nz + k, trapezoid volumes.
void orc_make_patch(int nx, int ny, int nz, double h, double *x, double *vol) {
  const double yOff = 0.5 * (double)(ny - 1) * h, zOff = 0.5 * (double)(nz - 1) * h;
  const double v = h * h * h;
  long p = 0;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k, ++p) {
        x[3 * p] = (double)i * h;
        x[3 * p + 1] = (double)j * h - yOff;
        x[3 * p + 2] = (double)k * h - zOff;
        vol[p] = v * ((i == 0 || i == nx - 1) ? 0.5 : 1.0) * ((j == 0 || j == ny - 1) ? 0.5 : 1.0) *
                 ((k == 0 || k == nz - 1) ? 0.5 : 1.0);
      }
}

// ParticleSets.cpp:12-50. side: 0 left, 1 right, 2 bottom, 3 top. Writes a 0/1 mask.
int orc_select_region(const double *x, long n, int side, double width, unsigned char *mask) {
  if (width <= 0.0 || n == 0)
    return fail("Boundary-region width and particle set must be nonzero.");
  const int comp = side < 2 ? 0 : 1;
  const bool minimum = side == 0 || side == 2;
  double extreme = minimum ? std::numeric_limits<double>::infinity()
                           : -std::numeric_limits<double>::infinity();
  for (long i = 0; i < n; ++i)
    extreme = minimum ? std::min(extreme, x[3 * i + comp]) : std::max(extreme, x[3 * i + comp]);
  const double tol = 1.0e-10 * std::max(1.0, std::abs(extreme));
  long count = 0;
  for (long i = 0; i < n; ++i) {
    const double dist = minimum ? x[3 * i + comp] - extreme : extreme - x[3 * i + comp];
    mask[i] = dist <= width + tol;
    count += mask[i];
  }
  if (count == 0)
    return fail("Boundary-region selection returned no particles.");
  return 0;
}

// BoundaryConditions.cpp:54-68 and :81-115 reduced to a per-particle, per-component
// code that records which prescription the sequence of writes in apply() leaves
// last: 0 free, 1 held at zero, 2 follows the ramp. Also returns anchor and rotation
// pin. loadComp 0 (x) or 1 (y); support/load sides as orc_select_region.
int orc_constraint_codes(const double *x, long n, int loadComp, int supportSide, int loadSide,
                         int regionLayers, double spacing, int planeStrain, unsigned char *codes,
                         long *anchorOut, long *pinOut) {
  const double width = (double)(regionLayers - 1) * spacing + 1.0e-9 * spacing;
  std::vector<unsigned char> sup(n), ld(n);
  if (orc_select_region(x, n, supportSide, width, sup.data()) ||
      orc_select_region(x, n, loadSide, width, ld.data()))
    return 1;
  double c[3] = {0.0, 0.0, 0.0};
  for (long i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d)
      c[d] += x[3 * i + d];
  for (double &v : c)
    v /= (double)n;
  // ParticleSets.cpp:52-71 over the support set in index order.
  long anchor = -1;
  double best = std::numeric_limits<double>::infinity();
  for (long i = 0; i < n; ++i) {
    if (!sup[i])
      continue;
    if (anchor < 0)
      anchor = i;
    double dist = 0.0;
    for (int d = 0; d < 3; ++d) {
      const double dl = x[3 * i + d] - c[d];
      dist += dl * dl;
    }
    if (dist < best) {
      best = dist;
      anchor = i;
    }
  }
  const int lateral = loadComp == 0 ? 1 : 0;
  long pin = anchor;
  for (long i = 0; i < 3 * n; ++i)
    codes[i] = 0;
  for (long i = 0; i < n; ++i)
    if (sup[i])
      codes[3 * i + loadComp] = 1;
  for (long i = 0; i < n; ++i)
    if (ld[i])
      codes[3 * i + loadComp] = 2;
  codes[3 * anchor + lateral] = 1;
  if (!planeStrain) {
    codes[3 * anchor + 2] = 1;
    double greatestY = -1.0e100;
    for (long i = 0; i < n; ++i) {
      if (!sup[i])
        continue;
      if (std::abs(x[3 * i] - x[3 * anchor]) < 1e-8 &&
          std::abs(x[3 * i + 2] - x[3 * anchor + 2]) < 1e-8 && x[3 * i + 1] > greatestY) {
        greatestY = x[3 * i + 1];
        pin = i;
      }
    }
    codes[3 * pin + 2] = 1;
  } else {
    for (long i = 0; i < n; ++i)
      codes[3 * i + 2] = 1;
  }
  if (anchorOut)
    *anchorOut = anchor;
  if (pinOut)
    *pinOut = pin;
  return 0;
}

// BoundaryConditions.cpp:70-79 — half-cosine ramp then hold.
void orc_ramp(double amplitude, double riseTime, double time, double *out3) {
  if (time >= riseTime) {
    out3[0] = amplitude;
    out3[1] = 0.0;
    out3[2] = 0.0;
    return;
  }
  const double pi = 3.14159265358979323846;
  const double phase = pi * time / riseTime;
  out3[0] = 0.5 * amplitude * (1.0 - std::cos(phase));
  out3[1] = 0.5 * amplitude * pi / riseTime * std::sin(phase);
  out3[2] = 0.5 * amplitude * pi * pi / (riseTime * riseTime) * std::cos(phase);
}

// BoundaryConditions.cpp:81-115 applied through the per-component codes.
void orc_apply_constraints(const unsigned char *codes, long n, const double *ramp3, double *u,
                           double *v, double *a) {
  for (long i = 0; i < 3 * n; ++i) {
    if (codes[i] == 1) {
      u[i] = 0.0;
      v[i] = 0.0;
      a[i] = 0.0;
    } else if (codes[i] == 2) {
      u[i] = ramp3[0];
      v[i] = ramp3[1];
      a[i] = ramp3[2];
    }
  }
}

// ---------------------------------------------------------------- the step loop

// One increment of BoundaryValueCaseRunner.cpp:207-229 (without its diagnostics):
// predict → constrain(t) → F → material(dt, (step−1)dt) → damage → force → correct →
// constrain(t). Every array is in/out state of the case. damage may be null (mode none).
// ramp3 holds (u, v, a) of the loaded band at time t = step·dt.
int orc_step(const double *x, const double *vol, long n, const int *rowptr, const int *col,
             const double *kinv, const unsigned char *codes, const double *ramp3, double dt,
             double density, int nstatev, int nprops, const double *props, double initiation,
             double failure, double *u, double *v, double *a, double *fN, double *vN,
             double *rN, double *strainN, double *stressN, double *statev, double *temp,
             double *eint, double *einel, double *damage, double *fOut, double *sigmaOut,
             double *forceOut) {
  orc_predict(u, v, a, n, dt);
  orc_apply_constraints(codes, n, ramp3, u, v, a);
  if (orc_deformation_gradient(x, u, vol, n, rowptr, col, kinv, fOut))
    return 1;
  if (orc_material_update(fOut, n, dt, nstatev, nprops, props, fN, vN, rN, strainN, stressN,
                          statev, temp, eint, einel, sigmaOut))
    return 1;
  if (damage && orc_damage_update(x, u, vol, n, rowptr, col, initiation, failure, damage, nullptr))
    return 1;
  if (orc_force_density(x, vol, n, rowptr, col, kinv, fOut, sigmaOut, damage, forceOut, nullptr))
    return 1;
  orc_correct(v, a, forceOut, n, dt, density);
  orc_apply_constraints(codes, n, ramp3, u, v, a);
  return 0;
}

} // extern "C"
