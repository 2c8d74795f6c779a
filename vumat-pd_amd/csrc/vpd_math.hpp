// Device-side FP64 tensor algebra and the point-wise constitutive pieces of the PD
// increment: Flanagan-Taylor corotational update, Voigt packing, Johnson-Cook radial
// return, first Piola stress. Every routine keeps the operand order of the reference
// expression it stands for (cited per function), and the library is compiled with
// -ffp-contract=off, so +, -, *, / and sqrt round exactly as the reference's CPU build
// does; sin, pow and log come from the device math library (within an ulp of libm).
//
// Tensors are 3x3 row-major in registers: t[3*r + c].
#pragma once

#include <hip/hip_runtime.h>

namespace vpd {

// Planes of an SoA field with stride N.
struct Planes {
  double *p;
  int64_t n;
  __device__ __forceinline__ double &operator()(int c, int64_t i) const { return p[c * n + i]; }
};

// Planes of component pairs: pair c of particle i is the 16-byte (component 2c, 2c+1) at
// p[c·n + i]. The bond passes gather neighbour data in this form: one 16-byte load per lane
// brings two components, and the 4 lanes of a quad, reading 4 consecutive particles, touch
// 64 contiguous bytes.
typedef double d2v __attribute__((ext_vector_type(2)));
struct PairPlanes {
  d2v *p;
  int64_t n;
  __device__ __forceinline__ d2v &operator()(int c, int64_t i) const { return p[c * n + i]; }
};

// Displacement as two pair planes, (u_x, u_y) and (u_z, 0): the F pass gathers u_j with two
// 16-byte loads instead of three 8-byte ones. u(c, i) addresses one component.
struct UPairs {
  double *p;
  int64_t n;
  __device__ __forceinline__ double &operator()(int c, int64_t i) const {
    return c < 2 ? p[2 * i + c] : p[2 * (n + i)];
  }
  __device__ __forceinline__ d2v pair(int h, int64_t i) const {
    return reinterpret_cast<const d2v *>(p)[h * n + i];
  }
};

struct T3 {
  double m[9];
};
struct V3 {
  double v[3];
};

__device__ __forceinline__ T3 t3_ident() {
  T3 t;
#pragma unroll
  for (int c = 0; c < 9; ++c)
    t.m[c] = (c % 4 == 0) ? 1.0 : 0.0;
  return t;
}

// PdCorrespondenceSkeleton.cpp:47-54 / CorotationalKinematics.cpp:27-34: o(i,j) starts
// at zero and accumulates a(i,k) b(k,j) for k = 0, 1, 2.
__device__ __forceinline__ T3 t3_mul(const T3 &a, const T3 &b) {
  T3 o;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k)
        s += a.m[3 * i + k] * b.m[3 * k + j];
      o.m[3 * i + j] = s;
    }
  return o;
}

// PdCorrespondenceSkeleton.cpp:55-59
__device__ __forceinline__ V3 t3_mulv(const T3 &a, const V3 &x) {
  V3 o;
  o.v[0] = a.m[0] * x.v[0] + a.m[1] * x.v[1] + a.m[2] * x.v[2];
  o.v[1] = a.m[3] * x.v[0] + a.m[4] * x.v[1] + a.m[5] * x.v[2];
  o.v[2] = a.m[6] * x.v[0] + a.m[7] * x.v[1] + a.m[8] * x.v[2];
  return o;
}

__device__ __forceinline__ T3 t3_transpose(const T3 &a) {
  T3 o;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      o.m[3 * i + j] = a.m[3 * j + i];
  return o;
}

// PdCorrespondenceSkeleton.cpp:63-66
__device__ __forceinline__ double t3_det(const T3 &t) {
  const double *a = t.m;
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}

// PdCorrespondenceSkeleton.cpp:67-76: cofactors (off-diagonal ones negated) over det.
__device__ __forceinline__ bool t3_inverse_skeleton(const T3 &t, T3 &o) {
  const double *a = t.m;
  const double d = t3_det(t);
  if (fabs(d) < 1.0e-14)
    return false;
  o.m[0] = (a[4] * a[8] - a[5] * a[7]) / d;
  o.m[1] = -(a[1] * a[8] - a[2] * a[7]) / d;
  o.m[2] = (a[1] * a[5] - a[2] * a[4]) / d;
  o.m[3] = -(a[3] * a[8] - a[5] * a[6]) / d;
  o.m[4] = (a[0] * a[8] - a[2] * a[6]) / d;
  o.m[5] = -(a[0] * a[5] - a[2] * a[3]) / d;
  o.m[6] = (a[3] * a[7] - a[4] * a[6]) / d;
  o.m[7] = -(a[0] * a[7] - a[1] * a[6]) / d;
  o.m[8] = (a[0] * a[4] - a[1] * a[3]) / d;
  return true;
}

// CorotationalKinematics.cpp:41-50: the same cofactors written with swapped operands.
__device__ __forceinline__ bool t3_inverse_corot(const T3 &t, T3 &o) {
  const double *a = t.m;
  const double d = t3_det(t);
  if (fabs(d) < 1.0e-14)
    return false;
  o.m[0] = (a[4] * a[8] - a[5] * a[7]) / d;
  o.m[1] = (a[2] * a[7] - a[1] * a[8]) / d;
  o.m[2] = (a[1] * a[5] - a[2] * a[4]) / d;
  o.m[3] = (a[5] * a[6] - a[3] * a[8]) / d;
  o.m[4] = (a[0] * a[8] - a[2] * a[6]) / d;
  o.m[5] = (a[2] * a[3] - a[0] * a[5]) / d;
  o.m[6] = (a[3] * a[7] - a[4] * a[6]) / d;
  o.m[7] = (a[1] * a[6] - a[0] * a[7]) / d;
  o.m[8] = (a[0] * a[4] - a[1] * a[3]) / d;
  return true;
}

// Error bits a kernel raises through the context's device flag.
enum : int {
  ERR_SINGULAR_SHAPE = 1,
  ERR_SINGULAR_COROT = 2,
  ERR_PIOLA_DET = 4,
  ERR_NEIGHBOR_RANGE = 8,
  ERR_NONFINITE_COORD = 16,
  ERR_NEIGHBOR_OVERFLOW = 32,
  ERR_VUMAT_DIMS = 64,
};

struct Corot {
  T3 R1, V1, Dbar;
};

// CorotationalKinematics.cpp:61-135 (Flanagan-Taylor). Also hands back L, W and Ω so the
// per-point API can form the relative-spin increment (:125-134); the fused material kernel
// drops them, since they feed only VUMAT arguments the Johnson-Cook routine never reads.
__device__ __forceinline__ bool corotational_full(const T3 &fN, const T3 &fNP1, const T3 &vNt,
                                                  const T3 &rN, double dt, Corot &out, T3 &L,
                                                  T3 &w, T3 &om) {
  T3 fDot;
#pragma unroll
  for (int i = 0; i < 9; ++i)
    fDot.m[i] = (fNP1.m[i] - fN.m[i]) / dt;
  T3 fNi;
  if (!t3_inverse_corot(fN, fNi))
    return false;
  L = t3_mul(fDot, fNi);
  T3 d;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      d.m[3 * r + c] = 0.5 * (L.m[3 * r + c] + L.m[3 * c + r]);
      w.m[3 * r + c] = 0.5 * (L.m[3 * r + c] - L.m[3 * c + r]);
    }
  const double *vN = vNt.m;
  const double zX =
      -vN[2] * d.m[3] - vN[5] * d.m[4] - vN[8] * d.m[5] + vN[1] * d.m[6] + vN[4] * d.m[7] +
      vN[7] * d.m[8];
  const double zY = vN[2] * d.m[0] + vN[5] * d.m[1] + vN[8] * d.m[2] - vN[0] * d.m[6] -
                    vN[3] * d.m[7] - vN[6] * d.m[8];
  const double zZ =
      -vN[1] * d.m[0] - vN[4] * d.m[1] - vN[7] * d.m[2] + vN[0] * d.m[3] + vN[3] * d.m[4] +
      vN[6] * d.m[5];
  const double wX = 0.5 * (w.m[7] - w.m[5]);
  const double wY = 0.5 * (w.m[2] - w.m[6]);
  const double wZ = 0.5 * (w.m[3] - w.m[1]);
  const double trV = vN[0] + vN[4] + vN[8];
  T3 m = t3_ident();
#pragma unroll
  for (int i = 0; i < 9; ++i)
    m.m[i] = m.m[i] * trV - vN[i];
  T3 c;
  if (!t3_inverse_corot(m, c))
    return false;
  const double oX = wX + c.m[0] * zX + c.m[1] * zY + c.m[2] * zZ;
  const double oY = wY + c.m[3] * zX + c.m[4] * zY + c.m[5] * zZ;
  const double oZ = wZ + c.m[6] * zX + c.m[7] * zY + c.m[8] * zZ;
  om.m[0] = 0.0;
  om.m[1] = -oZ;
  om.m[2] = oY;
  om.m[3] = oZ;
  om.m[4] = 0.0;
  om.m[5] = -oX;
  om.m[6] = -oY;
  om.m[7] = oX;
  om.m[8] = 0.0;
  const double oSq = oX * oX + oY * oY + oZ * oZ;
  T3 q = t3_ident();
  if (oSq > 1.0e-30) {
    const double mag = sqrt(oSq);
    const double s1 = sin(dt * mag) / mag;
    const double hs = sin(0.5 * dt * mag);
    const double s2 = 2.0 * hs * hs / oSq;
    const T3 o2 = t3_mul(om, om);
#pragma unroll
    for (int i = 0; i < 9; ++i)
      q.m[i] += s1 * om.m[i] + s2 * o2.m[i];
  }
  out.R1 = t3_mul(q, rN);
  const T3 ra = t3_mul(L, vNt);
  const T3 rb = t3_mul(vNt, om);
#pragma unroll
  for (int i = 0; i < 9; ++i)
    out.V1.m[i] = vN[i] + dt * (ra.m[i] - rb.m[i]);
  out.Dbar = t3_mul(t3_transpose(out.R1), t3_mul(d, out.R1));
  return true;
}

__device__ __forceinline__ bool corotational(const T3 &fN, const T3 &fNP1, const T3 &vNt,
                                             const T3 &rN, double dt, Corot &out) {
  T3 L, w, om;
  return corotational_full(fN, fNP1, vNt, rN, dt, out, L, w, om);
}

// TensorVoigtMap.cpp:9-16: [11, 22, 33, 23, 13, 12], shears symmetrised.
__device__ __forceinline__ void to_voigt(const T3 &f, double v[6]) {
  v[0] = f.m[0];
  v[1] = f.m[4];
  v[2] = f.m[8];
  v[3] = 0.5 * (f.m[5] + f.m[7]);
  v[4] = 0.5 * (f.m[2] + f.m[6]);
  v[5] = 0.5 * (f.m[1] + f.m[3]);
}
// TensorVoigtMap.cpp:18-20
__device__ __forceinline__ T3 from_voigt(const double v[6]) {
  T3 t;
  t.m[0] = v[0];
  t.m[1] = v[5];
  t.m[2] = v[4];
  t.m[3] = v[5];
  t.m[4] = v[1];
  t.m[5] = v[3];
  t.m[6] = v[4];
  t.m[7] = v[3];
  t.m[8] = v[2];
  return t;
}

// PdCorrespondenceSkeleton.cpp:83-91: P = J σ F⁻ᵀ.
__device__ __forceinline__ bool first_piola(const T3 &s, const T3 &f, T3 &p) {
  const double j = t3_det(f);
  if (fabs(j) < 1.0e-14)
    return false;
  T3 fi;
  if (!t3_inverse_skeleton(f, fi))
    return false;
  p = t3_mul(s, t3_transpose(fi));
#pragma unroll
  for (int i = 0; i < 9; ++i)
    p.m[i] *= j;
  return true;
}

// Johnson-Cook parameters (props 1..14 of materials/examples/JohnsonCook.f90:9) and the
// derived constants of :70-72.
struct JcParams {
  double young, poisson, a0, b0, rateC, hardN, thermalM, rate0, tref, tmelt, rho, beta, cp,
      workToHeat;
  double mu, lambda, tolerance;
};

// JohnsonCook.f90:160-170, split in the thermal factor 1 − θᵐ (which a return-map solve,
// at fixed temperature, evaluates once) and the rest, multiplied in the reference's order:
// (A + B ε̄ⁿ)·(1 − θᵐ)·(1 + C ln(ε̇/ε̇₀)).
__device__ __forceinline__ double jc_softening(const JcParams &p, double temp) {
  double theta = (temp - p.tref) / (p.tmelt - p.tref);
  theta = fmin(1.0, fmax(0.0, theta));
  return 1.0 - pow(theta, p.thermalM);
}
__device__ __forceinline__ double jc_yield_soft(const JcParams &p, double ep, double rate,
                                                double soft) {
  double er = rate;
  if (er == 0.0)
    er = p.rate0;
  return (p.a0 + p.b0 * pow(ep, p.hardN)) * soft * (1.0 + p.rateC * log(er / p.rate0));
}
__device__ __forceinline__ double jc_yield(const JcParams &p, double ep, double rate,
                                           double temp) {
  return jc_yield_soft(p, ep, rate, jc_softening(p, temp));
}

// Screen of the bisection's yield evaluation. The bisection's iterates depend only on the
// sign of each residual and on whether it is below the tolerance, and it stores only the last
// evaluated yield; so an iteration whose residual is far from 0 and from ±tol can take its
// decision from an estimate. The estimate forms (A + B ε̄ⁿ)·soft·(1 + C ln ρ), ρ = Δ/(dt ε̇₀),
// with log2 x = e + log2(m) for x = m·2^e (m in [0.5, 1), exponent exact, the FP32 hardware
// log2 of m rounded to FP32) and 2^t = 2^⌊t⌋·2^(t − ⌊t⌋) (the FP32 hardware exp2 of the
// fraction rounded to FP32); the rest is FP64. Error budget (FP32 steps; FP64 rounding is
// covered by the caller's 1e-14 term):
//  * log2 x: the hardware log2 errs by at most kScreenLog2Hw on every float in [0.5, 1]
//    (measured exhaustively, vpd_screen_accuracy), and rounding m to FP32 moves log2 m by at
//    most 2⁻²⁴/ln 2;
//  * 2^f: the hardware exp2 errs by a relative kScreenExp2Hw at most on every float in
//    [0, 1], and rounding t − ⌊t⌋ to FP32 moves it by 2⁻²⁵ at most;
//  so ε̄ⁿ errs by a relative e_pw = ln 2·(|n|·e_log + 2⁻²⁵) + kScreenExp2Hw and ln ρ by
// e_lr = ln 2·e_log absolutely (e_log = kScreenLog2Hw + 2⁻²⁴/ln 2), and the estimate by at most
//   |soft|·(|B ε̄ⁿ|·|1 + C ln ρ|·e_pw + |A + B ε̄ⁿ|·|C|·e_lr)
// to first order; `margin` is four times that.
// Returns false (no estimate) for arguments outside the range the FP32 steps cover.
constexpr double kScreenLog2Hw = 6.0e-8; // ≥ 5.945e-8 measured on MI355X (0.997·2⁻²⁴)
constexpr double kScreenExp2Hw = 8.4e-8; // ≥ 8.317e-8 measured on MI355X (0.698·2⁻²³)
__device__ __forceinline__ double screen_log2(double x) { // x > 0, normal
  const float m = (float)__builtin_amdgcn_frexp_mant(x);
  return (double)__builtin_amdgcn_frexp_exp(x) + (double)__builtin_amdgcn_logf(m);
}
// The estimate's two transcendental terms: ε̄ⁿ (pw) and ln ρ (lr, 0 for Δ = 0: the rate is
// then ε̇₀). False for arguments outside the range the FP32 steps cover (pow(0, n) for n ≤ 0
// included: left to the exact evaluation). Branch-free: each term is evaluated on its
// argument or, where that is out of range, on 1, and the results are selected, so that the
// checks (which in practice always pass) cost selects rather than divergent branches.
__device__ __forceinline__ bool screen_terms(const JcParams &p, double ep, double dep,
                                             double rateScale, double &pw, double &lr) {
  const bool epZero = ep == 0.0;
  const bool epOk = ep > 1e-300 && ep < 1e300;
  const double t = p.hardN * screen_log2(epOk ? ep : 1.0);
  const bool tOk = fabs(t) < 1000.0;
  const double tc = tOk ? t : 0.0;
  const double k = floor(tc);
  const double pwEst =
      __builtin_amdgcn_ldexp((double)__builtin_amdgcn_exp2f((float)(tc - k)), (int)k);
  pw = epZero ? 0.0 : pwEst;
  const bool depZero = dep == 0.0;
  const double rho = dep * rateScale;
  const bool rhoOk = rho > 1e-300 && rho < 1e300;
  const double lrEst = 0.6931471805599453 * screen_log2(rhoOk ? rho : 1.0);
  lr = depZero ? 0.0 : lrEst;
  return (epZero ? p.hardN > 0.0 : epOk && tOk) && (depZero || rhoOk);
}
// Four times the first-order error bound of the estimate (above).
__device__ __forceinline__ double screen_margin(const JcParams &p, double soft, double bpw,
                                                double hard, double rfac) {
  constexpr double ln2 = 0.6931471805599453;
  constexpr double eLog = kScreenLog2Hw + 5.9604644775390625e-8 / ln2; // + 2⁻²⁴/ln 2
  const double ePw = ln2 * (fabs(p.hardN) * eLog + 2.98023223876953125e-8) + kScreenExp2Hw;
  return 4.0 * fabs(soft) *
         (fabs(bpw) * fabs(rfac) * ePw + fabs(hard) * fabs(p.rateC) * (ln2 * eLog));
}
// y and margin are set also where it returns false (finite then, and not an estimate).
__device__ __forceinline__ bool jc_yield_screen(const JcParams &p, double ep, double dep,
                                                double rateScale, double soft, double &y,
                                                double &margin) {
  double pw, lr;
  const bool ok = screen_terms(p, ep, dep, rateScale, pw, lr);
  const double bpw = p.b0 * pw;
  const double hard = p.a0 + bpw;
  const double rfac = 1.0 + p.rateC * lr;
  y = hard * soft * rfac;
  margin = screen_margin(p, soft, bpw, hard, rfac);
  return ok;
}
// The estimate with its slope dY/dΔ at Δ = dep > 0 (jc_yield_slope's expressions on the
// estimated terms), for the FP32 phase of the Newton solvers.
__device__ __forceinline__ bool jc_yield_slope_screen(const JcParams &p, double epOld,
                                                      double dep, double rateScale,
                                                      double soft, double &y, double &dy,
                                                      double &margin) {
  const double ep = epOld + dep;
  double pw, lr;
  if (!screen_terms(p, ep, dep, rateScale, pw, lr))
    return false;
  const double bpw = p.b0 * pw;
  const double hard = p.a0 + bpw;
  const double rfac = 1.0 + p.rateC * lr;
  y = hard * soft * rfac;
  const double dHard = ep > 0.0 ? p.b0 * p.hardN * pw / ep : 0.0;
  dy = soft * (dHard * rfac + hard * p.rateC / dep);
  margin = screen_margin(p, soft, bpw, hard, rfac);
  return true;
}

// Return-map solvers of the Johnson-Cook stage.
//   RM_BISECTION  the exact bracketed bisection of JohnsonCook.f90:114-136 (parity mode):
//                 tolerance E·1e-9, 50 iterations, iteration-1 seed = state 5.
//   RM_BISECTION_TIGHT  the same scheme run to |r| <= rtol·q_trial (200 iterations): the
//                 plain return map at the fast modes' accuracy.
//   RM_NEWTON     safeguarded Newton on r(Δ) = (q − 3μΔ) − max(Y(ε̄+Δ, Δ/dt, T), Y_old)
//                 inside the bracket [0, q/3μ], seeded with state 5 (the reference's seed),
//                 falling back to bisection whenever a step leaves the bracket; its first
//                 steps use the screen's FP32 estimate of Y and dY/dΔ (while the estimate
//                 decides the residual's sign), the rest the FP64 yield; stops at
//                 |r| <= rtol·q_trial on an FP64 residual. Iterations count both kinds.
//   RM_LEARNED    RM_NEWTON seeded with an external Δε̄ᵖ estimate (the learned warm start).
// Both bisections screen their iterations (jc_yield_screen): an iteration evaluates the FP64
// yield only when the estimate cannot decide, or when the loop ends there; iterates, the
// stored yield and every result are those of evaluating every iteration exactly, bit for bit.
//   RM_BISECTION_UNSCREENED  RM_BISECTION with the FP64 yield at every iteration (the literal
//                 loop, kept to check the screened one against).
enum ReturnMap { RM_BISECTION = 0, RM_BISECTION_TIGHT = 1, RM_NEWTON = 2, RM_LEARNED = 3 };
// The material stage's form for the linear elastic VUMAT (not a return-map solver).
constexpr int RM_ELASTIC = 4;
constexpr int RM_BISECTION_UNSCREENED = 5;

// The linear isotropic elastic VUMAT of materials/examples/Elastic.f: props E, ν;
// λ = Eν/((1+ν)(1−2ν)), μ = E/(2(1+ν)); the 6×6 stiffness with λ + 2μ / λ on the normal
// block and 2μ on the shear diagonal.
struct ElasticParams {
  double lam, mu;
};

inline ElasticParams make_elastic(double e, double nu) {
  return ElasticParams{e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), e / (2.0 * (1.0 + nu))};
}

// σ_new(i) = σ_old(i) + Σ_j C(i, j) Δε(j), summed over j = 1..6 in order, zero entries
// included (Elastic.f:41-47); state, temperature and energies are left unchanged.
__device__ __forceinline__ void elastic_point(const ElasticParams &p, const double inc[6],
                                              const double sOld[6], double sNew[6]) {
  const double diag = p.lam + 2.0 * p.mu, shear = 2.0 * p.mu;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double s = sOld[i];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const double cij = (i < 3 && j < 3) ? (i == j ? diag : p.lam) : (i == j ? shear : 0.0);
      s = s + cij * inc[j];
    }
    sNew[i] = s;
  }
}

struct ReturnMapStats {
  int iterations = 0;
  int fallbacks = 0; // Newton steps replaced by bisection
};

// Y and dY/dΔ at Δ > 0 (JohnsonCook.f90:160-170 differentiated in Δ); soft = 1 − θᵐ.
__device__ __forceinline__ void jc_yield_slope(const JcParams &p, double epOld, double dep,
                                               double dt, double soft, double &Y, double &dY) {
  const double ep = epOld + dep;
  const double epn = pow(ep, p.hardN);
  const double hard = p.a0 + p.b0 * epn;
  const double rate = dep / dt;
  const double er = rate == 0.0 ? p.rate0 : rate;
  const double rfac = 1.0 + p.rateC * log(er / p.rate0);
  Y = hard * soft * rfac;
  const double dHard = ep > 0.0 ? p.b0 * p.hardN * epn / ep : 0.0;
  dY = soft * (dHard * rfac + (dep > 0.0 ? hard * p.rateC / dep : 0.0));
}

// The first iteration's thermal factor 1 − θᵐ and yield stress (the max with the stored
// one), as jc_elastic evaluates them; jc_point takes them instead of evaluating them again.
struct JcFirst {
  double soft, yield;
};

// Whether jc_point<MODE> ends at its first iteration, Δε̄ᵖ = 0: the init branch, or a trial
// stress that leaves the yield stress at Δ = 0 as the solve's first iteration accepts it.
// The bisections (BISECT; TIGHT for RM_BISECTION_TIGHT) end there when |r| < tol or r < 0
// (JohnsonCook.f90:119-127), the Newton forms when r does not exceed rtol·q. Evaluates
// exactly the expressions jc_point's first iteration does and returns them in `first`.
template <bool BISECT = false, bool TIGHT = false>
__device__ __forceinline__ bool jc_elastic(const JcParams &p, const double inc[6],
                                           const double stressOld[6], const double state[6],
                                           double rtol, JcFirst *first = nullptr) {
  if (state[0] == 0.0)
    return true;
  double trial[6];
  const double traceInc = inc[0] + inc[1] + inc[2];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    trial[i] = stressOld[i] + p.lambda * traceInc + 2.0 * p.mu * inc[i];
#pragma unroll
  for (int i = 3; i < 6; ++i)
    trial[i] = stressOld[i] + 2.0 * p.mu * inc[i];
  const double pressure = (trial[0] + trial[1] + trial[2]) / 3.0;
  double dev[6];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    dev[i] = trial[i] - pressure;
#pragma unroll
  for (int i = 3; i < 6; ++i)
    dev[i] = trial[i];
  const double qTrial =
      sqrt(1.5 * (dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2] +
                  2.0 * (dev[3] * dev[3] + dev[4] * dev[4] + dev[5] * dev[5])));
  const double soft = jc_softening(p, state[2]);
  const double yield = fmax(jc_yield_soft(p, state[1], 0.0, soft), state[3]);
  if (first) {
    first->soft = soft;
    first->yield = yield;
  }
  const double r0 = fmax(0.0, qTrial) - yield;
  if (BISECT) // at Δ = 0 the bisection's residual is q − Y (qNew = max(0, q))
    return fabs(r0) < (TIGHT ? rtol * qTrial : p.tolerance) || r0 < 0.0;
  return !(r0 > 0.0 && fabs(r0) >= rtol * qTrial);
}

// The elastic trial stress of JohnsonCook.f90:75-81 (Voigt, engineering shear increments).
__device__ __forceinline__ void jc_trial_stress(const JcParams &p, const double inc[6],
                                                const double stressOld[6], double trial[6]) {
  const double traceInc = inc[0] + inc[1] + inc[2];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    trial[i] = stressOld[i] + p.lambda * traceInc + 2.0 * p.mu * inc[i];
#pragma unroll
  for (int i = 3; i < 6; ++i)
    trial[i] = stressOld[i] + 2.0 * p.mu * inc[i];
}
// Its pressure, deviator and von Mises stress q (JohnsonCook.f90:108-112).
__device__ __forceinline__ double jc_deviator(const double trial[6], double &pressure,
                                              double dev[6]) {
  pressure = (trial[0] + trial[1] + trial[2]) / 3.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
    dev[i] = trial[i] - pressure;
#pragma unroll
  for (int i = 3; i < 6; ++i)
    dev[i] = trial[i];
  return sqrt(1.5 * (dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2] +
                     2.0 * (dev[3] * dev[3] + dev[4] * dev[4] + dev[5] * dev[5])));
}
// The update after the solve (JohnsonCook.f90:138-154): radial return by Δε̄ᵖ = epInc, the
// work and plastic-work energies, heating and the state variables.
__device__ __forceinline__ void jc_update(const JcParams &p, const double trial[6],
                                          double pressure, const double dev[6], double qTrial,
                                          const double inc[6], const double stressOld[6],
                                          double state[6], double &eint, double &einel,
                                          double stressNew[6], double epOld,
                                          double temperature, double epInc, double yield,
                                          int iteration) {
  double upd[6];
#pragma unroll
  for (int i = 0; i < 6; ++i)
    upd[i] = trial[i];
  if (qTrial > 0.0) { // JohnsonCook.f90:138-142
    const double scale = fmax(0.0, 1.0 - 3.0 * p.mu * epInc / qTrial);
#pragma unroll
    for (int i = 0; i < 3; ++i)
      upd[i] = pressure + scale * dev[i];
#pragma unroll
    for (int i = 3; i < 6; ++i)
      upd[i] = scale * dev[i];
  }
  const double qNew = fmax(0.0, qTrial - 3.0 * p.mu * epInc);
  double work = 0.0;
#pragma unroll
  for (int i = 0; i < 6; ++i)
    work = work + (stressOld[i] + upd[i]) * inc[i];
  const double workInc = 0.5 * work;
  const double plasticWork = 0.5 * epInc * qNew;
#pragma unroll
  for (int i = 0; i < 6; ++i)
    stressNew[i] = upd[i];
  eint = eint + workInc / p.rho;
  einel = einel + plasticWork / p.rho;
  state[0] = 1.0;
  state[1] = epOld + epInc;
  state[2] = temperature + p.workToHeat * p.beta * plasticWork / (p.rho * p.cp);
  state[3] = yield;
  state[4] = epInc;
  state[5] = state[5] + static_cast<double>(iteration);
}

// One material point of JohnsonCook.f90:74-156. in: stressOld[6], strainInc[6], state[6]
// (in/out), energies (in/out); writes stressNew[6]; returns the iteration count (0 for the
// init branch). The bookkeeping after the solve is the reference's for every solver.
// `first`, when given, holds jc_elastic's evaluation of the first iteration's 1 − θᵐ and
// yield for the same state (the same bits jc_point would compute). FIRST_EXIT: the caller
// knows (jc_elastic) that the solve ends at its first iteration with Δε̄ᵖ = 0 and passes
// `first`; the solver loops are then left out and that iteration's outcome is taken
// (iteration 1, Δε̄ᵖ = 0, the stored yield first->yield), the bits the full routine gives.
// G > 1 (8, 16, 32 or 64; the screened bisections only): the G lanes of an aligned group call
// this for the same point with the same inputs, and its bisection runs log2(G) iterations per
// round (below); every lane of the group returns the same results, the single lane's bits.
template <int MODE, bool FIRST_EXIT = false, int G = 1>
__device__ __forceinline__ int jc_point(const JcParams &p, double dt, const double inc[6],
                                        const double stressOld[6], double state[6],
                                        double &eint, double &einel, double stressNew[6],
                                        double seed = 0.0, double rtol = 1.0e-12,
                                        ReturnMapStats *stats = nullptr,
                                        const JcFirst *first = nullptr) {
  double trial[6];
  jc_trial_stress(p, inc, stressOld, trial);

  if (state[0] == 0.0) { // JohnsonCook.f90:92-102
    state[0] = 1.0;
    state[1] = 0.0;
    state[2] = p.tref;
    state[3] = 0.0;
    state[4] = 0.0;
    state[5] = 1.0;
#pragma unroll
    for (int i = 0; i < 6; ++i)
      stressNew[i] = trial[i];
    return 0;
  }
  const double epOld = state[1], temperature = state[2], oldYield = state[3];
  double pressure, dev[6];
  const double qTrial = jc_deviator(trial, pressure, dev);

  double epInc = 0.0, yield = oldYield;
  int iteration = 0;
  if constexpr (FIRST_EXIT) {
    iteration = 1;
    yield = first->yield;
  } else if (MODE == RM_BISECTION || MODE == RM_BISECTION_TIGHT ||
             MODE == RM_BISECTION_UNSCREENED) {
    constexpr bool SCREEN = MODE != RM_BISECTION_UNSCREENED;
    const double tol = MODE == RM_BISECTION_TIGHT ? rtol * qTrial : p.tolerance;
    const int maxIt = MODE == RM_BISECTION_TIGHT ? 200 : 50;
    const double soft = first ? first->soft : jc_softening(p, temperature); // every iteration
    const double rateScale = SCREEN ? 1.0 / (dt * p.rate0) : 0.0;
    double epLower = 0.0, epUpper = qTrial / (2.0 * p.mu);
    // JohnsonCook.f90:119-136: iterations 1..maxIt; iteration k evaluates the yield at the
    // current Δ, ends the loop when |r| < tol (or at Δ = 0 with r < 0), else halves the
    // bracket (iteration 1 then moves to the state-5 seed). Here each round first runs a
    // lane's screened iterations until one needs the FP64 yield (or is the last allowed),
    // then evaluates that one exactly: the FP64 evaluations of a wavefront's lanes fall in
    // the same round instead of one per iteration. The iterates and the stored yield are
    // the literal loop's.
    bool done = false, exact = true; // iteration 1 is evaluated exactly
    iteration = 1;
    static_assert(G == 1 || ((G == 8 || G == 16 || G == 32 || G == 64) && SCREEN),
                  "lane groups of 8, 16, 32 or 64");
    if constexpr (G > 1) {
      // Lane groups: iteration 1 as below, then rounds of L = log2(G) iterations. The G − 1
      // iterates the next L iterations can visit form a binary tree (heap order: node 1 is the
      // round's first iterate, node 2m + 1 follows node m when its residual is ≥ 0, 2m when
      // < 0); group lane m evaluates node m, whose iterate it reaches by applying the loop's
      // own bracket updates along the path its index encodes, with the single loop's screen
      // or exact evaluation. The group then walks the decided path with those updates; each
      // node's outcome (end, or the residual's sign) is what the single loop finds there, so
      // the iterates, the iteration count and the stored yield are the same bits.
      constexpr int L = G == 8 ? 3 : G == 16 ? 4 : G == 32 ? 5 : 6;
      const int lane = __lane_id(), gl = lane & (G - 1), base = lane & ~(G - 1);
      {
        const double epRate = epInc / dt;
        if (first) {
          yield = first->yield;
        } else {
          yield = jc_yield_soft(p, epOld + epInc, epRate, soft);
          yield = fmax(yield, oldYield);
        }
        const double qNew = fmax(0.0, qTrial - 3.0 * p.mu * epInc);
        const double residual = qNew - yield;
        if (fabs(residual) < tol || (epInc == 0.0 && residual < 0.0)) {
          done = true;
        } else {
          if (residual >= 0.0)
            epLower = fmax(epLower, epInc);
          else
            epUpper = fmin(epUpper, epInc);
          epInc = fmin(fmax(state[4], epLower), epUpper); // iteration 1 moves to the seed
          iteration = iteration + 1;
        }
      }
      while (!done) {
        int depth = -1;
        double lo = epLower, hi = epUpper, e = epInc;
        if (gl > 0) {
          depth = 31 - __builtin_clz((unsigned)gl);
          for (int d = depth - 1; d >= 0; --d) {
            if ((gl >> d) & 1)
              lo = fmax(lo, e);
            else
              hi = fmin(hi, e);
            e = 0.5 * (lo + hi);
          }
        }
        const int itNode = iteration + depth;
        bool nodeEnds = false, nodeUp = false;
        double nodeYield = 0.0;
        if (gl > 0 && itNode <= maxIt) {
          const double qNew = fmax(0.0, qTrial - 3.0 * p.mu * e);
          double ya, margin;
          const bool est = jc_yield_screen(p, epOld + e, e, rateScale, soft, ya, margin);
          const double ra = qNew - fmax(ya, oldYield);
          const double slack = margin + 1e-14 * (qNew + fabs(ya));
          const bool decided = itNode != maxIt && est && fabs(ra) >= tol + slack &&
                               !(e == 0.0 && ra < 0.0);
          if (decided) {
            nodeUp = ra >= 0.0;
          } else {
            double yv = jc_yield_soft(p, epOld + e, e / dt, soft);
            yv = fmax(yv, oldYield);
            const double residual = qNew - yv;
            nodeYield = yv;
            nodeEnds = fabs(residual) < tol || (e == 0.0 && residual < 0.0);
            nodeUp = residual >= 0.0;
          }
        }
        const uint64_t ends = __ballot(nodeEnds) >> base;
        const uint64_t ups = __ballot(nodeUp) >> base;
        int node = 1, last = 0;
        for (int d = 0; d < L; ++d) {
          if ((ends >> node) & 1ull) {
            last = node;
            done = true;
            break;
          }
          const bool up = (ups >> node) & 1ull;
          if (up)
            epLower = fmax(epLower, epInc);
          else
            epUpper = fmin(epUpper, epInc);
          epInc = 0.5 * (epLower + epUpper);
          if (iteration == maxIt) { // evaluated exactly (it is the last allowed)
            last = node;
            done = true;
            break;
          }
          iteration = iteration + 1;
          node = 2 * node + (up ? 1 : 0);
        }
        const double y = __shfl(nodeYield, base + last);
        if (done)
          yield = y;
      }
    }
    while (G == 1 && !done) {
      if (SCREEN) {
        while (!exact) {
          const double qNew = fmax(0.0, qTrial - 3.0 * p.mu * epInc);
          double ya, margin;
          const bool est = jc_yield_screen(p, epOld + epInc, epInc, rateScale, soft, ya, margin);
          const double ra = qNew - fmax(ya, oldYield);
          const double slack = margin + 1e-14 * (qNew + fabs(ya));
          // Same sign as the exact residual, which is not below tol and ends nothing.
          const bool decided = est && fabs(ra) >= tol + slack && !(epInc == 0.0 && ra < 0.0);
          if (!decided) {
            exact = true;
            break;
          }
          if (ra >= 0.0)
            epLower = fmax(epLower, epInc);
          else
            epUpper = fmin(epUpper, epInc);
          epInc = 0.5 * (epLower + epUpper);
          iteration = iteration + 1;
          exact = iteration == maxIt;
        }
      }
      const double epRate = epInc / dt;
      if (first && iteration == 1) { // Δ = 0: ε̄ + 0 = ε̄ and the rate is ε̇₀, as in jc_elastic
        yield = first->yield;
      } else {
        yield = jc_yield_soft(p, epOld + epInc, epRate, soft);
        yield = fmax(yield, oldYield);
      }
      const double qNew = fmax(0.0, qTrial - 3.0 * p.mu * epInc);
      const double residual = qNew - yield;
      if (fabs(residual) < tol || (epInc == 0.0 && residual < 0.0)) {
        done = true;
      } else {
        if (residual >= 0.0)
          epLower = fmax(epLower, epInc);
        else
          epUpper = fmin(epUpper, epInc);
        epInc = 0.5 * (epLower + epUpper);
        if (iteration == 1)
          epInc = fmin(fmax(state[4], epLower), epUpper);
        if (iteration == maxIt) {
          done = true;
        } else {
          iteration = iteration + 1;
          exact = !SCREEN || iteration == maxIt;
        }
      }
    }
  } else {
    // Iteration 1 tests Δ = 0 exactly as the reference does (elastic exit).
    iteration = 1;
    const double soft = first ? first->soft : jc_softening(p, temperature);
    yield = first ? first->yield : fmax(jc_yield_soft(p, epOld, 0.0, soft), oldYield);
    const double r0 = fmax(0.0, qTrial) - yield;
    if (r0 > 0.0 && fabs(r0) >= rtol * qTrial) {
      double lo = 0.0, hi = qTrial / (3.0 * p.mu);
      double d = MODE == RM_LEARNED ? seed : state[4];
      if (!(d > lo && d < hi))
        d = 0.5 * (lo + hi);
      int fallbacks = 0;
      // FP32 phase: Newton steps on the screen's estimate of the yield and its slope while
      // the estimated residual lies farther from 0 than twice the estimate's error bound (its
      // sign is then the exact residual's, so the bracket stays valid); the FP64 Newton
      // below finishes from there, usually in two evaluations.
      const double rateScale = 1.0 / (dt * p.rate0);
      for (int it = 0; it < 16; ++it) {
        double Y, dY, margin;
        if (!jc_yield_slope_screen(p, epOld, d, rateScale, soft, Y, dY, margin))
          break;
        const bool capped = Y < oldYield;
        const double r = (qTrial - 3.0 * p.mu * d) - (capped ? oldYield : Y);
        if (fabs(r) <= 2.0 * margin + 1e-14 * (qTrial + fabs(Y)))
          break;
        ++iteration;
        if (r > 0.0)
          lo = d;
        else
          hi = d;
        const double slope = -3.0 * p.mu - (capped ? 0.0 : dY);
        double next = d - r / slope;
        if (!(next > lo && next < hi)) {
          next = 0.5 * (lo + hi);
          ++fallbacks;
        }
        if (next == d)
          break;
        d = next;
      }
      for (int it = 0; it < 60; ++it) {
        ++iteration;
        double Y, dY;
        jc_yield_slope(p, epOld, d, dt, soft, Y, dY);
        const bool capped = Y < oldYield;
        const double Yc = capped ? oldYield : Y;
        const double r = (qTrial - 3.0 * p.mu * d) - Yc;
        yield = Yc;
        epInc = d;
        if (fabs(r) <= rtol * qTrial)
          break;
        if (r > 0.0)
          lo = d;
        else
          hi = d;
        const double slope = -3.0 * p.mu - (capped ? 0.0 : dY);
        double next = d - r / slope;
        if (!(next > lo && next < hi)) {
          next = 0.5 * (lo + hi);
          ++fallbacks;
        }
        if (next == d)
          break;
        d = next;
      }
      if (stats)
        stats->fallbacks += fallbacks;
    }
  }
  if (stats)
    stats->iterations += iteration;
  jc_update(p, trial, pressure, dev, qTrial, inc, stressOld, state, eint, einel, stressNew,
            epOld, temperature, epInc, yield, iteration);
  return iteration;
}

} // namespace vpd
