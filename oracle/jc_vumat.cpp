// ORACLE — test infrastructure only. Nothing in the product links, loads or calls this
// file; tests, the smoke check and the CPU-baseline leg of the benchmark use it as the
// checker. CPU restatement of the Johnson-Cook VUMAT example and of its C bridge.
//
// Follows materials/examples/JohnsonCook.f90:12-172 (vumat + jc_yield) and
// materials/bridge/vumat_c_wrapper_jinfo.f90:1-44 (vumat_adapter_call). Arrays are
// Fortran column-major: component c of point k lives at [k + nblock*c]. Every
// arithmetic expression keeps the Fortran evaluation order (left to right, powers
// first), so that with -ffp-contract=off the results agree bit for bit with a
// non-contracting build of the Fortran source (pow/log/sqrt come from libm in both).
//
// Parity pins: tests/test_oracle_material.py checks it against the amdflang build of the
// Fortran source under oracle/_ref (bitwise), the elastic-trial closed form, the
// init branch, the converged-residual bound |r| < E*1e-9, the epsilon=0/rate=0
// special case Y = A(1-theta^m), and the heat-unit ratio of
// tests/jc_thermal_units_test.cpp:26-34.
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace {

// JohnsonCook.f90:160-170 (jc_yield)
double jcYield(double a, double b, double c, double n, double m, double refRate, double refTemp,
               double meltTemp, double ep, double rate, double temp) {
  double theta = (temp - refTemp) / (meltTemp - refTemp);
  theta = std::fmin(1.0, std::fmax(0.0, theta));
  double effectiveRate = rate;
  if (effectiveRate == 0.0)
    effectiveRate = refRate;
  return (a + b * std::pow(ep, n)) * (1.0 - std::pow(theta, m)) *
         (1.0 + c * std::log(effectiveRate / refRate));
}

} // namespace

extern "C" {

// JohnsonCook.f90:12-156 with the bisection's stopping rule as a parameter: rtol <= 0 keeps
// the reference's |r| < E·1e-9 and 50 iterations; rtol > 0 stops at |r| < rtol·q_trial or
// maxIt iterations (the tightly converged solution the fast return maps are checked on). Returns 0, or 1
// for the dimension errors on which the Fortran routine executes STOP (JohnsonCook.f90:52-53).
int orc_vumat_jc_tol(int nblock, int ndir, int nshr, int nstatev, int nfieldv, int nprops,
                     const double *dtArray, const double *props, const double *strainInc,
                     const double *tempOld, const double *stretchOld, const double *defgradOld,
                     const double *fieldOld, const double *stressOld, const double *stateOld,
                     const double *enerInternOld, const double *enerInelasOld, double *tempNew,
                     double *stretchNew, double *defgradNew, double *fieldNew, double *stressNew,
                     double *stateNew, double *enerInternNew, double *enerInelasNew, double rtol,
                     int maxIt) {
  if (ndir != 3 || nshr != 3)
    return 1;
  if (nprops < 14 || nstatev < 6)
    return 1;
  const long nb = nblock;
  auto ix = [nb](long k, long c) { return k + nb * c; };

  const double young = props[0], poisson = props[1], a0 = props[2], b0 = props[3];
  const double rateC = props[4], hardN = props[5], thermalM = props[6], rate0 = props[7];
  const double tref = props[8], tmelt = props[9], rho = props[10], beta = props[11];
  const double cp = props[12], workToHeat = props[13];
  const double dt = dtArray[0];
  const double mu = young / (2.0 * (1.0 + poisson));
  const double lambda = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson));
  const double tolerance = young * 1.0e-9;

  for (long k = 0; k < nb; ++k) {
    double trial[6], deviator[6], updated[6], oldStress[6], inc[6];
    for (int i = 0; i < 6; ++i)
      inc[i] = strainInc[ix(k, i)];
    const double traceInc = inc[0] + inc[1] + inc[2];
    for (int i = 0; i < 3; ++i)
      trial[i] = stressOld[ix(k, i)] + lambda * traceInc + 2.0 * mu * inc[i];
    for (int i = 3; i < 6; ++i)
      trial[i] = stressOld[ix(k, i)] + 2.0 * mu * inc[i];

    for (int i = 0; i < 6; ++i)
      updated[i] = trial[i];
    for (int c = 0; c < nstatev; ++c)
      stateNew[ix(k, c)] = stateOld[ix(k, c)];
    tempNew[k] = tempOld[k];
    enerInternNew[k] = enerInternOld[k];
    enerInelasNew[k] = enerInelasOld[k];
    for (int c = 0; c < 6; ++c)
      stretchNew[ix(k, c)] = stretchOld[ix(k, c)];
    for (int c = 0; c < 9; ++c)
      defgradNew[ix(k, c)] = defgradOld[ix(k, c)];
    for (int c = 0; c < nfieldv; ++c)
      fieldNew[ix(k, c)] = fieldOld[ix(k, c)];

    // Init branch, JohnsonCook.f90:92-102: the first call is elastic only.
    if (stateOld[ix(k, 0)] == 0.0) {
      stateNew[ix(k, 0)] = 1.0;
      stateNew[ix(k, 1)] = 0.0;
      stateNew[ix(k, 2)] = tref;
      stateNew[ix(k, 3)] = 0.0;
      stateNew[ix(k, 4)] = 0.0;
      stateNew[ix(k, 5)] = 1.0;
      tempNew[k] = tref;
      for (int i = 0; i < 6; ++i)
        stressNew[ix(k, i)] = updated[i];
      continue;
    }

    for (int i = 0; i < 6; ++i)
      oldStress[i] = stressOld[ix(k, i)];
    const double epOld = stateOld[ix(k, 1)];
    const double temperature = stateOld[ix(k, 2)];
    const double oldYield = stateOld[ix(k, 3)];
    const double pressure = (trial[0] + trial[1] + trial[2]) / 3.0;
    for (int i = 0; i < 3; ++i)
      deviator[i] = trial[i] - pressure;
    for (int i = 3; i < 6; ++i)
      deviator[i] = trial[i];
    const double qTrial =
        std::sqrt(1.5 * (deviator[0] * deviator[0] + deviator[1] * deviator[1] +
                         deviator[2] * deviator[2] +
                         2.0 * (deviator[3] * deviator[3] + deviator[4] * deviator[4] +
                                deviator[5] * deviator[5])));

    // Bracketed bisection, JohnsonCook.f90:114-136.
    double epInc = 0.0, epLower = 0.0, epUpper = qTrial / (2.0 * mu);
    double yield = oldYield, qNew, residual;
    const double tol = rtol > 0.0 ? rtol * qTrial : tolerance;
    const int iterCap = rtol > 0.0 ? maxIt : 50;
    int iteration = 0;
    while (iteration < iterCap) {
      iteration = iteration + 1;
      const double epRate = epInc / dt;
      yield = jcYield(a0, b0, rateC, hardN, thermalM, rate0, tref, tmelt, epOld + epInc, epRate,
                      temperature);
      yield = std::fmax(yield, oldYield);
      qNew = std::fmax(0.0, qTrial - 3.0 * mu * epInc);
      residual = qNew - yield;
      if (std::fabs(residual) < tol)
        break;
      if (epInc == 0.0 && residual < 0.0)
        break;
      if (residual >= 0.0)
        epLower = std::fmax(epLower, epInc);
      else
        epUpper = std::fmin(epUpper, epInc);
      epInc = 0.5 * (epLower + epUpper);
      if (iteration == 1)
        epInc = std::fmin(std::fmax(stateOld[ix(k, 4)], epLower), epUpper);
    }

    // Radial scale and bookkeeping, JohnsonCook.f90:138-155.
    if (qTrial > 0.0) {
      const double scale = std::fmax(0.0, 1.0 - 3.0 * mu * epInc / qTrial);
      for (int i = 0; i < 3; ++i)
        updated[i] = pressure + scale * deviator[i];
      for (int i = 3; i < 6; ++i)
        updated[i] = scale * deviator[i];
    }
    qNew = std::fmax(0.0, qTrial - 3.0 * mu * epInc);
    double workSum = 0.0;
    for (int i = 0; i < 6; ++i)
      workSum = workSum + (oldStress[i] + updated[i]) * inc[i];
    const double workInc = 0.5 * workSum;
    const double plasticWork = 0.5 * epInc * qNew;
    for (int i = 0; i < 6; ++i)
      stressNew[ix(k, i)] = updated[i];
    enerInternNew[k] = enerInternOld[k] + workInc / rho;
    enerInelasNew[k] = enerInelasOld[k] + plasticWork / rho;
    stateNew[ix(k, 0)] = 1.0;
    stateNew[ix(k, 1)] = epOld + epInc;
    stateNew[ix(k, 2)] = temperature + workToHeat * beta * plasticWork / (rho * cp);
    stateNew[ix(k, 3)] = yield;
    stateNew[ix(k, 4)] = epInc;
    stateNew[ix(k, 5)] = stateOld[ix(k, 5)] + static_cast<double>(iteration);
    tempNew[k] = stateNew[ix(k, 2)];
  }
  return 0;
}

int orc_vumat_jc(int nblock, int ndir, int nshr, int nstatev, int nfieldv, int nprops,
                 const double *dtArray, const double *props, const double *strainInc,
                 const double *tempOld, const double *stretchOld, const double *defgradOld,
                 const double *fieldOld, const double *stressOld, const double *stateOld,
                 const double *enerInternOld, const double *enerInelasOld, double *tempNew,
                 double *stretchNew, double *defgradNew, double *fieldNew, double *stressNew,
                 double *stateNew, double *enerInternNew, double *enerInelasNew) {
  return orc_vumat_jc_tol(nblock, ndir, nshr, nstatev, nfieldv, nprops, dtArray, props,
                          strainInc, tempOld, stretchOld, defgradOld, fieldOld, stressOld,
                          stateOld, enerInternOld, enerInelasOld, tempNew, stretchNew,
                          defgradNew, fieldNew, stressNew, stateNew, enerInternNew,
                          enerInelasNew, 0.0, 50);
}

#ifdef VPD_ORACLE_PLUGIN
// materials/bridge/vumat_c_wrapper_jinfo.f90:1-44 — the 33-argument C entry. The bridge's
// jInfoArray, cmname, coordMp, charLength, density, relSpinInc and time arguments are not
// read by the JC routine; they are accepted and ignored. A STOP becomes exit(1), as the
// Fortran runtime does.
void vumat_adapter_call(int *nblock, int *ndir, int *nshr, int *nstatev, int *nfieldv,
                        int *nprops, int *lanneal, double *stepTime, double *totalTime,
                        double *dt, char *cmname, double *coordMp, double *charLength,
                        double *props, double *density, double *strainInc, double *relSpinInc,
                        double *tempOld, double *stretchOld, double *defgradOld,
                        double *fieldOld, double *stressOld, double *stateOld,
                        double *enerInternOld, double *enerInelasOld, double *tempNew,
                        double *stretchNew, double *defgradNew, double *fieldNew,
                        double *stressNew, double *stateNew, double *enerInternNew,
                        double *enerInelasNew) {
  (void)lanneal;
  (void)stepTime;
  (void)totalTime;
  (void)cmname;
  (void)coordMp;
  (void)charLength;
  (void)density;
  (void)relSpinInc;
  if (orc_vumat_jc(*nblock, *ndir, *nshr, *nstatev, *nfieldv, *nprops, dt, props, strainInc,
                   tempOld, stretchOld, defgradOld, fieldOld, stressOld, stateOld, enerInternOld,
                   enerInelasOld, tempNew, stretchNew, defgradNew, fieldNew, stressNew, stateNew,
                   enerInternNew, enerInelasNew) != 0) {
    std::fprintf(stderr, "JohnsonCook VUMAT requires ndir=3,nshr=3, 14 props and 6 states\n");
    std::exit(1);
  }
}
#endif

} // extern "C"
