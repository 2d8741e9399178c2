// pvglu_module.h -- single-diode PV module in explicit terminal form,
// shared by host (drop-in assembly) and device (batched stamping).
//
// I = Iph - I0 (exp((V + I Rs)/a) - 1) - (V + I Rs)/Rsh solved for I with the
// Lambert W function (Jain & Kapoor, Sol. Energy Mater. Sol. Cells 81 (2004)):
//   I = (Rsh (Iph + I0) - V)/(Rs + Rsh) - (a/Rs) W(theta),
//   ln theta = ln(Rs Rsh I0 / (a (Rs + Rsh))) + Rsh (Rs (Iph + I0) + V) / (a (Rs + Rsh)).
// W is evaluated from ln theta (no overflow for large V): for x = ln theta,
// w = W(e^x) solves w + ln w = x, refined by the Newton map
// w <- w (1 + x - ln w) / (1 + w)  (quadratic convergence).
// Derivatives: dI/dV = -1/(Rs+Rsh) - Rsh/(Rs (Rs+Rsh)) * W/(1+W),
//              dI/dIph = Rsh/(Rs+Rsh) / (1+W).
// The config modules: Iph = 9 G / 1000 A, I0 = 1e-10 A,
// a = n Vt = 1.5 V, Rs = 0.3 Ohm, Rsh log-uniform 200..600 Ohm.
#pragma once

#ifdef __HIPCC__
#define PVGLU_HD __host__ __device__ inline
#else
#define PVGLU_HD inline
#endif

#include <cmath>

namespace pvglu {

constexpr double kModuleI0 = 1e-10;
constexpr double kModuleA = 1.5;
constexpr double kModuleRs = 0.3;
constexpr double kIphPerIrradiance = 9.0 / 1000.0; // A per W/m^2
constexpr double kNortonG = 1e4;                   // S
constexpr double kNortonV = 600.0;                 // V

// W(e^x) for real x.
PVGLU_HD double lambertw_exp(double x) {
  if (x < -36.0) {
    // W(t) = t - t^2 + 1.5 t^3 ...; t < 2.4e-16 here
    const double t = exp(x);
    return t - t * t;
  }
  double w;
  if (x > 1.0) {
    const double lx = log(x);
    w = x - lx + lx / x;
  } else {
    const double t = exp(x); // t <= e
    w = t < 0.5 ? t * (1.0 - t) : 0.5;
  }
  for (int it = 0; it < 12; ++it) {
    const double wn = w * (1.0 + x - log(w)) / (1.0 + w);
    if (fabs(wn - w) <= 4e-16 * wn) {
      w = wn;
      break;
    }
    w = wn;
  }
  return w;
}

struct ModuleEval {
  double I, dIdV, dIdIph;
};

PVGLU_HD ModuleEval module_current(double V, double Iph, double Rsh) {
  const double Rs = kModuleRs, I0 = kModuleI0, a = kModuleA;
  const double RsRsh = Rs + Rsh;
  const double x = log(Rs * Rsh * I0 / (a * RsRsh)) +
                   Rsh * (Rs * (Iph + I0) + V) / (a * RsRsh);
  const double W = lambertw_exp(x);
  ModuleEval e;
  e.I = (Rsh * (Iph + I0) - V) / RsRsh - (a / Rs) * W;
  e.dIdV = -1.0 / RsRsh - Rsh / (Rs * RsRsh) * (W / (1.0 + W));
  e.dIdIph = (Rsh / RsRsh) / (1.0 + W);
  return e;
}

} // namespace pvglu
