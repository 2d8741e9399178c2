// pvglu_module.h -- single-diode PV module in explicit terminal form,
// shared by host (drop-in assembly) and device (batched stamping).
//
// I = Iph - I0 (exp((V + I Rs)/a) - 1) - (V + I Rs)/Rsh solved for I with the
// Lambert W function (Jain & Kapoor, Sol. Energy Mater. Sol. Cells 81 (2004)):
//   I = (Rsh (Iph + I0) - V)/(Rs + Rsh) - (a/Rs) W(theta),
//   ln theta = ln(Rs Rsh I0 / (a (Rs + Rsh))) + Rsh (Rs (Iph + I0) + V) / (a (Rs + Rsh)).
// W is evaluated from ln theta (no overflow for large V): for x = ln theta,
// w = W(e^x) solves w + ln w = x (see lambertw_exp).
// Derivatives: dI/dV = -1/(Rs+Rsh) - Rsh/(Rs (Rs+Rsh)) * W/(1+W),
//              dI/dIph = Rsh/(Rs+Rsh) / (1+W).
// The plant modules of the workloads: Iph = 9 G / 1000 A, I0 = 1e-10 A,
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

// One step of the fourth-order iteration of Fritsch, Shafer and Crowley
// (Comm. ACM 16 (1973) 123) for w + ln w = x:
//   z = x - w - ln w,  q = 2 (1+w) (1 + w + 2z/3),
//   w <- w (1 + z (q - z) / ((1+w) (q - 2z))).
template <class T> PVGLU_HD T fsc_step(T x, T w, T lw) {
  const T z = x - w - lw;
  const T w1 = T(1) + w;
  const T q = T(2) * w1 * (w1 + T(2.0 / 3.0) * z);
  return w * (T(1) + z * (q - z) / (w1 * (q - T(2) * z)));
}

// W(e^x) for real x, a fixed number of steps with no data-dependent branch
// for x >= -36 (a wavefront's lanes stay together).  The start is found in
// single precision: x > 1, the asymptotic w = x - ln x + ln x / x (within
// 8 % on (1, inf)); x <= 1, w = t/(1+t), t = e^x (within 27 %); both are
// computed and one selected, then two single-precision steps bring it
// within 2e-6 (the cancellation in z limits single precision there).  One
// double-precision step then leaves a relative error of order (2e-6)^4 plus
// its own rounding: at most 7e-15 on [-36, 3000] against a 40-digit Lambert
// W, the same bound as three double-precision steps from the start, with
// one double-precision logarithm and one division instead of four and five.
PVGLU_HD double lambertw_exp(double x) {
  if (x < -36.0) {
    // W(t) = t - t^2 + 1.5 t^3 ...; t < 2.4e-16 here
    const double t = exp(x);
    return t - t * t;
  }
  const float xf = (float)x;
  const float lx = logf(fmaxf(xf, 1.0f));
  const float wa = xf - lx + lx / xf;
  const float t = expf(fminf(xf, 1.0f));
  const float wb = t / (1.0f + t);
  float wf = xf > 1.0f ? wa : wb;
  wf = fsc_step<float>(xf, wf, logf(wf));
  wf = fsc_step<float>(xf, wf, logf(wf));
  const double w = (double)wf;
  return fsc_step<double>(x, w, log(w));
}

#ifdef __HIPCC__
// Launch shape of the [row][instance] kernels: a 256-thread block holds TW
// instance lanes (TW = the power of two >= B, at most 256) by 256/TW rows,
// so a wavefront's lanes all have work at any batch size (consecutive lanes:
// consecutive instances, then the next row) and a launch has about
// rows * B / 256 blocks.  (One block per row leaves most lanes idle when B is
// small, and hundreds of thousands of tiny blocks make the launch
// dispatch-bound.)
struct RowGrid {
  dim3 grid, block;
};
PVGLU_HD int row_tile_log2(int B) {
  return B >= 256 ? 8 : (B <= 1 ? 0 : 32 - __builtin_clz((unsigned)(B - 1)));
}
inline RowGrid row_grid(long rows, int B) {
  const int lt = row_tile_log2(B), rp = 256 >> lt;
  return {dim3((unsigned)((rows + rp - 1) / rp),
               (unsigned)((B + (1 << lt) - 1) >> lt)),
          dim3(256)};
}
#define PVGLU_ROW_INDEX(ROWS, B)                                               \
  const int lt_ = pvglu::row_tile_log2(B);                                     \
  const int row_ = (int)((blockIdx.x << (8 - lt_)) + (threadIdx.x >> lt_));    \
  const int inst =                                                             \
      (int)((blockIdx.y << lt_) + (threadIdx.x & ((1u << lt_) - 1u)));         \
  if (inst >= (B) || row_ >= (ROWS))                                           \
    return;                                                                    \
  const size_t gid = (size_t)row_ * (size_t)(B) + (size_t)inst
#endif

struct ModuleEval {
  double I, dIdV, dIdIph;
};

// The per-module constant of ln theta: ln(Rs Rsh I0 / (a (Rs + Rsh))).
PVGLU_HD double module_log_k(double Rsh) {
  const double Rs = kModuleRs, I0 = kModuleI0, a = kModuleA;
  return log(Rs * Rsh * I0 / (a * (Rs + Rsh)));
}

// module_current with the per-module constant lk = module_log_k(Rsh) given
// (computed once per module by the batched evaluators; the same operations).
PVGLU_HD ModuleEval module_current_lk(double V, double Iph, double Rsh,
                                      double lk) {
  const double Rs = kModuleRs, I0 = kModuleI0, a = kModuleA;
  const double RsRsh = Rs + Rsh;
  const double x = lk + Rsh * (Rs * (Iph + I0) + V) / (a * RsRsh);
  const double W = lambertw_exp(x);
  ModuleEval e;
  e.I = (Rsh * (Iph + I0) - V) / RsRsh - (a / Rs) * W;
  e.dIdV = -1.0 / RsRsh - Rsh / (Rs * RsRsh) * (W / (1.0 + W));
  e.dIdIph = (Rsh / RsRsh) / (1.0 + W);
  return e;
}

PVGLU_HD ModuleEval module_current(double V, double Iph, double Rsh) {
  return module_current_lk(V, Iph, Rsh, module_log_k(Rsh));
}

} // namespace pvglu
