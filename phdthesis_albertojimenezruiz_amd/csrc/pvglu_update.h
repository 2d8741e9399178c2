// pvglu_update.h -- the 011 damped element update shared by the plain
// Newton update (pvglu_newton_update) and the batched PV Newton.
#pragma once

#include <hip/hip_runtime.h>

#include "pvglu_launch.h"

namespace pvglu {

// 011 damping rule: the factor m halves from 1 until |cand - old| <= 1 V and
// neither value is NaN (is_value_change_justified,
// PV/011_calculate_new_iteration.cpp:6-17); each element's factor halves
// independently and the update fails once a factor is below 1e-13 (:44-74),
// leaving cand = old.  With m = 1 the candidate is exactly the new value.
__device__ __forceinline__ bool damped_step(double old, double nw,
                                            double &cand) {
  double m = 1.0;
  cand = old;
  for (int k = 0; k < 64; ++k) {
    if (m < 1e-13)
      break;
    cand = old * (1.0 - m) + nw * m;
    if (!isnan(old) && !isnan(cand) && !(fabs(old - cand) > 1.0))
      return true;
    m = m / 2.0;
  }
  cand = old;
  return false;
}

// The per-instance sum of |change| (the reference's `diff`, accumulated in
// long double, PV/011_calculate_new_iteration.cpp:39,76-85) is accumulated in
// fixed point: every |change| is at most 1 V (the damping rule), is rounded
// down to a multiple of 2^-44 V and added as a 64-bit integer.  Integer sums
// do not depend on their order, so the result is the same whatever the block
// partition (which depends on the batch size) or the order in which blocks
// finish; 2^64 * 2^-44 V = 1.0e6 V bounds the sum of 1e6 elements.
constexpr double kDeltaFixedScale = 17592186044416.0; // 2^44
__device__ __forceinline__ unsigned long long delta_to_fixed(double d) {
  return (unsigned long long)(d * kDeltaFixedScale);
}
__host__ __device__ inline double fixed_to_delta(unsigned long long s) {
  return (double)s * (1.0 / kDeltaFixedScale);
}

// Element-update launch shape (k_newton_update and the PV harness's fused
// update + module evaluation): a 256-thread block holds TW instance lanes
// (TW = the power of two >= B, at most 256) by RP = 256/TW element rows and
// covers RP * kUpdPerThread consecutive elements, row r taking e0 + r,
// e0 + r + RP, ...; the loads of kUpdBatch elements are issued together.
constexpr int kUpdPerThread = 8;
constexpr int kUpdBatch = 2;
__host__ __device__ inline int upd_tile_log2(int B) {
  return B >= 256 ? 8 : (B <= 1 ? 0 : 32 - __builtin_clz((unsigned)(B - 1)));
}

// Terminal tag of the correction form's element lists: bit 30 of pos[e] /
// neg[e] marks the one element that writes that node's next voltage (node
// indices stay below 2^30).
constexpr int kOwnTag = 1 << 30;

// Nothing loaded ahead for an element (the plain update's pre hook).
struct NoPre {};

// One block of the element update: v[e] <- damped_step(v[e], V_e) for the
// block's elements of the active instances, with V_e = x[pos[e]] - x[neg[e]]
// (voltage_drop_dc, PV/002_base_electric_system.cpp:21-33) or, in the
// correction form (CORR: x + dx being the new node voltages),
// V_e = (x[pos] + dx[pos]) - (x[neg] + dx[neg]); then post(e, instance,
// new value, V_e, pre(e, instance)), pre being loaded with the element's
// other operands.  Per instance the block's fixed-point sums and maxima of
// |change| are combined in LDS and added to delta_sum / delta_max with one
// atomic each; both results are independent of the order (integer sum; the
// maximum of non-negative doubles compared as integers).
// Loads: a batch of kUpdBatch elements issues all its loads in two rounds,
// first the terminals with everything that does not depend on them (old
// value, pre), then the node voltages at the terminals; a grounded terminal
// (index < 0) reads node 0 and takes 0 V by a select, so no branch sits
// between the loads (a branch there makes the compiler wait for every load
// in flight at its join).
// With xo (correction form), the node sums x + dx at the terminals tagged
// kOwnTag are also stored to xo: the next iteration's node voltages, the
// same additions bit for bit.  An instance frozen by the last control step
// (justf set) copies x to xo at those terminals instead, so that both node
// buffers hold its final voltages from then on.
template <bool CORR, typename Pre, typename Post>
__device__ __forceinline__ void
update_elements(int ne, int B, const int *__restrict__ pos,
                const int *__restrict__ neg, const double *__restrict__ x,
                const double *__restrict__ dx, double *__restrict__ v,
                unsigned long long *delta_sum, unsigned long long *delta_max,
                const int *active, int *status, Pre pre, Post post,
                double *__restrict__ xo, const int *__restrict__ justf) {
  __shared__ unsigned long long s_sum[256];
  __shared__ double s_max[256];
  const int lt = upd_tile_log2(B), rp = 256 >> lt;
  const int lane = (int)(threadIdx.x & ((1u << lt) - 1u));
  const int r = (int)(threadIdx.x >> lt);
  const int in = (int)(blockIdx.y << lt) + lane;
  const bool live = in < B && !(active && !active[in]);
  const int e0 = (int)blockIdx.x * rp * kUpdPerThread;
  const int e1 = min(ne, e0 + rp * kUpdPerThread);
  unsigned long long sum = 0;
  double mx = 0.0;
  auto at = [&](int k) { return (size_t)(k & (kOwnTag - 1)) * B + in; };
  if (live) {
    for (int eb = e0 + r; eb < e1; eb += kUpdBatch * rp) {
      int kp[kUpdBatch], kn[kUpdBatch];
      double old[kUpdBatch];
      decltype(pre(0, 0)) pd[kUpdBatch];
#pragma unroll
      for (int j = 0; j < kUpdBatch; ++j) {
        const int e = min(eb + j * rp, e1 - 1); // in bounds; unused if past e1
        kp[j] = pos[e];
        kn[j] = neg[e];
        old[j] = v[(size_t)e * B + in];
        pd[j] = pre(e, in);
      }
      double xp[kUpdBatch], xn[kUpdBatch];
#pragma unroll
      for (int j = 0; j < kUpdBatch; ++j) {
        const size_t ip = at(max(kp[j], 0)), iq = at(max(kn[j], 0));
        double a = x[ip], b = x[iq];
        if (CORR) {
          a = a + dx[ip];
          b = b + dx[iq];
        }
        xp[j] = kp[j] < 0 ? 0.0 : a;
        xn[j] = kn[j] < 0 ? 0.0 : b;
      }
#pragma unroll
      for (int j = 0; j < kUpdBatch; ++j) {
        const int e = eb + j * rp;
        if (e >= e1)
          break;
        if (CORR && xo) { // read next after the solve has streamed on
          if (kp[j] >= kOwnTag)
            __builtin_nontemporal_store(xp[j], xo + at(kp[j]));
          if (kn[j] >= kOwnTag)
            __builtin_nontemporal_store(xn[j], xo + at(kn[j]));
        }
        const double nw = xp[j] - xn[j];
        double cand;
        if (!damped_step(old[j], nw, cand))
          record_status(status, PVGLU_E_NOCONV, e, in);
        const double d = fabs(cand - old[j]);
        sum += delta_to_fixed(d <= 1.0 ? d : 1.0); // NaN: status recorded
        mx = fmax(mx, d);
        v[(size_t)e * B + in] = cand;
        post(e, in, cand, nw, pd[j]);
      }
    }
  } else if (CORR && xo && in < B && justf[in]) {
    for (int e = e0 + r; e < e1; e += rp) {
      const int a = pos[e], b = neg[e];
      if (a >= kOwnTag)
        xo[at(a)] = x[at(a)];
      if (b >= kOwnTag)
        xo[at(b)] = x[at(b)];
    }
  }
  s_sum[threadIdx.x] = sum;
  s_max[threadIdx.x] = mx;
  __syncthreads();
  if (r == 0 && live) {
    unsigned long long ts = 0;
    double tm = 0.0;
    for (int k = 0; k < rp; ++k) {
      ts += s_sum[(k << lt) + lane];
      tm = fmax(tm, s_max[(k << lt) + lane]);
    }
    atomicAdd(&delta_sum[in], ts);
    atomicMax(&delta_max[in], (unsigned long long)__double_as_longlong(tm));
  }
}

} // namespace pvglu
