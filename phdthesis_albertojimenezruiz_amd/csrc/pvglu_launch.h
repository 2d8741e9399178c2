// pvglu_launch.h -- kernel argument views of a device plan and the
// per-width launchers of the numeric kernels.
//
// The refactor / substitution kernels are templates over the instance width
// W (lanes of one chain on consecutive instances).  Each width is compiled in
// its own translation unit (pvglu_k<W>.hip instantiates Launch<W>), so the
// widths build in parallel; host code calls Launch<W> through these
// declarations only.
#pragma once

#include <hip/hip_runtime.h>

#include "pvglu.h"

namespace pvglu {

struct Pat {
  const int *__restrict__ cp;
  const int *__restrict__ ri;
  const int *__restrict__ op_ptr;
  const int *__restrict__ op_l;
  const int *__restrict__ op_u;
  const int *__restrict__ diag;
  const int *__restrict__ lrow_ptr;
  const int *__restrict__ lrow_slot;
  const int *__restrict__ lrow_col;
  const int *__restrict__ urow_ptr;
  const int *__restrict__ urow_slot;
  const int *__restrict__ urow_col;
  const int *__restrict__ lu_rec;
  const int *__restrict__ lu_seg;
  const int *__restrict__ lu_run;
  const int *__restrict__ lu_fast;
  const int *__restrict__ u_rec;
  const int *__restrict__ u_seg;
  const int *__restrict__ u_run;
  const int *__restrict__ u_fast;
  const int *__restrict__ l_rec;
  const int *__restrict__ l_seg;
  const int *__restrict__ l_run;
  const int *__restrict__ l_fast;
  const int *__restrict__ ut_rec;
  const int *__restrict__ ut_seg;
  const int *__restrict__ ut_run;
  const int *__restrict__ ut_fast;
  const int *__restrict__ lt_rec;
  const int *__restrict__ lt_seg;
  const int *__restrict__ lt_run;
  const int *__restrict__ lt_fast;
  const int *__restrict__ drf;
  int *status;
};

struct Lev {
  const int *__restrict__ level_chain_ptr;
  const int *__restrict__ chain_col_ptr;
  const int *__restrict__ cols;
  const int *__restrict__ comp_ptr;
  const int *__restrict__ comp_lv_begin;
  const int *__restrict__ comp_lv_end;
  const int *__restrict__ comp_lv_level;
};

enum Op { OP_REFACTOR = 0, OP_REFACTOR_FUSED, OP_LSOLVE, OP_USOLVE, OP_UT, OP_LT };

// HYBRID schedule: a run of narrow levels (few chains) in one launch.  Block
// -> (component, tile); the four level sets are walked in order
// (refactor, forward, backward, transposed-L) over the given level ranges,
// each with its operation, barriers between levels.
struct LvRange {
  int lo[4], hi[4], op[4];
};

// Multi-pattern work list: groups of instances with their own plans in one
// launch; block -> (group, component, instance tile).
constexpr int kMaxGroups = 6;
struct MultiArgs {
  int ngroups;
  int block_start[kMaxGroups + 1];
  int tiles[kMaxGroups];
  int B[kMaxGroups];
  int fuse[kMaxGroups];
  Pat pat[kMaxGroups];
  Lev lu[kMaxGroups], l[kMaxGroups], u[kMaxGroups], lt[kMaxGroups];
  const double *a[kMaxGroups];
  double *luv[kMaxGroups];
  const double *b[kMaxGroups];
  double *x[kMaxGroups];
};

__device__ __forceinline__ void record_status(int *status, int code, int col,
                                              int inst) {
  if (atomicCAS(status, 0, code) == 0) {
    status[1] = col;
    status[2] = inst;
  }
}

// 011 damping rule: the factor m halves from 1 until |cand - old| <= 1 V and
// neither value is NaN (is_value_change_justified,
// PV/011_calculate_new_iteration.cpp:6-17); each element's factor halves
// independently and the update fails once a factor is below 1e-13 (:44-74),
// leaving cand = old.
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

// Element-update launch shape (k_newton_update and the PV harness's fused
// update + module evaluation): a 256-thread block holds TW instance lanes
// (TW = the power of two >= B, at most 256) by RP = 256/TW element rows and
// covers RP * kUpdPerThread consecutive elements, row r taking e0 + r,
// e0 + r + RP, ...; the loads of kUpdBatch elements are issued together.
constexpr int kUpdPerThread = 16;
constexpr int kUpdBatch = 4;
__host__ __device__ inline int upd_tile_log2(int B) {
  return B >= 256 ? 8 : (B <= 1 ? 0 : 32 - __builtin_clz((unsigned)(B - 1)));
}

// One block of the element update: v[e] <- damped_step(v[e], x[pos[e]] -
// x[neg[e]]) for the block's elements of the active instances, then
// post(e, instance, new value); per instance the block's sums and maxima of
// |change| are combined in LDS in a fixed order and added to delta_sum /
// delta_max with one atomic each (the sum is order-dependent in the last
// bits across blocks; the maximum is exact).
template <typename Post>
__device__ __forceinline__ void
update_elements(int ne, int B, const int *__restrict__ pos,
                const int *__restrict__ neg, const double *__restrict__ x,
                double *__restrict__ v, double *delta_sum,
                unsigned long long *delta_max, const int *active, int *status,
                Post post) {
  __shared__ double s_sum[256], s_max[256];
  const int lt = upd_tile_log2(B), rp = 256 >> lt;
  const int lane = (int)(threadIdx.x & ((1u << lt) - 1u));
  const int r = (int)(threadIdx.x >> lt);
  const int in = (int)(blockIdx.y << lt) + lane;
  const bool live = in < B && !(active && !active[in]);
  const int e0 = (int)blockIdx.x * rp * kUpdPerThread;
  const int e1 = min(ne, e0 + rp * kUpdPerThread);
  double sum = 0.0, mx = 0.0;
  if (live) {
    for (int eb = e0 + r; eb < e1; eb += kUpdBatch * rp) {
      double nw[kUpdBatch], old[kUpdBatch];
#pragma unroll
      for (int j = 0; j < kUpdBatch; ++j) {
        const int e = min(eb + j * rp, e1 - 1); // in bounds; unused if past e1
        const int p = pos[e], m_ = neg[e];
        const double vp = p >= 0 ? x[(size_t)p * B + in] : 0.0;
        const double vn = m_ >= 0 ? x[(size_t)m_ * B + in] : 0.0;
        nw[j] = vp - vn; // voltage_drop_dc (PV/002_base_electric_system.cpp:21-33)
        old[j] = v[(size_t)e * B + in];
      }
#pragma unroll
      for (int j = 0; j < kUpdBatch; ++j) {
        const int e = eb + j * rp;
        if (e >= e1)
          break;
        double cand;
        if (!damped_step(old[j], nw[j], cand))
          record_status(status, PVGLU_E_NOCONV, e, in);
        const double d = fabs(cand - old[j]);
        sum += d;
        mx = fmax(mx, d);
        v[(size_t)e * B + in] = cand;
        post(e, in, cand);
      }
    }
  }
  s_sum[threadIdx.x] = sum;
  s_max[threadIdx.x] = mx;
  __syncthreads();
  if (r == 0 && live) {
    double ts = 0.0, tm = 0.0;
    for (int k = 0; k < rp; ++k) {
      ts += s_sum[(k << lt) + lane];
      tm = fmax(tm, s_max[(k << lt) + lane]);
    }
    atomicAdd(&delta_sum[in], ts);
    atomicMax(&delta_max[in], (unsigned long long)__double_as_longlong(tm));
  }
}

// Multi-pattern LEVEL / HYBRID steps: one launch covers the same step of
// every group (plans of one topology at different sizes, so their step
// sequences agree); blocks [block_start[g], block_start[g + 1]) belong to
// group g.  MultiLevel: one level (chains [first, first + nch) of a level
// set, operation op); MultiRun: a narrow run, block -> (component, tile).
struct MultiLevel {
  int ngroups;
  int block_start[kMaxGroups + 1];
  int first[kMaxGroups], nch[kMaxGroups], tiles[kMaxGroups], B[kMaxGroups];
  Pat pat[kMaxGroups];
  Lev lev[kMaxGroups];
  const double *a[kMaxGroups];
  double *luv[kMaxGroups];
  const double *b[kMaxGroups];
  double *x[kMaxGroups];
};

struct MultiRun {
  int ngroups;
  int block_start[kMaxGroups + 1];
  int tiles[kMaxGroups], B[kMaxGroups];
  LvRange R[kMaxGroups];
  Pat pat[kMaxGroups];
  Lev lu[kMaxGroups], l[kMaxGroups], u[kMaxGroups], lt[kMaxGroups];
  const double *a[kMaxGroups];
  double *luv[kMaxGroups];
  const double *b[kMaxGroups];
  double *x[kMaxGroups];
};

// Launchers of the numeric kernels at instance width W.  All are stream
// ordered and neither synchronize nor allocate.
template <int W> struct Launch {
  // one level (chains [first, first + nch) of a level set) with operation op;
  // grid: ceil(nch * tiles * W / 256) blocks of 256 threads
  // stream: non-temporal hints on the streaming accesses (W == 64 only)
  static void level(const Pat &pat, const Lev &lev, int op, int first,
                    int nch, int tiles, const double *A, double *LU,
                    const double *b, double *x, int B, hipStream_t s,
                    bool stream = false);
  // a run of narrow levels, block -> (component comp0 + k, tile)
  static void hybrid(const Pat &pat, const Lev &lu, const Lev &l,
                     const Lev &u, const Lev &lt, const LvRange &R, int comp0,
                     int ncomp, int tiles, int threads, const double *A,
                     double *LU, const double *b, double *x, int B,
                     hipStream_t s, bool stream = false);
  // the whole refactor / solve, block -> (component, tile)
  static void component(const Pat &pat, const Lev &lu, const Lev &l,
                        const Lev &u, const Lev &lt, int mode, bool fuse,
                        int ncomp, int tiles, int threads, const double *A,
                        double *LU, const double *b, double *x, int B,
                        hipStream_t s);
  // multi-pattern work list
  static void multi(const MultiArgs &m, int mode, int blocks, int threads,
                    hipStream_t s);
  // multi-pattern LEVEL / HYBRID steps (256-thread blocks for levels,
  // `threads` <= 256 for runs)
  static void level_multi(const MultiLevel &m, int op, int blocks,
                          hipStream_t s);
  static void run_multi(const MultiRun &m, int blocks, int threads,
                        hipStream_t s);
};

} // namespace pvglu
