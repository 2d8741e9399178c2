// pvglu_newton.hip -- Newton drivers of the config PV plants.
//
// Drop-in driver (pvglu_pv_newton_dropin): the reference's host call sequence
// (PV/000_PVPP_simulator.cpp:163-235) -- host assembly on the map-based
// matrix, symmetric + row permutation, CSC value fill, then the GPU refactor,
// GPU solve and the 011 damped update on the GPU -- one instance at a time.
//
// Batched driver (pvglu_pv_newton_batched): B scenarios of one plant share
// the pattern; every Newton iteration is one replay of a hipGraph:
//   k_eval (module I(V), G = -dI/dV, I_eq per element and instance)
//   -> k_stamp_a / k_stamp_b (gather of the companion stamps into the
//      permuted CSC values and right-hand side, in the host assembly's order)
//   -> refactor + solve (pvglu_refactor_solve)
//   -> 011 update (pvglu_newton_update).
// The adjoint (pvglu_pv_gradients) reuses the factors for A^T mu = dP/dx.
#include <hip/hip_runtime.h>

#include "pvglu_launch.h"
#include "pvglu_module.h"
#include "pvglu_pv.h"
#include "pvglu_update.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

using namespace pvglu;

#define HIP_TRY(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail(PVGLU_E_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                         \
  } while (0)

static const double kLadder[] = {100, 10, 1, 0.9, 0.8, 0.5, 0.4, 0.3, 0.2, 0.1,
                                 0.05, -1}; // PV/000_PVPP_simulator.cpp:101-102

namespace pvglu {

struct DropinDev {
  double *a = nullptr, *lu = nullptr, *b = nullptr, *x = nullptr, *v = nullptr,
         *dsum = nullptr, *dmax = nullptr;
  int *pos = nullptr, *neg = nullptr;
};

void dropin_dev_destroy(DropinDev *d) {
  if (!d)
    return;
  for (void *p : {(void *)d->a, (void *)d->lu, (void *)d->b, (void *)d->x,
                  (void *)d->v, (void *)d->dsum, (void *)d->dmax,
                  (void *)d->pos, (void *)d->neg})
    (void)hipFree(p);
  delete d;
}

struct NewtonDev {
  int B = 0;
  // pattern-side (shared by instances)
  double *base_a = nullptr, *base_b = nullptr, *rsh = nullptr,
         *cab_g = nullptr, *lk = nullptr; // lk: module_log_k per module
  int *a_ptr = nullptr, *a_mod = nullptr, *a_sgn = nullptr, *b_ptr = nullptr,
      *b_mod = nullptr, *b_sgn = nullptr, *pos = nullptr, *neg = nullptr,
      *rpos = nullptr, *rneg = nullptr, *cab_ra = nullptr, *cab_rb = nullptr,
      *cab_ca = nullptr, *cab_cb = nullptr, *bus_c = nullptr;
  // linear branches per permuted row (correction-form residual): terminals
  // ca, cb (permuted unknowns; cb < 0: a source of EMF l_e), signed
  // conductance l_g (+ at the row of ca, - at the row of cb)
  int *l_ptr = nullptr, *l_ca = nullptr, *l_cb = nullptr;
  double *l_g = nullptr, *l_e = nullptr;
  // correction form: the element terminals with the kOwnTag of the element
  // that writes each node's next voltage (upos, uneg) and the nodes no
  // element terminal reaches (left[nleft]), which the update's extra blocks
  // write
  int *upos = nullptr, *uneg = nullptr, *left = nullptr;
  int nleft = 0;
  // instance-side
  // ieq: the companion current source f - G v (direct form) or the branch
  // current at the new node voltages f + G (V - v) (correction form)
  double *G = nullptr, *v = nullptr, *geq = nullptr, *ieq = nullptr,
         *A = nullptr, *LU = nullptr, *b = nullptr, *x = nullptr,
         *X[2] = {nullptr, nullptr}, *dx = nullptr,
         *gvs = nullptr, *gP = nullptr,
         *mu = nullptr, *dG = nullptr, *dgc = nullptr, *lastmax = nullptr,
         *ladder = nullptr;
  // per-instance sum (2^-44 V fixed point) and max (double bits) of |change|
  unsigned long long *dsum = nullptr, *dmax = nullptr;
  int *stage = nullptr, *active = nullptr, *stall = nullptr, *iters = nullptr,
      *remaining = nullptr;
  // set by the control step that freezes an instance, cleared by the next
  // one: the update in between copies the instance's final node voltages
  // into the other node buffer
  int *justf = nullptr;
  int nstages = 12; // stages of the current solve (12 = ladder, 1 = last only)
  double tol = 1e-10;    // captured into the graph; re-captured when changed
  int form = PVGLU_NEWTON_CORRECTION; // form the graphs were captured for
  int skip = 1; // frozen-tile skipping the graphs were captured with
  // stamping on load (correction form): the records of the LU string
  // segments (kStampRec ints per kernel position; host copy kept for the
  // coverage of each batch), the A slots the refactor does not stamp for the
  // captured batch (k_stamp_a_list writes them) and whether the captured
  // graphs use it
  int *stamp_seg = nullptr, *a_part = nullptr;
  std::vector<int> h_stamp_seg;
  int npart = 0;
  bool stamped = false; // the captured graphs stamp on load
  // the rows whose correction-form right-hand side the stamped refactor does
  // not form (the residual kernel's row list when it forms any) and whether
  // the captured graphs use it
  int *res_list = nullptr;
  int nres = 0;
  bool res_on_load = false;
  int fused = 1;        // stamping on load requested when captured
  hipStream_t s = nullptr;
  // the pipelined poll of the batched loop: the remaining-instance count of
  // each burst lands in pinned host memory, behind an event
  int *hrem = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  // one graph per iteration; the correction form alternates two (parity p
  // reads node voltages X[p] and its update writes X[1 - p])
  hipGraphExec_t exec = nullptr, exec2 = nullptr;
  hipGraph_t graph = nullptr, graph2 = nullptr;
  std::vector<void *> allocs;
  std::vector<void *> inst_allocs;
  // compaction of the batched Newton: two alternate sets of the per-instance
  // state the iteration carries over (allocated at the first compaction of a
  // batch size) and index lists [3 * B] (survivors, leavers, their home
  // instances)
  struct InstSet {
    double *G = nullptr, *v = nullptr, *geq = nullptr, *ieq = nullptr,
           *X0 = nullptr, *X1 = nullptr, *dx = nullptr, *x = nullptr,
           *gvs = nullptr, *lastmax = nullptr;
    int *stage = nullptr, *stall = nullptr, *iters = nullptr;
  };
  InstSet alt[2];
  int *cidx = nullptr;
  std::vector<void *> alt_allocs;
};

void newton_dev_destroy(NewtonDev *d) {
  if (!d)
    return;
  for (hipGraphExec_t e : {d->exec, d->exec2})
    if (e)
      (void)hipGraphExecDestroy(e);
  for (hipGraph_t g : {d->graph, d->graph2})
    if (g)
      (void)hipGraphDestroy(g);
  if (d->s)
    (void)hipStreamDestroy(d->s);
  for (hipEvent_t e : d->ev)
    if (e)
      (void)hipEventDestroy(e);
  if (d->hrem)
    (void)hipHostFree(d->hrem);
  for (void *p : d->allocs)
    (void)hipFree(p);
  for (void *p : d->inst_allocs)
    (void)hipFree(p);
  for (void *p : d->alt_allocs)
    (void)hipFree(p);
  delete d;
}

} // namespace pvglu

// ----------------------------------------------------------------- kernels

// Rows per block of the row gathers (stamping, correction-form residual).
constexpr int kStampRows = 8;

// Companion model of each module (PV/002b_sistema_electrico_base_matriz.cpp:183-203):
// element current f(V) = -I(V) flows + -> -, G = f'(V), I_eq = f(V) - G V.
__global__ void k_eval(int nm, int B, const double *__restrict__ Girr,
                       const double *__restrict__ rsh,
                       const double *__restrict__ lk,
                       const double *__restrict__ v, double *__restrict__ geq,
                       double *__restrict__ ieq, const int *active) {
  PVGLU_ROW_INDEX(nm, B);
  if (active && !active[inst])
    return; // frozen instance: its system (and so its solution) stays as is
  const int e = row_;
  const double V = v[gid];
  const ModuleEval me =
      module_current_lk(V, kIphPerIrradiance * Girr[gid], rsh[e], lk[e]);
  const double f = -me.I, g = -me.dIdV;
  geq[gid] = g;
  ieq[gid] = f - g * V;
}

// The module operands of an element's evaluation, loaded with its update.
struct ModuleOperands {
  double g, rsh, lk;
};

// 011 update fused with the companion models of the next iteration: once a
// module's voltage is updated (update_elements), its conductance and current
// source for the next stamping are evaluated from the new voltage in
// registers, with k_eval's operations (so the same values), saving k_eval's
// read of the voltages and one launch per iteration.  An instance that the
// control kernel freezes right after this update gets an evaluation that
// nothing reads (stamping skips frozen instances); the control kernel does
// not change voltages.
// Direct form (CORR false, dx null): x holds the new node voltages;
// ieq <- f - G v.  Correction form (CORR): x + dx are the new node voltages (x = the previous ones,
// dx = the solve's correction) and ieq <- f + G (V - v), the linearized
// branch current at those node voltages, which the next iteration's residual
// gathers (V - v = 0 exactly when the step was not damped).  The update also
// writes those node voltages, x + dx, to xo for the next iteration: at the
// element terminals tagged kOwnTag (pos, neg are the tagged lists) and, in
// the blocks past the nblk element blocks, at the nodes left[0..nleft) that
// no element reaches.  The additions are those the next iteration would
// otherwise make over the whole vector, so the iterates are the same bits.
template <bool CORR>
__global__ void __launch_bounds__(256)
    k_update_eval(int ne, int B, const int *__restrict__ pos,
                  const int *__restrict__ neg, const double *__restrict__ x,
                  const double *__restrict__ dx, double *__restrict__ v,
                  unsigned long long *dsum, unsigned long long *dmax,
                  const int *active, int *status,
                  const double *__restrict__ Girr,
                  const double *__restrict__ rsh,
                  const double *__restrict__ lk, double *__restrict__ geq,
                  double *__restrict__ ieq, int *remaining,
                  double *__restrict__ xo, const int *__restrict__ justf,
                  int nblk, const int *__restrict__ left, int nleft) {
  // the control kernel after this one counts the unfinished instances
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0)
    *remaining = 0;
  if (CORR && (int)blockIdx.x >= nblk) { // nodes without an element terminal
    const int lt = upd_tile_log2(B), rp = 256 >> lt;
    const int in =
        (int)(blockIdx.y << lt) + (int)(threadIdx.x & ((1u << lt) - 1u));
    if (in >= B)
      return;
    const bool live = active[in] != 0, copy = !live && justf[in] != 0;
    if (!live && !copy)
      return;
    const int k0 = ((int)blockIdx.x - nblk) * rp * kUpdPerThread;
    const int k1 = min(nleft, k0 + rp * kUpdPerThread);
    for (int q = k0 + (int)(threadIdx.x >> lt); q < k1; q += rp) {
      const size_t i = (size_t)left[q] * B + in;
      xo[i] = live ? x[i] + dx[i] : x[i];
    }
    return;
  }
  update_elements<CORR>(
      ne, B, pos, neg, x, dx, v, dsum, dmax, active, status,
      [&](int e, int in) {
        return ModuleOperands{Girr[(size_t)e * B + in], rsh[e], lk[e]};
      },
      [&](int e, int in, double V, double Vx, const ModuleOperands &m) {
        const size_t gid = (size_t)e * B + in;
        const ModuleEval me =
            module_current_lk(V, kIphPerIrradiance * m.g, m.rsh, m.lk);
        const double f = -me.I, g = -me.dIdV;
        // read once, by the next refactor's streaming loads (or the
        // stamping kernels), after a whole sweep of other traffic
        __builtin_nontemporal_store(g, &geq[gid]);
        __builtin_nontemporal_store(CORR ? f + g * (Vx - V) : f - g * V,
                                    &ieq[gid]);
      },
      xo, justf);
}

// Correction-form right-hand side.  Per permuted row (equation of one node)
// and instance: minus the sum of the currents leaving the node through its
// branches, at the node voltages x (this iteration's: the previous iterate
// plus its correction, written by the previous update):
//   modules  ibr (from k_update_eval), + g_virtual (x[pos] - x[neg]) on
//            homotopy stages;
//   linear   l_g (x[ca] - x[cb]), cb < 0 meaning a source of EMF l_e (the
//            Norton bus: G_N (x - 600)).
// Each difference of node voltages is formed before it is scaled, so a
// string's ~600 V node voltages cancel exactly and the residual carries the
// rounding of the branch currents only.  Frozen instances get a zero
// right-hand side (their correction is not used: the update skips them).
// Layout: one block per kStampRows rows x up to 256 instances, as the
// stamping gathers; a string row reads its two modules' branch currents only.  `rows` (may be
// null: rows 0 .. n-1) lists the rows to compute: with the right-hand side
// formed on load by the stamped refactor, only the rows it does not form.
__global__ void __launch_bounds__(256)
    k_residual(int n, int B, const int *__restrict__ mptr,
               const int *__restrict__ mmod, const int *__restrict__ msgn,
               const int *__restrict__ lptr, const int *__restrict__ lca,
               const int *__restrict__ lcb, const double *__restrict__ lg,
               const double *__restrict__ le, const int *__restrict__ pos,
               const int *__restrict__ neg, const double *__restrict__ ibr,
               const double *__restrict__ gvp, const double *__restrict__ x,
               double *__restrict__ r, const int *active,
               const int *__restrict__ rows) {
  const int inst = (int)(blockIdx.y * blockDim.x + threadIdx.x);
  if (inst >= B)
    return;
  const int r0 = (int)blockIdx.x * kStampRows;
  const bool live = !active || active[inst];
  const double gv = live ? gvp[inst] : -1.0;
  auto node = [&](int k) {
    return k < 0 ? 0.0 : x[(size_t)k * B + inst];
  };
  // the block's rows (n = the number of rows listed) and their ranges
  int row[kStampRows], m0p[kStampRows], m1p[kStampRows], l0p[kStampRows],
      l1p[kStampRows];
  if (rows) {
#pragma unroll
    for (int q = 0; q < kStampRows; ++q)
      row[q] = rows[min(r0 + q, n - 1)];
#pragma unroll
    for (int q = 0; q < kStampRows; ++q) {
      m0p[q] = mptr[row[q]];
      m1p[q] = mptr[row[q] + 1];
      l0p[q] = lptr[row[q]];
      l1p[q] = lptr[row[q] + 1];
    }
  } else { // consecutive rows share their range bounds
    int mb[kStampRows + 1], lb[kStampRows + 1];
#pragma unroll
    for (int q = 0; q <= kStampRows; ++q) {
      mb[q] = mptr[min(r0 + q, n)];
      lb[q] = lptr[min(r0 + q, n)];
    }
#pragma unroll
    for (int q = 0; q < kStampRows; ++q) {
      row[q] = min(r0 + q, n - 1);
      m0p[q] = mb[q];
      m1p[q] = mb[q + 1];
      l0p[q] = lb[q];
      l1p[q] = lb[q + 1];
    }
  }
  if (!live) {
    for (int q = 0; q < kStampRows && r0 + q < n; ++q)
      r[(size_t)row[q] * B + inst] = 0.0;
    return;
  }
  bool simple = gv <= 0;
  bool anylin = false;
#pragma unroll
  for (int q = 0; q < kStampRows; ++q) {
    simple = simple && m1p[q] - m0p[q] <= 2 && l1p[q] - l0p[q] <= 1;
    anylin = anylin || l1p[q] > l0p[q];
  }
  if (simple) {
    // every load first (absent terms load a valid stand-in), then combine
    double i0[kStampRows], i1[kStampRows], xa[kStampRows], xb[kStampRows];
#pragma unroll
    for (int q = 0; q < kStampRows; ++q) {
      const int cm = m1p[q] - m0p[q];
      const int m0 = cm > 0 ? mmod[m0p[q]] : 0;
      const int m1 = cm > 1 ? mmod[m0p[q] + 1] : m0;
      i0[q] = ibr[(size_t)m0 * B + inst];
      i1[q] = ibr[(size_t)m1 * B + inst];
      xa[q] = xb[q] = 0.0;
    }
    if (anylin) { // combiner and bus rows: node voltages of the linear branch
#pragma unroll
      for (int q = 0; q < kStampRows; ++q)
        if (l1p[q] > l0p[q]) {
          xa[q] = node(lca[l0p[q]]);
          const int cb = lcb[l0p[q]];
          xb[q] = cb >= 0 ? node(cb) : le[l0p[q]];
        }
    }
#pragma unroll
    for (int q = 0; q < kStampRows; ++q) {
      if (r0 + q >= n)
        break;
      double acc = 0.0;
      const int cm = m1p[q] - m0p[q];
      if (cm > 0)
        acc = acc - (double)msgn[m0p[q]] * i0[q];
      if (cm > 1)
        acc = acc - (double)msgn[m0p[q] + 1] * i1[q];
      if (l1p[q] > l0p[q])
        acc = acc - lg[l0p[q]] * (xa[q] - xb[q]);
      r[(size_t)row[q] * B + inst] = acc;
    }
    return;
  }
  for (int q = 0; q < kStampRows && r0 + q < n; ++q) {
    double acc = 0.0;
    for (int k = m0p[q]; k < m1p[q]; ++k) {
      const int e = mmod[k];
      double ie = ibr[(size_t)e * B + inst];
      if (gv > 0)
        ie = ie + gv * (node(pos[e]) - node(neg[e]));
      acc = acc - (double)msgn[k] * ie;
    }
    for (int k = l0p[q]; k < l1p[q]; ++k) {
      const int cb = lcb[k];
      const double xb = cb >= 0 ? node(cb) : le[k];
      acc = acc - lg[k] * (node(lca[k]) - xb);
    }
    r[(size_t)row[q] * B + inst] = acc;
  }
}

// Stamping gathers.  A block covers kStampRows consecutive rows (slots of A,
// or rows of b) for up to 256 instances (stamp_grid), so the rows' metadata
// is uniform in the block and comes through scalar loads.  (The row_grid
// shape, several row groups per wavefront, fills every lane at small batches
// but turns those loads into per-lane vector loads, and measured slower: 129
// vs 104 us for A at 32 instances of the 400-inverter plant.)  When every row
// of the block has at most two contributing modules (all rows of a PV plant),
// the block loads all its module values first and then combines them row by
// row, so a block is one memory round trip; otherwise the rows are combined
// one at a time.  The
// operation order per row is fixed (contributions in map order).

inline RowGrid stamp_grid(long rows, int B) {
  const int bs = B >= 256 ? 256 : ((B + 63) / 64) * 64;
  return {dim3((unsigned)((rows + kStampRows - 1) / kStampRows),
               (unsigned)((B + bs - 1) / bs)),
          dim3(bs)};
}

// out[row] = base[row] + sum_k sign_k * val[mod_k] (+ sign_k * gv per term
// when gv > 0; SUB: minus instead of plus, without gv).
template <bool SUB>
__device__ __forceinline__ void stamp_rows(int rows, int B,
                                           const double *__restrict__ base,
                                           const int *__restrict__ ptr,
                                           const int *__restrict__ mod,
                                           const int *__restrict__ sgn,
                                           const double *__restrict__ val,
                                           double gv, double *__restrict__ out,
                                           int r0, int inst) {
  int kp[kStampRows + 1];
#pragma unroll
  for (int r = 0; r <= kStampRows; ++r)
    kp[r] = ptr[min(r0 + r, rows)];
  bool simple = true;
#pragma unroll
  for (int r = 0; r < kStampRows; ++r)
    simple = simple && kp[r + 1] - kp[r] <= 2;
  auto term = [&](double acc, int k, double v) {
    const double sg = (double)sgn[k];
    if constexpr (SUB)
      return acc - sg * v;
    acc = acc + sg * v;
    if (gv > 0)
      acc = acc + sg * gv;
    return acc;
  };
  if (simple) {
    double v0[kStampRows], v1[kStampRows], bs[kStampRows];
#pragma unroll
    for (int r = 0; r < kStampRows; ++r) {
      const int row = min(r0 + r, rows - 1);
      bs[r] = base[row];
      // a row with fewer terms loads module 0 in their place (always valid)
      const int cnt = kp[r + 1] - kp[r];
      const int m0 = cnt > 0 ? mod[kp[r]] : 0;
      const int m1 = cnt > 1 ? mod[kp[r] + 1] : m0;
      v0[r] = val[(size_t)m0 * B + inst];
      v1[r] = val[(size_t)m1 * B + inst];
    }
#pragma unroll
    for (int r = 0; r < kStampRows; ++r) {
      if (r0 + r >= rows)
        break;
      const int cnt = kp[r + 1] - kp[r];
      double acc = bs[r];
      if (cnt > 0)
        acc = term(acc, kp[r], v0[r]);
      if (cnt > 1)
        acc = term(acc, kp[r] + 1, v1[r]);
      out[(size_t)(r0 + r) * B + inst] = acc;
    }
    return;
  }
  for (int r = 0; r < kStampRows && r0 + r < rows; ++r) {
    double acc = base[r0 + r];
    for (int k = kp[r]; k < kp[r + 1]; ++k)
      acc = term(acc, k, val[(size_t)mod[k] * B + inst]);
    out[(size_t)(r0 + r) * B + inst] = acc;
  }
}

// A[q] = linear part + sum over the modules stamping slot q, each adding
// sign*G and, on homotopy stages, sign*g_virtual.
__global__ void __launch_bounds__(256)
    k_stamp_a(int nnz, int B, const double *__restrict__ base,
              const int *__restrict__ ptr, const int *__restrict__ mod,
              const int *__restrict__ sgn, const double *__restrict__ geq,
              const double *__restrict__ gvp, double *__restrict__ A,
              const int *active) {
  const int inst = (int)(blockIdx.y * blockDim.x + threadIdx.x);
  if (inst >= B || (active && !active[inst]))
    return;
  // gvp: homotopy conductance of this instance's stage
  stamp_rows<false>(nnz, B, base, ptr, mod, sgn, geq, gvp[inst], A,
                    (int)blockIdx.x * kStampRows, inst);
}

// k_stamp_a over the slots list[0..cnt) only (the slots a stamping-on-load
// refactor does not form itself), same operations per slot.
__global__ void __launch_bounds__(256)
    k_stamp_a_list(int cnt, const int *__restrict__ list, int B,
                   const double *__restrict__ base, const int *__restrict__ ptr,
                   const int *__restrict__ mod, const int *__restrict__ sgn,
                   const double *__restrict__ geq,
                   const double *__restrict__ gvp, double *__restrict__ A,
                   const int *active) {
  const int inst = (int)(blockIdx.y * blockDim.x + threadIdx.x);
  if (inst >= B || (active && !active[inst]))
    return;
  const double gv = gvp[inst];
  const int i0 = (int)blockIdx.x * kStampRows;
  for (int r = 0; r < kStampRows && i0 + r < cnt; ++r) {
    const int q = list[i0 + r];
    double acc = base[q];
    for (int k = ptr[q]; k < ptr[q + 1]; ++k) {
      const double sg = (double)sgn[k];
      acc = acc + sg * geq[(size_t)mod[k] * B + inst];
      if (gv > 0)
        acc = acc + sg * gv;
    }
    A[(size_t)q * B + inst] = acc;
  }
}

// b[r] = linear sources - sum over modules of sign * I_eq
__global__ void __launch_bounds__(256)
    k_stamp_b(int n, int B, const double *__restrict__ base,
              const int *__restrict__ ptr, const int *__restrict__ mod,
              const int *__restrict__ sgn, const double *__restrict__ ieq,
              double *__restrict__ b, const int *active) {
  const int inst = (int)(blockIdx.y * blockDim.x + threadIdx.x);
  if (inst >= B || (active && !active[inst]))
    return;
  stamp_rows<true>(n, B, base, ptr, mod, sgn, ieq, 0.0, b,
                   (int)blockIdx.x * kStampRows, inst);
}

// dP/dx on the bus unknowns (permuted): P = sum_bus V G_N (V - 600);
// zero elsewhere (k_zero first, then k_power_grad_bus)
__global__ void k_zero(long count, double *p) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid < count)
    p[gid] = 0.0;
}

// Compaction of a [row][instance] array (rows x Bc, instance-minor) to the
// instances sel[0..Bn) (rows x Bn): out[r][j] = in[r][sel[j]].
template <typename T>
__global__ void k_gather(long rows, int Bc, int Bn, const int *__restrict__ sel,
                         const T *__restrict__ in, T *__restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * Bn)
    return;
  const long r = i / Bn;
  const long src = r * Bc + sel[i - r * Bn];
  out[i] = in[src];
}

// The inverse for the instances leaving a compacted batch: out[r][to[k]] =
// in[r][from[k]], out with Bh instances per row (the caller's batch).
template <typename T>
__global__ void k_scatter(long rows, int Bc, int Bh, int cnt,
                          const int *__restrict__ from,
                          const int *__restrict__ to, const T *__restrict__ in,
                          T *__restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * cnt)
    return;
  const long r = i / cnt;
  const long k = i - r * cnt;
  const long src = r * Bc + from[k];
  out[r * Bh + to[k]] = in[src];
}

__global__ void k_power_grad_bus(int nbus, int B, const int *__restrict__ bus_c,
                                 const double *__restrict__ x, double *gP) {
  PVGLU_ROW_INDEX(nbus, B);
  (void)gid;
  const int k = row_;
  const size_t idx = (size_t)bus_c[k] * B + inst;
  gP[idx] = kNortonG * (2.0 * x[idx] - kNortonV);
}

// dP/dG_m = 9e-3 dI/dIph (mu[+] - mu[-]); mu in row-permuted order
__global__ void k_grad_modules(int nm, int B, const double *__restrict__ Girr,
                               const double *__restrict__ rsh,
                               const double *__restrict__ v,
                               const int *__restrict__ rpos,
                               const int *__restrict__ rneg,
                               const double *__restrict__ mu, double *dG) {
  PVGLU_ROW_INDEX(nm, B);
  const int e = row_;
  const ModuleEval me =
      module_current(v[gid], kIphPerIrradiance * Girr[gid], rsh[e]);
  const double lp = rpos[e] >= 0 ? mu[(size_t)rpos[e] * B + inst] : 0.0;
  const double ln = rneg[e] >= 0 ? mu[(size_t)rneg[e] * B + inst] : 0.0;
  dG[gid] = kIphPerIrradiance * me.dIdIph * (lp - ln);
}

// dP/dg_c = -(mu[a] - mu[b]) (V_a - V_b)
__global__ void k_grad_cables(int nc, int B, const int *__restrict__ ra,
                              const int *__restrict__ rb,
                              const int *__restrict__ ca,
                              const int *__restrict__ cb,
                              const double *__restrict__ x,
                              const double *__restrict__ mu, double *dg) {
  PVGLU_ROW_INDEX(nc, B);
  const int c = row_;
  const double dv = x[(size_t)ca[c] * B + inst] - x[(size_t)cb[c] * B + inst];
  const double dl = mu[(size_t)ra[c] * B + inst] - mu[(size_t)rb[c] * B + inst];
  dg[gid] = -dl * dv;
}


// module_log_k per module, once per plant (the evaluators then skip one
// logarithm per module and instance)
__global__ void k_module_lk(int nm, const double *__restrict__ rsh,
                            double *__restrict__ lk) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nm)
    lk[e] = module_log_k(rsh[e]);
}

__global__ void k_fill_double(double *p, int n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = v;
}

// Per-instance Newton state at the start of a solve.
__global__ void k_newton_init(int B, double gv0, int *stage, double *gvs,
                              int *active, int *stall, int *iters,
                              double *lastmax, unsigned long long *dsum,
                              unsigned long long *dmax, int *justf) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= B)
    return;
  justf[k] = 0;
  dsum[k] = 0; // the update accumulates into these; control clears them
  dmax[k] = 0;
  stage[k] = 0;
  gvs[k] = gv0;
  active[k] = 1;
  stall[k] = 0;
  iters[k] = 0;
  lastmax[k] = 0.0;
}

// Per-instance control after each iteration: the reference's loop
// (PV/000_PVPP_simulator.cpp:163-181) run independently for every instance.
// Intermediate homotopy stages end when sum |delta| < 1e-6 n; the last stage
// ends when max |delta| <= tol, or -- at the FP64 noise floor of the system --
// after stall_max iterations with max |delta| <= noise.  Finished instances
// are frozen (the update kernel skips them); `remaining` counts the others.
__global__ void k_newton_control(int B, int nstages,
                                 const double *__restrict__ ladder,
                                 double stage_crit, double tol, double noise,
                                 int stall_max, unsigned long long *dsum,
                                 unsigned long long *dmax, int *stage,
                                 double *gvs, int *active, int *stall,
                                 int *iters, double *lastmax, int *remaining,
                                 int *justf) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= B)
    return;
  if (!active[k]) {
    if (justf) // the update since the freeze has copied its node voltages
      justf[k] = 0;
    return;
  }
  iters[k]++;
  const int st = stage[k];
  const double m = __longlong_as_double((long long)dmax[k]);
  const double sum = fixed_to_delta(dsum[k]);
  dsum[k] = 0; // cleared for the next iteration's update
  dmax[k] = 0;
  lastmax[k] = m;
  if (st < nstages - 1) {
    if (sum < stage_crit) {
      stage[k] = st + 1;
      gvs[k] = ladder[st + 1];
    }
    atomicAdd(remaining, 1);
    return;
  }
  if (!(m > tol)) { // converged (NaN is caught by the status word)
    active[k] = 0;
    if (justf)
      justf[k] = 1;
    return;
  }
  if (m <= noise) {
    if (++stall[k] >= stall_max) {
      active[k] = 0;
      if (justf)
        justf[k] = 1;
      return;
    }
  } else {
    stall[k] = 0;
  }
  atomicAdd(remaining, 1);
}

// ------------------------------------------------------------ host helpers

namespace {

template <typename T>
int dupload(std::vector<void *> &allocs, T **dst, const std::vector<T> &src) {
  void *p = nullptr;
  HIP_TRY(hipMalloc(&p, sizeof(T) * std::max<size_t>(src.size(), 1)));
  allocs.push_back(p);
  if (!src.empty())
    HIP_TRY(hipMemcpy(p, src.data(), sizeof(T) * src.size(),
                      hipMemcpyHostToDevice));
  *dst = (T *)p;
  return PVGLU_OK;
}

template <typename T>
int dalloc(std::vector<void *> &allocs, T **dst, size_t count) {
  void *p = nullptr;
  HIP_TRY(hipMalloc(&p, sizeof(T) * std::max<size_t>(count, 1)));
  allocs.push_back(p);
  *dst = (T *)p;
  return PVGLU_OK;
}

// slot of entry (r, c) in the plan (CSR search, then CSR->CSC map)
int slot_of(const pvglu_plan *P, int r, int c) {
  const int *b = P->ci.data() + P->rp[r], *e = P->ci.data() + P->rp[r + 1];
  const int *it = std::lower_bound(b, e, c);
  if (it == e || *it != c)
    return -1;
  return P->csr2csc[(int)(it - P->ci.data())];
}

// rowmap: original row -> permuted row; colmap: unknown -> permuted unknown
void perm_maps(const pvglu_pv *pv, std::vector<int> &rowmap,
               std::vector<int> &colmap) {
  rowmap.resize(pv->n);
  colmap.resize(pv->n);
  for (int i = 0; i < pv->n; i++) {
    colmap[i] = pv->Pinv[i];
    rowmap[i] = pv->Prow[pv->Pinv[i]];
  }
}

int newton_setup(pvglu_pv *pv) {
  if (pv->nd)
    return PVGLU_OK;
  int st = pv_prepare(pv);
  if (st)
    return st;
  const pvglu_plan *P = pv->plan;
  const int n = pv->n, nnz = P->nnz, nm = pv->nmod;
  std::vector<int> rowmap, colmap;
  perm_maps(pv, rowmap, colmap);
  // linear part in permuted CSC slots / rows
  std::vector<double> base_a(nnz, 0.0), base_b(n, 0.0);
  for (int r = 0; r < n; r++)
    for (auto &e : pv->lin_A.rows[r]) {
      const int q = slot_of(P, rowmap[r], colmap[e.first]);
      if (q < 0)
        return fail(PVGLU_E_ARG, "newton_setup: linear stamp outside pattern");
      base_a[q] = e.second;
    }
  for (int r = 0; r < n; r++)
    base_b[rowmap[r]] = pv->lin_b[r];
  // module contributions per slot and per row, in element order
  std::vector<std::vector<std::pair<int, int>>> as(nnz), bs(n);
  for (int e = 0; e < nm; e++) {
    const int p = pv->mod_pos[e], m = pv->mod_neg[e];
    auto add = [&](int r, int c, int s) {
      const int q = slot_of(P, rowmap[r], colmap[c]);
      as[q].push_back({e, s});
    };
    if (p >= 0)
      add(p, p, 1);
    if (m >= 0)
      add(m, m, 1);
    if (p >= 0 && m >= 0) {
      add(p, m, -1);
      add(m, p, -1);
    }
    if (p >= 0)
      bs[rowmap[p]].push_back({e, 1});
    if (m >= 0)
      bs[rowmap[m]].push_back({e, -1});
  }
  auto flat = [](const std::vector<std::vector<std::pair<int, int>>> &l,
                 std::vector<int> &ptr, std::vector<int> &mod,
                 std::vector<int> &sgn) {
    ptr.assign(1, 0);
    for (auto &v : l) {
      for (auto &x : v) {
        mod.push_back(x.first);
        sgn.push_back(x.second);
      }
      ptr.push_back((int)mod.size());
    }
  };
  std::vector<int> a_ptr, a_mod, a_sgn, b_ptr, b_mod, b_sgn;
  flat(as, a_ptr, a_mod, a_sgn);
  flat(bs, b_ptr, b_mod, b_sgn);
  // Stamping-on-load records: a tridiagonal string segment of the refactor
  // (with the fused forward row) whose three entries per column are, in
  // every column, one module's conductance (U above), another's (L below)
  // and their sum (diagonal), with zero linear parts, the modules advancing
  // by constant steps and the signs and the diagonal's term order fixed.
  std::vector<int> sseg((size_t)n * kStampRec, 0);
  for (int t = 0; t < n; ++t) {
    const int *sg = &P->fast_lu_seg[(size_t)t * 4];
    const int len = sg[0];
    if (len <= 0 || !(sg[3] & 8))
      continue;
    const int slot0 = P->fast_lu_rec[(size_t)t * 16], dslot = sg[1];
    int r[kStampRec] = {0, 0, 0, 0, 0, 0, 0, 0};
    bool ok = true;
    for (int i = 0; i < len && ok; ++i) {
      const int q = slot0 + i * dslot;
      auto cnt = [&](int qq) { return a_ptr[qq + 1] - a_ptr[qq]; };
      if (base_a[q] != 0.0 || base_a[q + 1] != 0.0 || base_a[q + 2] != 0.0 ||
          cnt(q) != 1 || cnt(q + 1) != 2 || cnt(q + 2) != 1) {
        ok = false;
        break;
      }
      const int mu = a_mod[a_ptr[q]], ml = a_mod[a_ptr[q + 2]];
      const int m1 = a_mod[a_ptr[q + 1]], m2 = a_mod[a_ptr[q + 1] + 1];
      const bool first_u = m1 == mu && m2 == ml;
      if (!first_u && !(m1 == ml && m2 == mu)) {
        ok = false;
        break;
      }
      const int bits = (a_sgn[a_ptr[q]] < 0 ? 1 : 0) |
                       (a_sgn[a_ptr[q + 2]] < 0 ? 2 : 0) |
                       (a_sgn[a_ptr[q + 1]] < 0 ? 4 : 0) |
                       (a_sgn[a_ptr[q + 1] + 1] < 0 ? 8 : 0) |
                       (first_u ? 16 : 0);
      if (i == 0) {
        r[1] = mu;
        r[3] = ml;
        r[5] = bits;
      } else if (i == 1) {
        r[2] = mu - r[1];
        r[4] = ml - r[3];
      }
      ok = ok && bits == r[5] && mu == r[1] + i * r[2] &&
           ml == r[3] + i * r[4];
    }
    if (ok) {
      r[0] = 1;
      std::copy(r, r + kStampRec, &sseg[(size_t)t * kStampRec]);
    }
  }
  std::vector<int> pos(nm), neg(nm), rpos(nm), rneg(nm);
  for (int e = 0; e < nm; e++) {
    pos[e] = pv->mod_pos[e] >= 0 ? colmap[pv->mod_pos[e]] : -1;
    neg[e] = pv->mod_neg[e] >= 0 ? colmap[pv->mod_neg[e]] : -1;
    rpos[e] = pv->mod_pos[e] >= 0 ? rowmap[pv->mod_pos[e]] : -1;
    rneg[e] = pv->mod_neg[e] >= 0 ? rowmap[pv->mod_neg[e]] : -1;
  }
  // the correction form's update writes each node's next voltage at the
  // first element terminal (in element order) on the node; the nodes that no
  // terminal reaches (the DC buses behind their cables) are listed
  if (n >= kOwnTag)
    return fail(PVGLU_E_ARG, "batched Newton: %d unknowns (at most 2^30 - 1)", n);
  std::vector<int> upos(pos), uneg(neg), left;
  {
    std::vector<char> owned(n, 0);
    for (int e = 0; e < nm; e++) {
      if (pos[e] >= 0 && !owned[pos[e]]) {
        owned[pos[e]] = 1;
        upos[e] |= kOwnTag;
      }
      if (neg[e] >= 0 && !owned[neg[e]]) {
        owned[neg[e]] = 1;
        uneg[e] |= kOwnTag;
      }
    }
    for (int k = 0; k < n; k++)
      if (!owned[k])
        left.push_back(k);
  }
  std::vector<int> cra(pv->ncab), crb(pv->ncab), cca(pv->ncab), ccb(pv->ncab),
      busc(pv->I);
  for (int c = 0; c < pv->ncab; c++) {
    cra[c] = rowmap[pv->cab_a[c]];
    crb[c] = rowmap[pv->cab_b[c]];
    cca[c] = colmap[pv->cab_a[c]];
    ccb[c] = colmap[pv->cab_b[c]];
  }
  for (int k = 0; k < pv->I; k++)
    busc[k] = colmap[pv->bus[k]];
  // linear branches per permuted row, in a fixed order: cables (both ends),
  // then the Norton source of a bus
  struct LinTerm {
    int ca, cb;
    double g, e;
  };
  std::vector<std::vector<LinTerm>> lin(n);
  for (int c = 0; c < pv->ncab; c++) {
    const int a = pv->cab_a[c], b = pv->cab_b[c];
    lin[rowmap[a]].push_back({colmap[a], colmap[b], pv->cab_g[c], 0.0});
    lin[rowmap[b]].push_back({colmap[a], colmap[b], -pv->cab_g[c], 0.0});
  }
  for (int bidx : pv->bus)
    lin[rowmap[bidx]].push_back({colmap[bidx], -1, kNortonG, kNortonV});
  std::vector<int> l_ptr(1, 0), l_ca, l_cb;
  std::vector<double> l_g, l_e;
  for (auto &row : lin) {
    for (auto &t : row) {
      l_ca.push_back(t.ca);
      l_cb.push_back(t.cb);
      l_g.push_back(t.g);
      l_e.push_back(t.e);
    }
    l_ptr.push_back((int)l_ca.size());
  }
  // Right-hand side on load (record [6]): every row of the segment has
  // exactly the two module branches of its column's U-above and L-below
  // entries and no linear branch, in a branch order and with signs that are
  // the same in every row.
  for (int t = 0; t < n; ++t) {
    int *r = &sseg[(size_t)t * kStampRec];
    if (!r[0])
      continue;
    const int *sg = &P->fast_lu_seg[(size_t)t * 4];
    const int len = sg[0], dcol = sg[2];
    const int col0 = P->fast_lu_rec[(size_t)t * 16 + 1];
    // a chained string (column i's L-below module is column i+1's U-above
    // one) whose forwarded entry is the L entry (the kernel's fixed case)
    const bool chained = r[2] == r[4] && r[1] + r[2] == r[3];
    if (!chained || ((sg[3] >> 4) & 3) != 2) {
      r[6] = 0;
      continue;
    }
    int bits = -1;
    for (int i = 0; i < len; ++i) {
      const int row = col0 + i * dcol;
      const int mu = r[1] + i * r[2], ml = r[3] + i * r[4];
      const int b0 = b_ptr[row];
      if (b_ptr[row + 1] - b0 != 2 || l_ptr[row + 1] != l_ptr[row]) {
        bits = -1;
        break;
      }
      const int m0 = b_mod[b0], m1 = b_mod[b0 + 1];
      if (!((m0 == mu && m1 == ml) || (m0 == ml && m1 == mu))) {
        bits = -1;
        break;
      }
      // the modules' other terminals: the rows of the U-above and L-below
      // entries (node voltages of homotopy stages), the previous node of the
      // segment and, but in its last column, the next one (the last
      // column's is recorded in [7])
      auto joins = [&](int e, int a, int c) {
        return (pos[e] == a && neg[e] == c) || (pos[e] == c && neg[e] == a);
      };
      const int q = P->fast_lu_rec[(size_t)t * 16] + i * sg[1];
      const int rprev = P->ri[q], rnext = P->ri[q + 2];
      if (rprev != row - dcol || (i + 1 < len && rnext != row + dcol) ||
          !joins(mu, rprev, row) || !joins(ml, row, rnext)) {
        bits = -1;
        break;
      }
      if (i + 1 == len)
        r[7] = rnext;
      const int rb = 1 | (m0 == mu ? 2 : 0) | (b_sgn[b0] < 0 ? 4 : 0) |
                     (b_sgn[b0 + 1] < 0 ? 8 : 0);
      if (i > 0 && rb != bits) {
        bits = -1;
        break;
      }
      bits = rb;
    }
    r[6] = bits > 0 ? bits : 0;
  }
  auto *d = new NewtonDev();
  pv->nd = d;
  auto &A = d->allocs;
  if ((st = dupload(A, &d->base_a, base_a)) || (st = dupload(A, &d->base_b, base_b)) ||
      (st = dupload(A, &d->rsh, pv->mod_rsh)) || (st = dupload(A, &d->cab_g, pv->cab_g)) ||
      (st = dupload(A, &d->a_ptr, a_ptr)) || (st = dupload(A, &d->a_mod, a_mod)) ||
      (st = dupload(A, &d->a_sgn, a_sgn)) || (st = dupload(A, &d->b_ptr, b_ptr)) ||
      (st = dupload(A, &d->b_mod, b_mod)) || (st = dupload(A, &d->b_sgn, b_sgn)) ||
      (st = dupload(A, &d->pos, pos)) || (st = dupload(A, &d->neg, neg)) ||
      (st = dupload(A, &d->upos, upos)) || (st = dupload(A, &d->uneg, uneg)) ||
      (st = dupload(A, &d->left, left)) ||
      (st = dupload(A, &d->rpos, rpos)) || (st = dupload(A, &d->rneg, rneg)) ||
      (st = dupload(A, &d->cab_ra, cra)) || (st = dupload(A, &d->cab_rb, crb)) ||
      (st = dupload(A, &d->cab_ca, cca)) || (st = dupload(A, &d->cab_cb, ccb)) ||
      (st = dupload(A, &d->bus_c, busc)) || (st = dupload(A, &d->l_ptr, l_ptr)) ||
      (st = dupload(A, &d->l_ca, l_ca)) || (st = dupload(A, &d->l_cb, l_cb)) ||
      (st = dupload(A, &d->l_g, l_g)) || (st = dupload(A, &d->l_e, l_e)) ||
      (st = dupload(A, &d->ladder,
                    std::vector<double>(kLadder, kLadder + 12))) ||
      (st = dupload(A, &d->lk, std::vector<double>(pv->nmod, 0.0))) ||
      (st = dupload(A, &d->stamp_seg, sseg)) ||
      (st = dupload(A, &d->res_list, std::vector<int>(std::max(n, 1), 0))) ||
      (st = dupload(A, &d->a_part, std::vector<int>(std::max(nnz, 1), 0))))
    return st;
  d->h_stamp_seg.swap(sseg);
  d->nleft = (int)left.size();
  HIP_TRY(hipStreamCreate(&d->s));
  for (hipEvent_t &e : d->ev)
    HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  HIP_TRY(hipHostMalloc((void **)&d->hrem, 2 * sizeof(int)));
  if (pv->nmod > 0) {
    hipLaunchKernelGGL(k_module_lk, dim3((pv->nmod + 255) / 256), dim3(256), 0,
                       d->s, pv->nmod, d->rsh, d->lk);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(d->s));
  }
  return PVGLU_OK;
}

void drop_graphs(NewtonDev *d) {
  for (hipGraphExec_t *e : {&d->exec, &d->exec2})
    if (*e) {
      (void)hipGraphExecDestroy(*e);
      *e = nullptr;
    }
  for (hipGraph_t *g : {&d->graph, &d->graph2})
    if (*g) {
      (void)hipGraphDestroy(*g);
      *g = nullptr;
    }
}

int newton_alloc_instances(pvglu_pv *pv, int B) {
  NewtonDev *d = pv->nd;
  if (d->B == B)
    return PVGLU_OK;
  drop_graphs(d);
  for (void *p : d->inst_allocs)
    (void)hipFree(p);
  d->inst_allocs.clear();
  for (void *p : d->alt_allocs)
    (void)hipFree(p);
  d->alt_allocs.clear();
  d->alt[0] = d->alt[1] = NewtonDev::InstSet{};
  d->cidx = nullptr;
  d->B = 0;
  const size_t nm = pv->nmod, n = pv->n, nnz = pv->plan->nnz, nc = pv->ncab;
  auto &I = d->inst_allocs;
  int st;
  if ((st = dalloc(I, &d->G, nm * B)) || (st = dalloc(I, &d->v, nm * B)) ||
      (st = dalloc(I, &d->geq, nm * B)) || (st = dalloc(I, &d->ieq, nm * B)) ||
      (st = dalloc(I, &d->A, nnz * B)) || (st = dalloc(I, &d->LU, nnz * B)) ||
      (st = dalloc(I, &d->b, n * B)) || (st = dalloc(I, &d->x, n * B)) ||
      (st = dalloc(I, &d->X[0], n * B)) || (st = dalloc(I, &d->X[1], n * B)) ||
      (st = dalloc(I, &d->dx, n * B)) ||
      (st = dalloc(I, &d->dsum, (size_t)B)) || (st = dalloc(I, &d->dmax, (size_t)B)) ||
      (st = dalloc(I, &d->gP, n * B)) || (st = dalloc(I, &d->mu, n * B)) ||
      (st = dalloc(I, &d->dG, nm * B)) || (st = dalloc(I, &d->dgc, nc * B)) ||
      (st = dalloc(I, &d->gvs, (size_t)B)) || (st = dalloc(I, &d->lastmax, (size_t)B)) ||
      (st = dalloc(I, &d->stage, (size_t)B)) || (st = dalloc(I, &d->active, (size_t)B)) ||
      (st = dalloc(I, &d->stall, (size_t)B)) || (st = dalloc(I, &d->iters, (size_t)B)) ||
      (st = dalloc(I, &d->justf, (size_t)B)) || (st = dalloc(I, &d->remaining, 1)))
    return st;
  d->B = B;
  return PVGLU_OK;
}

inline dim3 grid_for(long count) { return dim3((unsigned)((count + 255) / 256)); }

// eval + stamps on stream s
// Companion models of the instances in `active` (all when null).
void launch_eval(pvglu_pv *pv, int B, const int *active) {
  NewtonDev *d = pv->nd;
  const RowGrid ge = row_grid(pv->nmod, B);
  hipLaunchKernelGGL(k_eval, ge.grid, ge.block, 0, d->s, pv->nmod, B, d->G,
                     d->rsh, d->lk, d->v, d->geq, d->ieq, active);
}

// Stamping of A and b from the companion models; eval: evaluate them first.
int launch_stamp(pvglu_pv *pv, int B, const int *active, bool eval) {
  NewtonDev *d = pv->nd;
  const int n = pv->n, nnz = pv->plan->nnz;
  const RowGrid ga = stamp_grid(nnz, B), gb = stamp_grid(n, B);
  if (eval)
    launch_eval(pv, B, active);
  hipLaunchKernelGGL(k_stamp_a, ga.grid, ga.block, 0, d->s, nnz, B, d->base_a,
                     d->a_ptr, d->a_mod, d->a_sgn, d->geq, d->gvs, d->A, active);
  hipLaunchKernelGGL(k_stamp_b, gb.grid, gb.block, 0, d->s, n, B, d->base_b,
                     d->b_ptr, d->b_mod, d->b_sgn, d->ieq, d->b, active);
  HIP_TRY(hipGetLastError());
  return PVGLU_OK;
}

// Stall rule of the last stage: max |delta| <= noise for kStallMax
// iterations.  The noise level is the FP64 floor of each form: the direct
// form's iterates carry the solve's rounding of the ~600 V node voltages
// (up to ~1e-9 V behind weak modules), the correction form's only the
// rounding of the branch currents (~1e-12 V).
constexpr double kNoiseDirect = 1e-8;
constexpr double kNoiseCorrection = 1e-11;
constexpr int kStallMax = 20;
inline double noise_of(int form) {
  return form == PVGLU_NEWTON_DIRECT ? kNoiseDirect : kNoiseCorrection;
}

// Stamping on load for a batch of B (correction form, switch on): the LU
// levels that run as level launches stamp their string segments with a
// record; the other A slots are listed for k_stamp_a_list.  Called before
// capture (it uploads the list).
int prepare_stamping(pvglu_pv *pv, int B) {
  NewtonDev *d = pv->nd;
  d->stamped = false;
  d->res_on_load = false;
  d->nres = 0;
  d->npart = 0;
  if (!pv->fused_stamping || d->form != PVGLU_NEWTON_CORRECTION)
    return PVGLU_OK;
  pvglu_plan *P = pv->plan;
  std::vector<int> levels;
  int st = stamped_lu_levels(P, B, levels);
  if (st)
    return st;
  if (levels.empty())
    return PVGLU_OK;
  std::vector<char> covered(P->nnz, 0);
  std::vector<int> rrow(pv->n, 0);
  bool anyres = false;
  const KernelLevels &K = P->k_lu;
  for (int lv : levels)
    for (int ch = K.level_chain_ptr[lv]; ch < K.level_chain_ptr[lv + 1]; ++ch)
      for (int t = K.chain_col_ptr[ch]; t < K.chain_col_ptr[ch + 1]; ++t) {
        const int *r = &d->h_stamp_seg[(size_t)t * kStampRec];
        if (!r[0])
          continue;
        const int len = P->fast_lu_seg[(size_t)t * 4];
        const int dslot = P->fast_lu_seg[(size_t)t * 4 + 1];
        const int dcol = P->fast_lu_seg[(size_t)t * 4 + 2];
        const int slot0 = P->fast_lu_rec[(size_t)t * 16];
        const int col0 = P->fast_lu_rec[(size_t)t * 16 + 1];
        for (int i = 0; i < len; ++i) {
          for (int k = 0; k < 3; ++k)
            covered[slot0 + i * dslot + k] = 1;
          if (r[6] & 1) {
            rrow[col0 + i * dcol] = 1;
            anyres = true;
          }
        }
      }
  if (anyres) {
    std::vector<int> rest;
    for (int r = 0; r < pv->n; ++r)
      if (!rrow[r])
        rest.push_back(r);
    if (!rest.empty())
      HIP_TRY(hipMemcpy(d->res_list, rest.data(), sizeof(int) * rest.size(),
                        hipMemcpyHostToDevice));
    d->nres = (int)rest.size();
  }
  d->res_on_load = anyres;
  std::vector<int> part;
  for (int q = 0; q < P->nnz; ++q)
    if (!covered[q])
      part.push_back(q);
  if (!part.empty())
    HIP_TRY(hipMemcpy(d->a_part, part.data(), sizeof(int) * part.size(),
                      hipMemcpyHostToDevice));
  d->npart = (int)part.size();
  d->stamped = true;
  return PVGLU_OK;
}

// One Newton iteration on d->s.  Direct form: stamping of A and b, solve for
// the node voltages x.  Correction form (parity p): stamping of A, the
// branch-current residual at the node voltages X[p], solve for the
// correction dx; the update writes X[p] + dx to X[1 - p].  Then the fused
// update + evaluation and the per-instance control.
int launch_iteration(pvglu_pv *pv, int B, int parity) {
  NewtonDev *d = pv->nd;
  const bool corr = d->form == PVGLU_NEWTON_CORRECTION;
  const int n = pv->n, nnz = pv->plan->nnz;
  // the companion models of this iteration come from the previous
  // iteration's fused update (or the evaluation before the first replay)
  int st;
  const double *xnew = d->x, *dx = nullptr;
  if (!corr) {
    if ((st = launch_stamp(pv, B, d->active, false)))
      return st;
    // frozen instances: A and b unchanged, so their solution stays put; a
    // tile of them is skipped
    if ((st = refactor_solve_masked(pv->plan, d->A, d->LU, d->b, d->x, B, d->s,
                                    d->skip ? d->active : nullptr)))
      return st;
  } else {
    const bool res = d->stamped && d->res_on_load;
    const int nrow = res ? d->nres : n;
    const RowGrid ga = stamp_grid(d->stamped ? d->npart : nnz, B),
                  gb = stamp_grid(nrow, B);
    if (!d->stamped)
      hipLaunchKernelGGL(k_stamp_a, ga.grid, ga.block, 0, d->s, nnz, B,
                         d->base_a, d->a_ptr, d->a_mod, d->a_sgn, d->geq,
                         d->gvs, d->A, d->active);
    else if (d->npart > 0) // the rest is stamped on load by the refactor
      hipLaunchKernelGGL(k_stamp_a_list, ga.grid, ga.block, 0, d->s, d->npart,
                         d->a_part, B, d->base_a, d->a_ptr, d->a_mod,
                         d->a_sgn, d->geq, d->gvs, d->A, d->active);
    // this iteration's node voltages X[parity] were written by the previous
    // update (or zeroed before the first replay)
    if (nrow > 0)
      hipLaunchKernelGGL(k_residual, gb.grid, gb.block, 0, d->s, nrow, B,
                         d->b_ptr, d->b_mod, d->b_sgn, d->l_ptr, d->l_ca,
                         d->l_cb, d->l_g, d->l_e, d->pos, d->neg, d->ieq,
                         d->gvs, d->X[parity], d->b, d->active,
                         res ? d->res_list : nullptr);
    HIP_TRY(hipGetLastError());
    // a tile of frozen instances is skipped (the update does not read their
    // corrections)
    const int *act = d->skip ? d->active : nullptr;
    if (d->stamped) {
      const StampOnLoad so{d->stamp_seg, d->geq, d->gvs,
                           d->res_on_load ? d->ieq : nullptr,
                           d->res_on_load ? d->X[parity] : nullptr,
                           d->res_on_load ? d->active : nullptr};
      if ((st = refactor_solve_stamped(pv->plan, so, d->A, d->LU, d->b, d->dx,
                                       B, d->s, act)))
        return st;
    } else if ((st = refactor_solve_masked(pv->plan, d->A, d->LU, d->b, d->dx,
                                           B, d->s, act))) {
      return st;
    }
    xnew = d->X[parity];
    dx = d->dx;
  }
  int *status = device_status(pv->plan);
  if (!status)
    return fail(PVGLU_E_STATE, "batched Newton: plan has no device state");
  const int nm = pv->nmod, lt = upd_tile_log2(B);
  const int per_block = (256 >> lt) * kUpdPerThread;
  const int nblk = (nm + per_block - 1) / per_block,
            nlb = corr ? (d->nleft + per_block - 1) / per_block : 0;
  // dsum / dmax were cleared by k_newton_init or the previous control kernel
  hipLaunchKernelGGL(corr ? k_update_eval<true> : k_update_eval<false>,
                     dim3(nblk + nlb, (B + (1 << lt) - 1) >> lt),
                     dim3(256), 0, d->s, nm, B, corr ? d->upos : d->pos,
                     corr ? d->uneg : d->neg, xnew, dx, d->v, d->dsum, d->dmax,
                     d->active, status, d->G, d->rsh, d->lk, d->geq, d->ieq,
                     d->remaining, corr ? d->X[1 - parity] : nullptr, d->justf,
                     nblk, d->left, d->nleft);
  const double *ladder = d->ladder + (12 - d->nstages);
  hipLaunchKernelGGL(k_newton_control, grid_for(B), dim3(256), 0, d->s, B,
                     d->nstages, ladder, 1e-6 * pv->n, d->tol,
                     noise_of(d->form), kStallMax, d->dsum, d->dmax, d->stage,
                     d->gvs, d->active, d->stall, d->iters, d->lastmax,
                     d->remaining, corr ? d->justf : nullptr);
  HIP_TRY(hipGetLastError());
  return PVGLU_OK;
}

int capture_one(pvglu_pv *pv, int B, int parity, hipGraph_t *graph,
                hipGraphExec_t *exec) {
  NewtonDev *d = pv->nd;
  HIP_TRY(hipStreamBeginCapture(d->s, hipStreamCaptureModeThreadLocal));
  int st = launch_iteration(pv, B, parity);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(d->s, &g);
  if (st) {
    if (g)
      (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess)
    return fail(PVGLU_E_HIP, "graph capture: %s", hipGetErrorString(e));
  *graph = g;
  HIP_TRY(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
  return PVGLU_OK;
}

int build_graph(pvglu_pv *pv, int B) {
  NewtonDev *d = pv->nd;
  if (d->exec)
    return PVGLU_OK;
  int st = prepare_stamping(pv, B);
  if (st)
    return st;
  st = capture_one(pv, B, 0, &d->graph, &d->exec);
  if (!st && d->form == PVGLU_NEWTON_CORRECTION)
    st = capture_one(pv, B, 1, &d->graph2, &d->exec2);
  if (st)
    drop_graphs(d);
  return st;
}

} // namespace

// --------------------------------------------------- batch compaction
//
// Instances finish their Newton solves at different iterations (the cfg2
// cold start: 248 to 484 iterations, median 375).  When at a poll at most
// three quarters of the working batch is still active, the finished
// instances' results go to their columns of the caller's arrays (the "home"
// set: G, v, x, iters, lastmax at batch B) and the active ones are gathered
// into a denser working batch in one of two alternate sets (in the
// correction form the node voltages X[parity] hold every instance's current
// iterate, or its final voltages once frozen).  The graphs are captured
// again for the smaller batch.  Every kernel's
// arithmetic per instance is independent of the batch width and of the
// other instances, so results are bit-identical with or without compaction.

namespace {

using InstSet = NewtonDev::InstSet;

__global__ void k_fill_i32(int n, int *p, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = v;
}

InstSet work_set(const NewtonDev *d) {
  InstSet w;
  w.G = d->G, w.v = d->v, w.geq = d->geq, w.ieq = d->ieq;
  w.X0 = d->X[0], w.X1 = d->X[1], w.dx = d->dx, w.x = d->x;
  w.gvs = d->gvs, w.lastmax = d->lastmax;
  w.stage = d->stage, w.stall = d->stall, w.iters = d->iters;
  return w;
}

void set_work(NewtonDev *d, const InstSet &w) {
  d->G = w.G, d->v = w.v, d->geq = w.geq, d->ieq = w.ieq;
  d->X[0] = w.X0, d->X[1] = w.X1, d->dx = w.dx, d->x = w.x;
  d->gvs = w.gvs, d->lastmax = w.lastmax;
  d->stage = w.stage, d->stall = w.stall, d->iters = w.iters;
}

int alloc_alt(pvglu_pv *pv) {
  NewtonDev *d = pv->nd;
  if (d->cidx)
    return PVGLU_OK;
  const size_t nm = pv->nmod, n = pv->n, B = d->B;
  auto &I = d->alt_allocs;
  int st;
  for (InstSet &a : d->alt)
    if ((st = dalloc(I, &a.G, nm * B)) || (st = dalloc(I, &a.v, nm * B)) ||
        (st = dalloc(I, &a.geq, nm * B)) || (st = dalloc(I, &a.ieq, nm * B)) ||
        (st = dalloc(I, &a.X0, n * B)) || (st = dalloc(I, &a.X1, n * B)) ||
        (st = dalloc(I, &a.dx, n * B)) || (st = dalloc(I, &a.x, n * B)) ||
        (st = dalloc(I, &a.gvs, B)) || (st = dalloc(I, &a.lastmax, B)) ||
        (st = dalloc(I, &a.stage, B)) || (st = dalloc(I, &a.stall, B)) ||
        (st = dalloc(I, &a.iters, B)))
      return st;
  return dalloc(I, &d->cidx, 3 * B);
}

template <typename T>
void gather(hipStream_t s, long rows, int Bc, int Bn, const int *sel,
            const T *in, T *out) {
  const long cnt = rows * Bn;
  if (cnt > 0)
    hipLaunchKernelGGL(k_gather<T>, dim3((unsigned)((cnt + 255) / 256)),
                       dim3(256), 0, s, rows, Bc, Bn, sel, in, out);
}

template <typename T>
void scatter(hipStream_t s, long rows, int Bc, int Bh, int cnt, const int *from,
             const int *to, const T *in, T *out) {
  const long tot = rows * cnt;
  if (tot > 0)
    hipLaunchKernelGGL(k_scatter<T>, dim3((unsigned)((tot + 255) / 256)),
                       dim3(256), 0, s, rows, Bc, Bh, cnt, from, to, in, out);
}

struct Compaction {
  pvglu_pv *pv;
  InstSet home; // the caller's batch, in the caller's instance order
  int B, Bc, next = 0;
  std::vector<int> map; // working instance -> caller instance
  bool active = false;  // the working batch is not the home set
  Compaction(pvglu_pv *pv, int B)
      : pv(pv), home(work_set(pv->nd)), B(B), Bc(B), map(B) {
    for (int k = 0; k < B; ++k)
      map[k] = k;
  }
  // on every exit: the home set back in place (after finish() it holds every
  // instance's result; on an error path it is incomplete) and the graphs of
  // the compacted batch dropped
  ~Compaction() {
    if (active) {
      set_work(pv->nd, home);
      drop_graphs(pv->nd);
    }
  }
  int caller_instance(int k) const { return k >= 0 && k < Bc ? map[k] : k; }

  // Results of the working instances from[0..cnt) (device list) to the home
  // columns to[0..cnt).  parity: the node-voltage buffer the next replay
  // would read.
  void results_home(int cnt, const int *from, const int *to, int parity,
                    bool corr) {
    NewtonDev *d = pv->nd;
    const InstSet w = work_set(d);
    scatter<double>(d->s, pv->nmod, Bc, B, cnt, from, to, w.v, home.v);
    scatter<double>(d->s, pv->n, Bc, B, cnt, from, to,
                    corr ? (parity ? w.X1 : w.X0) : w.x, home.x);
    scatter<int>(d->s, 1, Bc, B, cnt, from, to, w.iters, home.iters);
    scatter<double>(d->s, 1, Bc, B, cnt, from, to, w.lastmax, home.lastmax);
  }

  int compact(int parity, bool corr) {
    NewtonDev *d = pv->nd;
    int code = 0, col = 0, inst = 0, st;
    if ((st = pvglu_status(pv->plan, &code, &col, &inst)))
      return st;
    if (code) // stop at the first recorded error
      return fail(code, "batched Newton: status %d at column/element %d, instance %d",
                  code, col, caller_instance(inst));
    std::vector<int> act(Bc);
    HIP_TRY(hipMemcpy(act.data(), d->active, sizeof(int) * Bc,
                      hipMemcpyDeviceToHost));
    std::vector<int> idx, from, to; // idx: survivors, then leavers
    for (int k = 0; k < Bc; ++k)
      if (act[k])
        idx.push_back(k);
    const int Bn = (int)idx.size(), L = Bc - Bn;
    if (Bn == 0 || L == 0)
      return PVGLU_OK;
    for (int k = 0; k < Bc; ++k)
      if (!act[k]) {
        from.push_back(k);
        to.push_back(map[k]);
      }
    idx.insert(idx.end(), from.begin(), from.end());
    idx.insert(idx.end(), to.begin(), to.end());
    if ((st = alloc_alt(pv)))
      return st;
    HIP_TRY(hipMemcpy(d->cidx, idx.data(), sizeof(int) * idx.size(),
                      hipMemcpyHostToDevice));
    const int *sel = d->cidx, *dfrom = d->cidx + Bn, *dto = d->cidx + Bn + L;
    results_home(L, dfrom, dto, parity, corr);
    const InstSet w = work_set(d);
    InstSet nx = d->alt[next];
    const long nm = pv->nmod, n = pv->n;
    gather<double>(d->s, nm, Bc, Bn, sel, w.G, nx.G);
    gather<double>(d->s, nm, Bc, Bn, sel, w.v, nx.v);
    gather<double>(d->s, nm, Bc, Bn, sel, w.geq, nx.geq);
    gather<double>(d->s, nm, Bc, Bn, sel, w.ieq, nx.ieq);
    if (corr) // the correction is formed and used within a replay
      gather<double>(d->s, n, Bc, Bn, sel, parity ? w.X1 : w.X0,
                     parity ? nx.X1 : nx.X0);
    // direct form: the next replay's solve writes every x
    gather<double>(d->s, 1, Bc, Bn, sel, w.gvs, nx.gvs);
    gather<double>(d->s, 1, Bc, Bn, sel, w.lastmax, nx.lastmax);
    gather<int>(d->s, 1, Bc, Bn, sel, w.stage, nx.stage);
    gather<int>(d->s, 1, Bc, Bn, sel, w.stall, nx.stall);
    gather<int>(d->s, 1, Bc, Bn, sel, w.iters, nx.iters);
    hipLaunchKernelGGL(k_fill_i32, grid_for(Bn), dim3(256), 0, d->s, Bn,
                       d->active, 1);
    hipLaunchKernelGGL(k_fill_i32, grid_for(Bn), dim3(256), 0, d->s, Bn,
                       d->justf, 0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(d->s));
    // the home set's x keeps the caller's results; the working set brings
    // its own buffer (written by the direct form's solves)
    set_work(d, nx);
    next ^= 1;
    std::vector<int> nmap(Bn);
    for (int j = 0; j < Bn; ++j)
      nmap[j] = map[idx[j]];
    map.swap(nmap);
    Bc = Bn;
    active = true;
    drop_graphs(d);
    return build_graph(pv, Bc);
  }

  // every working instance's results home; the home set back in place
  int finish(int parity, bool corr) {
    NewtonDev *d = pv->nd;
    std::vector<int> idx(2 * Bc);
    for (int k = 0; k < Bc; ++k) {
      idx[k] = k;
      idx[Bc + k] = map[k];
    }
    HIP_TRY(hipMemcpy(d->cidx, idx.data(), sizeof(int) * idx.size(),
                      hipMemcpyHostToDevice));
    results_home(Bc, d->cidx, d->cidx + Bc, parity, corr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(d->s));
    set_work(d, home);
    drop_graphs(d);
    active = false;
    return PVGLU_OK;
  }
};

} // namespace

// ------------------------------------------------------------- batched API

static int batched_solve(pvglu_pv *pv, int B, const double *g, double *v_module,
                         double tol, int max_iters, int ladder,
                         double *stats) {
  int st = newton_setup(pv);
  if (st)
    return st;
  if ((st = newton_alloc_instances(pv, B)))
    return st;
  NewtonDev *d = pv->nd;
  const size_t nm = pv->nmod;
  // pick the kernel schedule before capture (autotune allocates + syncs)
  if (pv->plan->tuned_batch != B && pv->plan->sched_width == 0)
    if ((st = pvglu_autotune(pv->plan, B, 3, nullptr)))
      return st;
  HIP_TRY(hipMemcpyAsync(d->G, g, sizeof(double) * nm * B,
                         hipMemcpyHostToDevice, d->s));
  HIP_TRY(hipMemcpyAsync(d->v, v_module, sizeof(double) * nm * B,
                         hipMemcpyHostToDevice, d->s));
  if ((st = pvglu_status_reset(pv->plan, d->s)))
    return st;
  const int nst = ladder ? 12 : 1;
  if (d->exec && (d->nstages != nst || d->tol != tol || d->form != pv->form ||
                  d->skip != pv->skip_frozen || d->fused != pv->fused_stamping))
    drop_graphs(d);
  d->nstages = nst;
  d->tol = tol;
  d->form = pv->form;
  d->skip = pv->skip_frozen;
  d->fused = pv->fused_stamping;
  const bool corr = d->form == PVGLU_NEWTON_CORRECTION;
  hipLaunchKernelGGL(k_newton_init, grid_for(B), dim3(256), 0, d->s, B,
                     kLadder[12 - nst], d->stage, d->gvs, d->active, d->stall,
                     d->iters, d->lastmax, d->dsum, d->dmax, d->justf);
  if (corr) {
    // node voltages start at 0: the first iteration's right-hand side is
    // the companion b (its correction x = A^-1 b)
    const long cnt = (long)pv->n * B;
    hipLaunchKernelGGL(k_zero, grid_for(cnt), dim3(256), 0, d->s, cnt, d->X[0]);
  }
  launch_eval(pv, B, d->active); // the first replay's companion models
  HIP_TRY(hipGetLastError());
  if ((st = build_graph(pv, B)))
    return st;
  // The host polls the remaining-instance counter after every kPoll graph
  // replays: a replay after an instance has converged leaves it unchanged
  // (frozen: no evaluation, stamping or update; in the direct form the
  // refactor + solve of its unchanged system reproduces its solution bit for
  // bit, in the correction form its node voltages stay in both buffers and
  // its right-hand side is zero), so the extra replays of a burst cost time
  // only, and the GPU no longer idles for a host round trip after each
  // iteration.  The poll is pipelined: the next burst is queued before the
  // host waits for the previous one's count, so the GPU does not idle for
  // the round trip of a poll either (a burst queued behind the one that
  // finished every instance changes nothing; its replays are not counted).
  // The correction form alternates its two graphs (replay j reads
  // X[j % 2]).  With skip_frozen, a poll that finds at most three quarters
  // of the working batch active compacts it (Compaction above), after the
  // burst behind it.
  constexpr int kPoll = 4;
  int iters = 0, remaining = 1, parity = 0;
  Compaction cm(pv, B);
  auto launch_burst = [&](int slot) -> int {
    const int burst = std::min(kPoll, max_iters - iters);
    for (int k = 0; k < burst; ++k) {
      HIP_TRY(hipGraphLaunch(parity ? d->exec2 : d->exec, d->s));
      if (corr)
        parity ^= 1;
    }
    iters += burst;
    HIP_TRY(hipMemcpyAsync(&d->hrem[slot], d->remaining, sizeof(int),
                           hipMemcpyDeviceToHost, d->s));
    HIP_TRY(hipEventRecord(d->ev[slot], d->s));
    return PVGLU_OK;
  };
  auto t0 = std::chrono::steady_clock::now();
  int slot = 0;
  if ((st = launch_burst(slot)))
    return st;
  for (;;) {
    int ahead = 0; // replays queued behind the polled burst
    if (iters < max_iters) {
      const int before = iters;
      if ((st = launch_burst(slot ^ 1)))
        return st;
      ahead = iters - before;
    }
    HIP_TRY(hipEventSynchronize(d->ev[slot]));
    remaining = d->hrem[slot];
    const bool compact =
        pv->skip_frozen && remaining > 0 && 4 * remaining <= 3 * cm.Bc;
    if (remaining > 0 && ahead > 0 && !compact) {
      slot ^= 1; // poll the burst behind next
      continue;
    }
    HIP_TRY(hipStreamSynchronize(d->s));
    if (ahead > 0) {
      if (remaining == 0)
        iters -= ahead; // replays of finished instances only
      else
        remaining = d->hrem[slot ^ 1];
    }
    if (remaining == 0)
      break;
    if (iters >= max_iters) {
      // hand back the current iterate for diagnosis
      if (cm.active && (st = cm.finish(parity, corr)))
        return st;
      (void)hipMemcpy(v_module, d->v, sizeof(double) * nm * B,
                      hipMemcpyDeviceToHost);
      return fail(PVGLU_E_NOCONV, "batched Newton: %d iterations, %d instances left",
                  iters, remaining);
    }
    if (pv->skip_frozen && 4 * remaining <= 3 * cm.Bc)
      if ((st = cm.compact(parity, corr)))
        return st;
    slot = 0;
    if ((st = launch_burst(slot)))
      return st;
  }
  if (cm.active) {
    if ((st = cm.finish(parity, corr)))
      return st;
  } else if (corr) {
    // X[parity] holds every instance's final node voltages: the update of
    // the replay that froze it wrote them there, and the update of the next
    // replay, if any, copied them to the other buffer
    HIP_TRY(hipMemcpyAsync(d->x, d->X[parity], sizeof(double) * pv->n * B,
                           hipMemcpyDeviceToDevice, d->s));
    HIP_TRY(hipStreamSynchronize(d->s));
  }
  auto t1 = std::chrono::steady_clock::now();
  int code = 0, col = 0, inst = 0;
  if ((st = pvglu_status(pv->plan, &code, &col, &inst)))
    return st;
  if (code) // recorded in the last working batch
    return fail(code, "batched Newton: status %d at column/element %d, instance %d",
                code, col, cm.caller_instance(inst));
  HIP_TRY(hipMemcpy(v_module, d->v, sizeof(double) * nm * B,
                    hipMemcpyDeviceToHost));
  if (stats) {
    std::vector<int> it(B);
    std::vector<double> lm(B);
    HIP_TRY(hipMemcpy(it.data(), d->iters, sizeof(int) * B, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(lm.data(), d->lastmax, sizeof(double) * B,
                      hipMemcpyDeviceToHost));
    const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    stats[0] = *std::max_element(it.begin(), it.end());
    stats[1] = ms;
    stats[2] = ms / std::max(iters, 1);
    stats[3] = *std::max_element(lm.begin(), lm.end());
  }
  return PVGLU_OK;
}

static int fetch_solution(pvglu_pv *pv, int B, double *x_out,
                          double *power_out) {
  NewtonDev *d = pv->nd;
  const int n = pv->n;
  std::vector<double> xp((size_t)n * B);
  HIP_TRY(hipMemcpy(xp.data(), d->x, sizeof(double) * xp.size(),
                    hipMemcpyDeviceToHost));
  if (x_out)
    for (int i = 0; i < n; i++)
      memcpy(x_out + (size_t)i * B, xp.data() + (size_t)pv->Pinv[i] * B,
             sizeof(double) * B);
  if (power_out)
    for (int k = 0; k < B; k++) {
      double p = 0;
      for (int bidx : pv->bus) {
        const double V = xp[(size_t)pv->Pinv[bidx] * B + k];
        p += V * (kNortonG * (V - kNortonV));
      }
      power_out[k] = p;
    }
  return PVGLU_OK;
}

extern "C" int pvglu_pv_newton_batched(pvglu_pv *pv, int batch, const double *g,
                                       double *v_module, double *x_out,
                                       double *power_out, double tol,
                                       int max_iters, int ladder,
                                       double *stats) {
  begin_call();
  if (!pv || batch <= 0 || !g || !v_module)
    return fail(PVGLU_E_ARG, "newton_batched: bad arguments");
  int st = batched_solve(pv, batch, g, v_module, tol,
                         max_iters > 0 ? max_iters : 5000, ladder, stats);
  if (st)
    return st;
  return fetch_solution(pv, batch, x_out, power_out);
}

extern "C" int pvglu_pv_gradients(pvglu_pv *pv, int batch, const double *g,
                                  double *v_module, double *power_out,
                                  double *dP_dG, double *dP_dgc, double tol,
                                  int max_iters) {
  begin_call();
  if (!pv || batch <= 0 || !g || !v_module)
    return fail(PVGLU_E_ARG, "gradients: bad arguments");
  int st = batched_solve(pv, batch, g, v_module, tol,
                         max_iters > 0 ? max_iters : 5000, 0, nullptr);
  if (st)
    return st;
  NewtonDev *d = pv->nd;
  const int B = batch, n = pv->n, nm = pv->nmod, nc = pv->ncab;
  // Jacobian at the converged element voltages: stamp + refactor (the
  // companion matrix at V_e = V_branch(x) is dF/dx); x is the Newton's
  // converged node voltages
  if ((st = launch_stamp(pv, B, nullptr, true)))
    return st;
  if ((st = pvglu_refactor(pv->plan, d->A, d->LU, B, d->s)))
    return st;
  hipLaunchKernelGGL(k_zero, grid_for((long)n * B), dim3(256), 0, d->s,
                     (long)n * B, d->gP);
  const RowGrid gbus = row_grid(pv->I, B), gm = row_grid(nm, B),
                gc = row_grid(nc, B);
  hipLaunchKernelGGL(k_power_grad_bus, gbus.grid, gbus.block, 0, d->s, pv->I, B,
                     d->bus_c, d->x, d->gP);
  if ((st = pvglu_solve_transpose(pv->plan, d->LU, d->gP, d->mu, B, d->s)))
    return st;
  hipLaunchKernelGGL(k_grad_modules, gm.grid, gm.block, 0, d->s, nm, B, d->G,
                     d->rsh, d->v, d->rpos, d->rneg, d->mu, d->dG);
  hipLaunchKernelGGL(k_grad_cables, gc.grid, gc.block, 0, d->s, nc, B,
                     d->cab_ra, d->cab_rb, d->cab_ca, d->cab_cb, d->x, d->mu,
                     d->dgc);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(d->s));
  int code = 0, col = 0, inst = 0;
  if ((st = pvglu_status(pv->plan, &code, &col, &inst)))
    return st;
  if (code)
    return fail(code, "gradients: status %d", code);
  if (dP_dG)
    HIP_TRY(hipMemcpy(dP_dG, d->dG, sizeof(double) * nm * B,
                      hipMemcpyDeviceToHost));
  if (dP_dgc)
    HIP_TRY(hipMemcpy(dP_dgc, d->dgc, sizeof(double) * nc * B,
                      hipMemcpyDeviceToHost));
  return fetch_solution(pv, B, nullptr, power_out);
}

// ------------------------------------------------------------- drop-in API

namespace pvglu {

int dropin_newton(Circuit &c, pvglu_plan *P, const std::vector<int> &Pv,
                  const std::vector<int> &Pinv, const std::vector<int> &Prow,
                  DropinDev *&dd, int form, double tol, int ladder,
                  int fixed_iters, double *x_out, double *stats) {
  (void)Pv;
  const int n = c.unknowns(), nnz = P->nnz, nm = c.nonlinear_count();
  if (!dd) {
    auto *d = new DropinDev();
    dd = d;
    HIP_TRY(hipMalloc(&d->a, sizeof(double) * nnz));
    HIP_TRY(hipMalloc(&d->lu, sizeof(double) * nnz));
    HIP_TRY(hipMalloc(&d->b, sizeof(double) * n));
    HIP_TRY(hipMalloc(&d->x, sizeof(double) * n));
    HIP_TRY(hipMalloc(&d->v, sizeof(double) * std::max(nm, 1)));
    HIP_TRY(hipMalloc(&d->dsum, sizeof(double)));
    HIP_TRY(hipMalloc(&d->dmax, sizeof(double)));
    HIP_TRY(hipMalloc(&d->pos, sizeof(int) * std::max(nm, 1)));
    HIP_TRY(hipMalloc(&d->neg, sizeof(int) * std::max(nm, 1)));
    // element terminals as permuted unknowns (-1: ground)
    std::vector<int> pos(nm), neg(nm);
    for (int e = 0; e < nm; e++) {
      const auto t = c.nonlinear_nodes(e);
      pos[e] = t.first ? Pinv[c.x_of(t.first)] : -1;
      neg[e] = t.second ? Pinv[c.x_of(t.second)] : -1;
    }
    if (nm) {
      HIP_TRY(hipMemcpy(d->pos, pos.data(), sizeof(int) * nm, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(d->neg, neg.data(), sizeof(int) * nm, hipMemcpyHostToDevice));
    }
  }
  DropinDev *d = dd;
  if (nm)
    HIP_TRY(hipMemcpy(d->v, c.v_iterate.data(), sizeof(double) * nm,
                      hipMemcpyHostToDevice));
  int st;
  if ((st = pvglu_status_reset(P, nullptr)))
    return st;
  const bool corr = form == PVGLU_NEWTON_CORRECTION;
  const bool inv = c.inverter_count() > 0;
  // correction form: unknowns of the last iteration (circuit order; 0 at the
  // start, so the first right-hand side is b) and their permuted copy
  std::vector<double> vals(nnz), bvec, tmp(n), xprev(n, 0.0), xcol(n), xu(n);
  double t_host = 0, t_dev = 0;
  auto T0 = std::chrono::steady_clock::now();
  int it = 0;
  const int nst = (ladder && fixed_iters <= 0) ? 12 : 1;
  const double *stages = (ladder && fixed_iters <= 0) ? kLadder : kLadder + 11;
  for (int si = 0; si < nst; si++) {
    const bool last = si == nst - 1;
    long double diff = 1e10;
    int stage_it = 0, stall = 0;
    for (;;) {
      if (fixed_iters > 0) {
        if (stage_it >= fixed_iters)
          break;
      } else if (!last && diff < 1e-6L * n) {
        break;
      }
      if (++it > 5000)
        return fail(PVGLU_E_NOCONV, "Did not converge"); // 000:179-181
      stage_it++;
      auto h0 = std::chrono::steady_clock::now();
      // create_matrix_and_vector + permutations + CSC value fill (000:185-210)
      SparseMap A;
      c.create_matrix_and_vector(stages[si], A, bvec);
      if (corr)
        c.residual(stages[si], xprev, bvec);
      A = A.interchange_row_columns(Pinv.data());
      A = A.exchange_rows(Prow.data());
      for (int i = 0; i < n; i++)
        tmp[Pinv[i]] = bvec[i];
      for (int i = 0; i < n; i++)
        bvec[Prow[i]] = tmp[i];
      for (int col = 0; col < n; col++)
        for (int q = P->cp[col]; q < P->cp[col + 1]; q++)
          vals[q] = A.D(P->ri[q], col); // fill_values_CSC_only (010:142-155)
      auto h1 = std::chrono::steady_clock::now();
      HIP_TRY(hipMemcpy(d->a, vals.data(), sizeof(double) * nnz, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(d->b, bvec.data(), sizeof(double) * n, hipMemcpyHostToDevice));
      if ((st = pvglu_refactor(P, d->a, d->lu, 1, nullptr)) ||
          (st = pvglu_solve(P, d->lu, d->b, d->x, 1, nullptr)))
        return st;
      if (corr || inv)
        HIP_TRY(hipMemcpy(tmp.data(), d->x, sizeof(double) * n,
                          hipMemcpyDeviceToHost));
      if (corr) {
        // x = x_prev + d, in circuit and permuted order
        for (int i = 0; i < n; i++) {
          xprev[i] = xprev[i] + tmp[Pinv[i]];
          xcol[Pinv[i]] = xprev[i];
        }
        HIP_TRY(hipMemcpy(d->x, xcol.data(), sizeof(double) * n,
                          hipMemcpyHostToDevice));
      }
      double dsum = 0, dmax = 0;
      if (nm) {
        if ((st = pvglu_newton_update(P, nm, 1, d->pos, d->neg, d->x, d->v, d->dsum,
                                      d->dmax, nullptr, nullptr)))
          return st;
        HIP_TRY(hipMemcpy(c.v_iterate.data(), d->v, sizeof(double) * nm,
                          hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&dsum, d->dsum, sizeof(double), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&dmax, d->dmax, sizeof(double), hipMemcpyDeviceToHost));
      }
      if (inv) {
        // the inverters' (V2, I1) from this solution (011:88-109)
        if (corr)
          xu = xprev;
        else
          for (int i = 0; i < n; i++)
            xu[i] = tmp[Pinv[i]];
        double imax = 0.0;
        dsum += c.refresh_inverters(xu, imax);
        dmax = std::fmax(dmax, imax);
      }
      auto h2 = std::chrono::steady_clock::now();
      t_host += std::chrono::duration<double, std::milli>(h1 - h0).count();
      t_dev += std::chrono::duration<double, std::milli>(h2 - h1).count();
      diff = dsum;
      if (std::isnan(dsum))
        return fail(PVGLU_E_NOCONV, "drop-in Newton: NaN update");
      if (fixed_iters <= 0 && last) {
        if (dmax <= tol)
          break;
        stall = dmax <= noise_of(form) ? stall + 1 : 0; // FP64 noise floor
        if (stall >= kStallMax)
          break;
      }
    }
  }
  int code = 0, col = 0, inst = 0;
  if ((st = pvglu_status(P, &code, &col, &inst)))
    return st;
  if (code)
    return fail(code, "drop-in Newton: status %d at %d", code, col);
  if (x_out) {
    HIP_TRY(hipMemcpy(tmp.data(), d->x, sizeof(double) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++)
      x_out[i] = tmp[Pinv[i]]; // apply_permutation(x, P) (000:233)
  }
  auto T1 = std::chrono::steady_clock::now();
  if (stats) {
    stats[0] = it;
    stats[1] = t_host;
    stats[2] = t_dev;
    stats[3] = std::chrono::duration<double, std::milli>(T1 - T0).count();
  }
  return PVGLU_OK;
}

} // namespace pvglu

extern "C" int pvglu_pv_newton_dropin(pvglu_pv *pv, const double *g,
                                      double *v_module, double *x_out,
                                      double tol, int ladder, int fixed_iters,
                                      double *stats) {
  begin_call();
  if (!pv || !g || !v_module)
    return fail(PVGLU_E_ARG, "newton_dropin: bad arguments");
  int st = pv_prepare(pv);
  if (st)
    return st;
  const int nm = pv->nmod;
  for (int e = 0; e < nm; e++)
    pv->mod_iph[e] = kIphPerIrradiance * g[e];
  pv->circuit.v_iterate.assign(v_module, v_module + nm);
  if ((st = dropin_newton(pv->circuit, pv->plan, pv->P, pv->Pinv, pv->Prow,
                          pv->dd, pv->form, tol, ladder, fixed_iters, x_out,
                          stats)))
    return st;
  memcpy(v_module, pv->circuit.v_iterate.data(), sizeof(double) * nm);
  return PVGLU_OK;
}

// ------------------------------------------------------ stamping entry points

extern "C" pvglu_plan *pvglu_pv_plan(pvglu_pv *pv) {
  begin_call();
  if (!pv) {
    fail(PVGLU_E_ARG, "pv_plan: null");
    return nullptr;
  }
  if (pv_prepare(pv))
    return nullptr;
  return pv->plan;
}

extern "C" int pvglu_pv_stamp(pvglu_pv *pv, int batch, const double *g,
                              const double *v, double g_virtual, double *a,
                              double *b, void *stream) {
  begin_call();
  if (!pv || batch <= 0 || !g || !v || !a || !b)
    return fail(PVGLU_E_ARG, "pv_stamp: bad arguments");
  int st = newton_setup(pv);
  if (st)
    return st;
  NewtonDev *d = pv->nd;
  const int nm = pv->nmod, n = pv->n, nnz = pv->plan->nnz, B = batch;
  hipStream_t s = (hipStream_t)stream;
  // element companion values in caller-independent scratch sized for B
  double *geq = nullptr, *ieq = nullptr;
  if (d->B >= B) {
    geq = d->geq;
    ieq = d->ieq;
  } else {
    if ((st = newton_alloc_instances(pv, B)))
      return st;
    geq = d->geq;
    ieq = d->ieq;
  }
  hipLaunchKernelGGL(k_fill_double, grid_for(B), dim3(256), 0, s, d->gvs, B,
                     g_virtual);
  const RowGrid ge = row_grid(nm, B),
                ga = stamp_grid(nnz, B), gb = stamp_grid(n, B);
  hipLaunchKernelGGL(k_eval, ge.grid, ge.block, 0, s, nm, B, g, d->rsh, d->lk,
                     v, geq, ieq, nullptr);
  hipLaunchKernelGGL(k_stamp_a, ga.grid, ga.block, 0, s, nnz, B, d->base_a,
                     d->a_ptr, d->a_mod, d->a_sgn, geq, d->gvs, a, nullptr);
  hipLaunchKernelGGL(k_stamp_b, gb.grid, gb.block, 0, s, n, B, d->base_b,
                     d->b_ptr, d->b_mod, d->b_sgn, ieq, b, nullptr);
  HIP_TRY(hipGetLastError());
  return PVGLU_OK;
}

extern "C" int pvglu_pv_set_newton_form(pvglu_pv *pv, int form) {
  begin_call();
  if (!pv || (form != PVGLU_NEWTON_CORRECTION && form != PVGLU_NEWTON_DIRECT))
    return fail(PVGLU_E_ARG, "set_newton_form: bad arguments");
  pv->form = form;
  return PVGLU_OK;
}

extern "C" int pvglu_pv_set_skip_frozen(pvglu_pv *pv, int on) {
  begin_call();
  if (!pv)
    return fail(PVGLU_E_ARG, "set_skip_frozen: null");
  pv->skip_frozen = on ? 1 : 0;
  return PVGLU_OK;
}

extern "C" int pvglu_pv_set_stamp_on_load(pvglu_pv *pv, int on) {
  begin_call();
  if (!pv)
    return fail(PVGLU_E_ARG, "set_stamp_on_load: null");
  pv->fused_stamping = on ? 1 : 0;
  return PVGLU_OK;
}

extern "C" int pvglu_pv_get_newton_form(const pvglu_pv *pv) {
  if (!pv)
    return fail(PVGLU_E_ARG, "get_newton_form: null");
  return pv->form;
}

extern "C" int pvglu_pv_newton_info(const pvglu_pv *pv, int batch, int *iters,
                                    double *last_max_delta) {
  begin_call();
  if (!pv || !pv->nd || batch <= 0 || batch > pv->nd->B)
    return fail(PVGLU_E_STATE, "newton_info: no batched solve of this size");
  if (iters)
    HIP_TRY(hipMemcpy(iters, pv->nd->iters, sizeof(int) * batch,
                      hipMemcpyDeviceToHost));
  if (last_max_delta)
    HIP_TRY(hipMemcpy(last_max_delta, pv->nd->lastmax, sizeof(double) * batch,
                      hipMemcpyDeviceToHost));
  return PVGLU_OK;
}
