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
  // Optional per-instance activity flags (null: every instance runs).  A
  // W-instance tile none of whose instances is active skips the launch's
  // work: the batched Newton's frozen instances, whose system and solution
  // no longer change.
  const int *act;
  // Stamping on load (OP_REFACTOR_STAMPED launches only, null otherwise):
  // per kernel position t of the LU levels, kStampRec ints describing how
  // the A values of the series-string segment starting there follow from
  // the module conductances geq [module][instance] and the homotopy
  // conductance gvs[instance] (see tri_lu_segment_stamped).
  const int *stamp_seg;
  const double *geq;
  const double *gvs;
  // Correction-form right-hand side formed on load (records with sr[6] bit
  // 0): the module branch currents ieq [module][instance], the node voltages
  // xk [row][instance] of the iteration (homotopy stages) and the batched
  // Newton's instance activity flags (a frozen instance's right-hand side is
  // zero).  Null when the launch has no such records.
  const double *ieq;
  const double *xk;
  const int *live;
};

// Ints per kernel position of a stamping record: [0] 1 when the segment's A
// values are stamped on load; [1] module of the U-above entry of the first
// column, [2] its step per column; [3] module of the L-below entry, [4] its
// step; [5] bits: 0/1/2/3 sign of the U-above / L-below / first diagonal /
// second diagonal term negative, 4 the diagonal's first term is the U-above
// module.  The entries' linear parts are zero.  [6] the correction-form
// right-hand side of the segment's rows formed on load: bit 0 on, bit 1 the
// row's first branch current is the U-above module's, bits 2/3 the first /
// second term's sign negative (the row has exactly these two module
// branches and no linear one, and the modules' other terminals are the
// segment's previous and next nodes); [7] the row of the last column's
// L-below entry (its next node).
constexpr int kStampRec = 8;

struct Lev {
  const int *__restrict__ level_chain_ptr;
  const int *__restrict__ chain_col_ptr;
  const int *__restrict__ cols;
  const int *__restrict__ comp_ptr;
  const int *__restrict__ comp_lv_begin;
  const int *__restrict__ comp_lv_end;
  const int *__restrict__ comp_lv_level;
};

// OP_REFACTOR_STAMPED: OP_REFACTOR_FUSED whose string segments with a
// stamping record form their A values from the module conductances instead
// of loading them (Pat::stamp_seg).
enum Op {
  OP_REFACTOR = 0,
  OP_REFACTOR_FUSED,
  OP_LSOLVE,
  OP_USOLVE,
  OP_UT,
  OP_LT,
  OP_REFACTOR_STAMPED
};

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
