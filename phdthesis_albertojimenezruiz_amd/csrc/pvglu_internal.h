// pvglu_internal.h -- internal types shared by the host symbolic code, the
// PV plant harness and the HIP kernels.
#pragma once

#include "pvglu.h"

#include <cstdint>
#include <string>
#include <vector>

namespace pvglu {

// Star columns (updates only on the diagonal, strided operands) take at most
// this many updates: the star path's register budget in the narrow-run
// kernel, and the plan's limit for flagging a column as a star.
constexpr int kStarMaxUpdates = 32;

// Thread-local error message + status helper.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();
// clear_error() plus discarding a HIP error recorded by an earlier runtime
// call outside this entry point (another library's, or a teardown's that
// ignores its status): the entry points that launch kernels check
// hipGetLastError() after their launches, and that check must see only
// their own.  Defined in pvglu_device.hip.
void begin_call();

// pvglu_refactor_solve whose launches skip every instance tile with no
// active instance (active: device flags, one per instance, or null; the
// batched Newton's frozen instances), with the schedule's width narrowed to
// the batch when it is wider (a batch compacted below the width it was
// tuned for).  Defined in pvglu_device.hip.
int refactor_solve_masked(pvglu_plan *P, const double *a, double *lu,
                          const double *b, double *x, int batch, void *stream,
                          const int *active);

// Stamping on load: seg holds kStampRec ints per kernel position of the LU
// levels (pvglu_launch.h), geq the module conductances [module][instance],
// gvs the homotopy conductance per instance; ieq and live (optional) the
// module branch currents and instance flags of the records that also form
// the correction-form right-hand side of their rows.
struct StampOnLoad {
  const int *seg;
  const double *geq;
  const double *gvs;
  const double *ieq = nullptr; // branch currents (right-hand side on load)
  const double *xk = nullptr;  // node voltages of the iteration
  const int *live = nullptr;   // instance activity flags
};
// refactor_solve_masked whose LU level launches (the levels that
// stamped_lu_levels lists for this batch) form the A values of the string
// segments with a stamping record from geq and gvs; every other A value is
// read from `a` as usual.
int refactor_solve_stamped(pvglu_plan *P, const StampOnLoad &so,
                           const double *a, double *lu, const double *b,
                           double *x, int batch, void *stream,
                           const int *active);
// The LU levels that a refactor + solve of `batch` instances runs as level
// launches under the plan's schedule (empty for COMPONENT, or without the
// fused forward substitution): the levels a stamped call stamps on load.
int stamped_lu_levels(pvglu_plan *P, int batch, std::vector<int> &levels);

// Levels of chains of columns in the flattened form of
// PV/008_convert_chains_columns_levels_into_vector.cpp:4-30.
struct LevelSet {
  std::vector<int> start_levels; // nlev + 1 (offsets into chains list)
  std::vector<int> start_chains; // nchains + 1 (offsets into cols)
  std::vector<int> chains;       // n columns
  int nlev() const { return (int)start_levels.size() - 1; }
  int nchains() const { return (int)start_chains.size() - 1; }
};

// Dependency rules of PV/006_get_dependencies_GLU.cpp plus the transposed L
// rule used by the adjoint.
enum DepKind { DEP_GLU = 0, DEP_L = 1, DEP_U = 2, DEP_LT = 3 };

// deps as CSR (ascending entries).  Host, O(nnz).
void dependencies(DepKind kind, int n, const std::vector<int> &rp,
                  const std::vector<int> &ci, std::vector<int> &dptr,
                  std::vector<int> &didx);
// 007 + 008.  Returns false on a dependency cycle.
bool levels_with_chains(int n, const std::vector<int> &dptr,
                        const std::vector<int> &didx, LevelSet &out);

// Per-kernel view of a level set, reordered so that the chains of one
// independent component are contiguous inside each level.
struct KernelLevels {
  std::vector<int> level_chain_ptr; // nlev+1, offsets into chain_col_ptr
  std::vector<int> chain_col_ptr;   // nchains+1, offsets into cols
  std::vector<int> cols;            // n
  // component view: comp_ptr[c]..comp_ptr[c+1] indexes comp_lv_* entries,
  // each a (chain_begin, chain_end) range of one level of component c.
  std::vector<int> comp_ptr;
  std::vector<int> comp_lv_begin, comp_lv_end, comp_lv_level;
  int max_comp_chains = 0; // widest level of any component
};

struct DevicePlan; // defined in the HIP translation unit

} // namespace pvglu

struct pvglu_plan {
  int n = 0, nnz = 0;
  // Filled pattern, CSR and CSC views with the maps of PV/010_create_CSR_CSC.cpp
  std::vector<int> rp, ci, cp, ri, csr2csc, csc2csr, csr_diag, csc_diag;
  // Reference-order level sets (exported for parity)
  pvglu::LevelSet lu, l, u, lt;
  std::vector<int> max_per_level;

  // Compiled gather schedule ------------------------------------------------
  // refactor: for CSC slot q, ops [op_ptr[q], op_ptr[q+1]) are (L slot, U slot)
  // pairs in ascending elimination column.
  std::vector<int> op_ptr, op_l, op_u;
  // forward substitution, row i: L(i,j) slots and columns j ascending
  std::vector<int> lrow_ptr, lrow_slot, lrow_col;
  // backward substitution, row i: U(i,j) slots and columns j descending
  std::vector<int> urow_ptr, urow_slot, urow_col;
  std::vector<int> diag_slot;
  bool fuse_lsolve = false;
  // per column: the diagonal's updates and row c of the forward sweep share
  // their L operands (one pass computes both sums)
  std::vector<int> col_flags; // bit 0 fused diagonal/forward row, bit 1 star column
  int ncomp = 0;
  std::vector<int> comp_of; // component id per column
  pvglu::KernelLevels k_lu, k_l, k_u, k_lt;
  // Register-forwarding programs (see compile_fast_programs): one 16-int
  // record per refactor column and one 8-int record per backward-substitution
  // row (and one per forward-substitution row of the L levels and per column
  // of the transposed sweeps), in kernel
  // order, the end of the register run holding each position
  // (0: generic column) and a per-chain whole-chain flag.
  std::vector<int> fast_lu_rec, fast_lu_run, fast_lu_chain, fast_lu_seg;
  std::vector<int> fast_u_rec, fast_u_run, fast_u_chain, fast_u_seg;
  std::vector<int> fast_l_rec, fast_l_run, fast_l_chain, fast_l_seg;
  std::vector<int> fast_ut_rec, fast_ut_run, fast_ut_chain, fast_ut_seg;
  std::vector<int> fast_lt_rec, fast_lt_run, fast_lt_chain, fast_lt_seg;
  int fast_lu_chains = 0, fast_u_chains = 0, fast_l_chains = 0;
  int fast_ut_chains = 0, fast_lt_chains = 0;
  int fast_lu_segs = 0, fast_u_segs = 0;
  uint64_t hash = 0;

  // schedule choice
  int sched_variant = PVGLU_SCHED_AUTO;
  int sched_width = 0;
  int sched_groups = 1; // component groups on separate streams (1, 2, 4)
  int cache_policy = 0;  // PVGLU_CACHE_AUTO / STREAMING / DEFAULT
  int tuned_batch = -1;

  pvglu::DevicePlan *dev = nullptr;
};

namespace pvglu {
// Build the kernel views (component-contiguous chains) of a level set.
void build_kernel_levels(const LevelSet &ls, const std::vector<int> &comp_of,
                         int ncomp, KernelLevels &out);
// Small-column register programs for the refactor (with the fused forward
// substitution) and the backward substitution.
void compile_fast_programs(pvglu_plan *P);
// Device teardown hook (defined in the HIP unit).
void device_plan_destroy(pvglu_plan *plan);
// The plan's device status word [code, column, instance] (null before the
// plan's first device call).
int *device_status(pvglu_plan *plan);
} // namespace pvglu
