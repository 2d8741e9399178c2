/*
 * pvglu.h -- C-ABI of the MI355X (gfx950) PV Newton-step linear-solve path.
 *
 * Plain C types only (pointers, sizes, an opaque stream handle that is a
 * hipStream_t).  Every entry point returns an int status (PVGLU_OK or a
 * negative PVGLU_E_* code) and records a message for pvglu_last_error() on
 * the calling thread.  Citations are to the reference entry point that each
 * function replaces (PV/ = PVPP_Plant_Simulator/).
 *
 * Value layout on the device ("slot-major, instance-minor"):
 *   values[slot * batch + instance],  slot = CSC position in the plan,
 *   vectors[row * batch + instance],  row  = permuted unknown index.
 * With batch == 1 this is the reference's plain CSC / dense vector layout.
 *
 * Numeric entry points (refactor, solve, transposed solve, Newton update,
 * stamping) are stream-ordered: they neither synchronize nor allocate, so a
 * caller can capture them in a hipGraph.  Numerical failures (pivot below
 * 1e-8, damping below 1e-13) are recorded in a device status word that
 * pvglu_status() reads (that call synchronizes the stream).
 */
#ifndef PVGLU_H
#define PVGLU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PVGLU_OK = 0,
  PVGLU_E_SINGULAR = -1, /* |pivot| < 1e-8 (PV/GLU_CSC_CPU_levels_strings.cpp:26-30) */
  PVGLU_E_NODIAG = -2,   /* column without a diagonal entry (:15-20) */
  PVGLU_E_HIP = -3,      /* HIP runtime error (PV/gpu_util.h:39-47 aborts) */
  PVGLU_E_ARG = -4,      /* invalid argument or pattern */
  PVGLU_E_NOCONV = -5,   /* Newton damping < 1e-13 or iteration cap
                            (PV/011_calculate_new_iteration.cpp:58-62,
                             PV/000_PVPP_simulator.cpp:179-181) */
  PVGLU_E_STATE = -6     /* call out of order (e.g. plan not uploaded) */
};

typedef struct pvglu_plan pvglu_plan;
typedef struct pvglu_multi pvglu_multi;

const char *pvglu_last_error(void);
void pvglu_free(void *p);
/* Library version and the device architecture the HIP code targets. */
const char *pvglu_version(void);
/* Number of visible HIP devices (0 when none; never errors). */
int pvglu_device_count(void);

/* ------------------------------------------------------------ symbolic */

/* Fill-in discovery on the GPU.  Replaces obtain_fillin_matrix_CUDA
 * (PV/005a_obtain_fillin_matrix_CUDA.cu:18-236) and its prefix sum
 * call_cumsum_CUDA (PV/005b_cumsum_CUDA.cu:22-31).  Input: the reordered
 * square pattern as CSR (sorted rows).  Output: the filled pattern, allocated
 * by the library (release with pvglu_free). */
int pvglu_obtain_fillin(int n, const int *rowptr, const int *colidx,
                        int **filled_rowptr, int **filled_colidx,
                        int *filled_nnz, void *stream);

/* Fill-in discovery that allocates nothing.  No counterpart in the
 * reference: for callers that cannot call hipMalloc (code under hipGraph
 * capture elsewhere in their stream, tight serving loops, harnesses that
 * forbid device allocation).  Same passes and result as
 * pvglu_obtain_fillin, with every device buffer in one file-scope device
 * array of the library (about 21 MB, held on each device the library runs
 * on until the process ends) and the output written to caller buffers:
 * filled_rowptr (n + 1 ints) and filled_colidx (capacity ints).  Limits:
 * n <= 524288 rows and 2097152 nonzeros in the input, every intermediate
 * and the filled pattern; beyond them the call fails with PVGLU_E_ARG and a
 * message naming the limit.  When the filled pattern has more than capacity
 * nonzeros the call fails with PVGLU_E_ARG and *filled_nnz holds the size
 * needed.  Synchronizes the stream once per pass (as pvglu_obtain_fillin).
 * Calls that overlap on one device share the array and race. */
int pvglu_obtain_fillin_noalloc(int n, const int *rowptr, const int *colidx,
                                int *filled_rowptr, int *filled_colidx,
                                int capacity, int *filled_nnz, void *stream);

/* Device prefix sum (inclusive, in place) over n int32 values resident on
 * the device.  Replaces call_cumsum_CUDA (PV/005b_cumsum_CUDA.cu:22-31).
 * Stream-ordered; one workgroup walks the array in 4096-value tiles with a
 * running carry, so it needs no scratch beyond the array itself. */
int pvglu_inclusive_scan_device(int *d_values, int n, void *stream);

/* Symbolic plan over a filled pattern: fill_fillins (PV/005c_Rellenar_Fillins.cpp:4-14),
 * get_dependencies_GLU_3_0 / _L / _U (PV/006_get_dependencies_GLU.cpp:5-77),
 * convert_dependence_list_in_chains_of_columns (PV/007_convert_dependencies_in_levels.cpp:146-210),
 * convert_chains_columns_levels_into_vector (PV/008_...:4-30),
 * get_max_numbers_per_column (PV/009_...:3-31),
 * create_CSR_CSC_matrices (PV/010_create_CSR_CSC.cpp:15-128), i.e. the
 * pattern half of prepare_matrices (PV/000_PVPP_simulator.cpp:33-90), plus the
 * compiled gather schedule the kernels execute.  Host only; no device work.
 * Fails with PVGLU_E_ARG if the pattern is not closed under elimination,
 * lacks a diagonal, or its dependency graph has a cycle. */
int pvglu_plan_create(int n, const int *filled_rowptr,
                      const int *filled_colidx, pvglu_plan **out);
void pvglu_plan_destroy(pvglu_plan *plan);

/* Copies the plan's arrays to the current HIP device (the device half of
 * struct_CPU_CUDA_arrays::create_CUDA, PV/factorize_electrical_system.h:63-98). */
int pvglu_plan_upload(pvglu_plan *plan);

/* Sizes and exported arrays, for parity tests and callers that keep the
 * reference's struct_CPU_CUDA_arrays fields. */
enum {
  PVGLU_INFO_N = 0,
  PVGLU_INFO_NNZ,
  PVGLU_INFO_NLEV_LU,
  PVGLU_INFO_NCHAIN_LU,
  PVGLU_INFO_NLEV_L,
  PVGLU_INFO_NCHAIN_L,
  PVGLU_INFO_NLEV_U,
  PVGLU_INFO_NCHAIN_U,
  PVGLU_INFO_NLEV_LT,
  PVGLU_INFO_NCHAIN_LT,
  PVGLU_INFO_NOPS,
  PVGLU_INFO_FUSED_LSOLVE,
  PVGLU_INFO_NCOMP,
  PVGLU_INFO_FAST_LU_CHAINS, /* refactor chains on the register-forwarding path */
  PVGLU_INFO_FAST_U_CHAINS,  /* backward-substitution chains on that path */
  PVGLU_INFO_FAST_LU_SEGMENTS, /* strided record segments, refactor */
  PVGLU_INFO_FAST_U_SEGMENTS,  /* strided record segments, backward */
  PVGLU_INFO_COUNT
};
int pvglu_plan_info(const pvglu_plan *plan, int what);

enum {
  PVGLU_ARR_CSR_ROWPTR = 0,     /* mat_A_CSR_start_rows                n+1 */
  PVGLU_ARR_CSR_COLIDX,         /* mat_A_CSR_position_columns          nnz */
  PVGLU_ARR_CSC_COLPTR,         /* mat_A_CSC_start_columns             n+1 */
  PVGLU_ARR_CSC_ROWIDX,         /* mat_A_CSC_position_rows             nnz */
  PVGLU_ARR_CSR2CSC,            /* mat_A_CSR_corresponding_value_in_CSC nnz */
  PVGLU_ARR_CSC2CSR,            /* mat_A_CSC_corresponding_value_in_CSR nnz */
  PVGLU_ARR_CSR_DIAG,           /* mat_A_CSR_start_from_diagonal       n */
  PVGLU_ARR_CSC_DIAG,           /* mat_A_CSC_start_from_diagonal       n */
  PVGLU_ARR_LU_START_LEVELS,    /* start_levels_unified_LU             nlev+1 */
  PVGLU_ARR_LU_START_CHAINS,    /* start_chains_columns_unified_LU     nch+1 */
  PVGLU_ARR_LU_CHAINS,          /* chains_columns_unified_LU           n */
  PVGLU_ARR_L_START_LEVELS,
  PVGLU_ARR_L_START_CHAINS,
  PVGLU_ARR_L_CHAINS,
  PVGLU_ARR_U_START_LEVELS,
  PVGLU_ARR_U_START_CHAINS,
  PVGLU_ARR_U_CHAINS,
  PVGLU_ARR_MAX_PER_LEVEL,      /* max_elements_per_column             nlev_lu */
  /* kernel programs of this build (no counterpart in the reference): */
  PVGLU_ARR_COL_FLAGS,          /* per column: bit 0 diagonal and forward row
                                   share their L operands, bit 1 star column  n */
  PVGLU_ARR_LU_SEGMENTS,        /* refactor register program, per position in
                                   kernel order: (length, slot step, column
                                   step, flags) at a segment start, else 0;
                                   flags: bit 0/1 operands forwarded (with /
                                   without the forward row), bit 2/3 the
                                   tridiagonal step (bits 4-5: the previous
                                   column's forwarded entry)          4n */
  PVGLU_ARR_COUNT
};
/* Length of an exported array, or a negative status. */
int pvglu_plan_array_len(const pvglu_plan *plan, int which);
int pvglu_plan_export(const pvglu_plan *plan, int which, int *out);

/* On-disk plan: the pattern half of struct_CPU_CUDA_arrays
 * (PV/factorize_electrical_system.h:3-61) -- every pvglu_plan_export array --
 * and optionally the ordering (P, Prow; NULL to omit), with a checksum.
 * Loading rebuilds the plan from the stored filled pattern and fails with
 * PVGLU_E_ARG if any stored array differs from the rebuilt one; P / Prow (n
 * ints each, may be NULL) receive the stored ordering (identity if none was
 * stored).  pvglu_plan_file_n returns n from a file's header.  No counterpart
 * in the reference, which recomputes the symbolic phase every run. */
int pvglu_plan_save(const pvglu_plan *plan, const char *path, const int *P,
                    const int *Prow);
int pvglu_plan_file_n(const char *path);
int pvglu_plan_load(const char *path, pvglu_plan **out, int *P, int *Prow);

/* Kernel schedule: variant and instances-per-chain tile width.
 *   PVGLU_SCHED_LEVEL      one kernel per level (reference launch structure,
 *                          PV/factorize_electrical_system_CUDA.cu:67-144)
 *   PVGLU_SCHED_COMPONENT  one launch for everything: one workgroup per
 *                          (independent component, instance tile) runs all
 *                          levels with workgroup barriers
 *   PVGLU_SCHED_HYBRID     wide levels one kernel each; every run of two or
 *                          more consecutive narrow levels (at most 4 chains
 *                          per component) one COMPONENT-style launch
 *   PVGLU_SCHED_AUTO       HYBRID from batch 16 up, COMPONENT below (for
 *                          plans with 8 or more independent components)
 * width: instances handled side by side per chain (1..64, power of two;
 * 0: the largest of 64, 32, 16, 4, 1 not above the batch). */
enum {
  PVGLU_SCHED_AUTO = 0,
  PVGLU_SCHED_LEVEL = 1,
  PVGLU_SCHED_COMPONENT = 2,
  PVGLU_SCHED_HYBRID = 3
};
int pvglu_set_schedule(pvglu_plan *plan, int variant, int width);
int pvglu_get_schedule(const pvglu_plan *plan, int *variant, int *width);

/* Component groups of the LEVEL and HYBRID schedules (1, 2 or 4; default 1).
 * With G > 1 the independent components are split into G groups whose
 * launches run on the plan's own side streams, forked from and joined back
 * into the caller's stream with events, each group starting after the
 * previous group's first level.  On a stream under capture the setting is
 * ignored (one group).  The side streams and events belong to the plan: two
 * numeric calls on one plan from different host threads at the same time
 * need G = 1.  No counterpart in the reference. */
int pvglu_set_stream_groups(pvglu_plan *plan, int groups);

/* Cache policy of the width-32 and width-64 LEVEL and HYBRID kernels' value
 * streams.
 * STREAMING marks the loads of A, b and the substitution operands and the
 * stores of L entries and forward results non-temporal (they are not read
 * again before eviction); DEFAULT leaves every access to the cache; AUTO
 * (default) streams when the batch's algorithmic bytes exceed twice the
 * 256 MB Infinity Cache.  pvglu_autotune tries both.  No counterpart in the
 * reference. */
enum {
  PVGLU_CACHE_AUTO = 0,
  PVGLU_CACHE_STREAMING = 1,
  PVGLU_CACHE_DEFAULT = 2
};
int pvglu_set_cache_policy(pvglu_plan *plan, int policy);
int pvglu_get_stream_groups(const pvglu_plan *plan);

/* Times the schedule candidates (variant x width x component groups x cache
 * policy) for this batch size on scratch buffers (allocates and
 * synchronizes: call it outside any capture) and keeps the fastest.  Results
 * are cached per (pattern hash, batch): with best_ms NULL a cached choice is
 * taken without timing; with best_ms set the candidates are timed again and
 * *best_ms receives the fastest median. */
int pvglu_autotune(pvglu_plan *plan, int batch, int reps, double *best_ms);

/* ------------------------------------------------------------- numeric */

/* Level-scheduled numeric refactorization without pivoting.  Replaces
 * factorize_electrical_system_CUDA (PV/factorize_electrical_system_CUDA.cu:61-152)
 * and its kernel (PV/factorize_matrix_GLU_CUDA.cuh:51-105,161-211).  Reads A
 * values, writes L\U values (unit L below the diagonal, U on and above); a may
 * equal lu (in place, as the reference).  Per column the update order is the
 * sequential CPU order (GLU_CSC_CPU_sequential, PV/GLU_CSC_CPU_levels_strings.cpp:108-132). */
int pvglu_refactor(pvglu_plan *plan, const double *a, double *lu, int batch,
                   void *stream);

/* L then U level-scheduled substitution.  Replaces triangular_substitution_LU_CUDA
 * (PV/triangular_substitution.cu:232-271).  x may equal b. */
int pvglu_solve(pvglu_plan *plan, const double *lu, const double *b, double *x,
                int batch, void *stream);

/* Refactor followed by the solve, with the forward substitution riding in the
 * refactor sweep when the pattern allows it. */
int pvglu_refactor_solve(pvglu_plan *plan, const double *a, double *lu,
                         const double *b, double *x, int batch, void *stream);

/* Adjoint: solves A^T y = g with the factors of A (U^T then L^T).  No
 * reference counterpart.  y may equal g. */
int pvglu_solve_transpose(pvglu_plan *plan, const double *lu, const double *g,
                          double *y, int batch, void *stream);

/* Reads and clears nothing: returns the first recorded numerical failure
 * (code PVGLU_OK if none) with its column and instance.  Synchronizes. */
int pvglu_status(pvglu_plan *plan, int *code, int *column, int *instance);
/* Clears the device status word (a fill kernel; capture-safe). */
int pvglu_status_reset(pvglu_plan *plan, void *stream);

/* Multi-pattern batch: groups of instances sharing one plan each (mixed plant
 * sizes), refactored and solved in one launch from a single work list. */
int pvglu_multi_create(int ngroups, pvglu_plan *const *plans,
                       const int *batches, pvglu_multi **out);
void pvglu_multi_destroy(pvglu_multi *m);
int pvglu_multi_refactor_solve(pvglu_multi *m, const double *const *a,
                               double *const *lu, const double *const *b,
                               double *const *x, void *stream);
int pvglu_multi_solve_transpose(pvglu_multi *m, const double *const *lu,
                                const double *const *g, double *const *y,
                                void *stream);

/* -------------------------------------------------------------- Newton */

/* Damped per-element voltage update of calculate_new_iteration
 * (PV/011_calculate_new_iteration.cpp:37-86) for ne nonlinear elements in
 * each of batch instances.  x: permuted node voltages [row][instance];
 * pos/neg: permuted row of each element's terminals (-1 = ground);
 * v_elem: element voltages [elem][instance], updated in place.
 * delta_sum / delta_max: per-instance sum and max of |delta| (device,
 * batch entries each).  The sum is accumulated in 2^-44 V fixed point (each
 * |delta| <= 1 V rounded down to a multiple of 2^-44 V, integer additions),
 * so it is the same whatever the block order or batch size; the reference
 * sums in long double (PV/011_calculate_new_iteration.cpp:39,76-85).  active (device, batch ints, may be NULL): instances
 * with 0 are left untouched (converged).  Damping below 1e-13 records
 * PVGLU_E_NOCONV in the status word of `plan`. */
int pvglu_newton_update(pvglu_plan *plan, int ne, int batch, const int *pos,
                        const int *neg, const double *x, double *v_elem,
                        double *delta_sum, double *delta_max,
                        const int *active, void *stream);

/* ---------------------------------------------------- PV plant harness */

/* The config plants: inverters x strings x modules, single-diode modules in
 * series, each string's combiner tied by a cable to its inverter's DC bus,
 * each bus a Norton source (1e4 S, 600 V).  Built through the reference's
 * circuit API (add_linear_element / add_nonlinear_voltage_source,
 * PV/002_base_electric_system.cpp:78-108,164-173). */
typedef struct pvglu_pv pvglu_pv;
int pvglu_pv_create(int inverters, int strings, int modules, uint64_t seed,
                    pvglu_pv **out);
void pvglu_pv_destroy(pvglu_pv *pv);
enum {
  PVGLU_PV_N = 0,        /* unknowns (node voltages) */
  PVGLU_PV_NMODULES,     /* nonlinear elements */
  PVGLU_PV_NCABLES,
  PVGLU_PV_NNZ_A,        /* nonzeros of the MNA matrix */
  PVGLU_PV_COUNT
};
int pvglu_pv_info(const pvglu_pv *pv, int what);
/* Element data: module terminals (+,-) as node indices (-1 ground), Rsh per
 * module, cable ends and conductance, bus nodes. */
int pvglu_pv_modules(const pvglu_pv *pv, int *pos, int *neg, double *rsh);
int pvglu_pv_cables(const pvglu_pv *pv, int *a, int *b, double *g);
int pvglu_pv_buses(const pvglu_pv *pv, int *bus);
/* Module positions on the 1 km^2 grid (metres). */
int pvglu_pv_positions(const pvglu_pv *pv, double *x, double *y);
/* Irradiance (W/m^2) per module at time t (s) for a seeded cloud scenario
 * (8 Gaussian shadow blobs over a 1000 W/m^2 clear sky); scenario < 0 gives
 * the uniform 1000 W/m^2 sky. */
int pvglu_pv_irradiance(const pvglu_pv *pv, int64_t scenario, double t,
                        double *g_out);

/* Device cloud generator (the analogue of PV/001_clouds.cu:96-155 for the
 * config cloud model): irradiance of `batch` instances, instance k at
 * scenarios[k] (< 0: uniform 1000 W/m^2) and time times[k] s, into g_dev
 * [module][instance] on `stream`.  scenarios and times are host arrays.  The
 * first call (and a call with a larger batch) allocates the plant's
 * parameter buffer; the call waits for its own upload before returning. */
int pvglu_pv_irradiance_device(pvglu_pv *pv, int batch,
                               const int64_t *scenarios, const double *times,
                               double *g_dev, void *stream);
/* MNA pattern of the unpermuted system (CSR, sorted) and AMD + row-swap
 * reordering of it: reorder_matrix (PV/004d_reorder_matrix.cpp:27-47). */
int pvglu_pv_pattern(const pvglu_pv *pv, int *rowptr, int *colidx);
/* The plant's symbolic plan (built on first use: reorder, GPU fill-in,
 * plan_create, upload); owned by the plant. */
pvglu_plan *pvglu_pv_plan(pvglu_pv *pv);
/* Device companion-model stamping for `batch` instances: irradiance g and
 * module voltages v ([module][instance], device) -> permuted CSC values a
 * ([slot][instance]) and right-hand side b ([row][instance]), the
 * create_matrix_and_vector + permutation + fill_values_CSC_only steps
 * (PV/002b_sistema_electrico_base_matriz.cpp:141-203, PV/000_PVPP_simulator.cpp:185-210,
 * PV/010_create_CSR_CSC.cpp:142-155) as one gather per slot.  g_virtual > 0
 * adds the homotopy conductance across every module.  Stream-ordered. */
int pvglu_pv_stamp(pvglu_pv *pv, int batch, const double *g, const double *v,
                   double g_virtual, double *a, double *b, void *stream);
/* Module current I(V) (terminal form) and its derivatives. */
int pvglu_pv_module_current(double V, double Iph, double rsh, double *I,
                            double *dIdV, double *dIdIph);

/* Ordering: AMD 1.1 restatement on A+A^T (PV/004a_algorithm_amd_function.cpp:10-50)
 * followed by the row-only zero-diagonal swaps (PV/004b_remove_zeros_diagonal.cpp:12-100).
 * P[k] = original index at position k; Prow as the reference records it. */
int pvglu_amd_order(int n, const int *rowptr, const int *colidx, int *P);
int pvglu_reorder(int n, const int *rowptr, const int *colidx, int *P,
                  int *Prow);

/* Newton step form of the PV drivers below (drop-in, batched, gradients).
 *   PVGLU_NEWTON_DIRECT      the reference's step: the companion system
 *                            A(v) x = b(v) solved for the node voltages
 *                            (PV/000_PVPP_simulator.cpp:185-230 with
 *                            PV/002b_sistema_electrico_base_matriz.cpp:141-203).
 *   PVGLU_NEWTON_CORRECTION  (default) the same Newton step written as a
 *                            correction: A(v) d = b(v) - A(v) x_prev,
 *                            x = x_prev + d, the right-hand side evaluated
 *                            branch by branch (KCL residual of the linearized
 *                            circuit at x_prev; x_prev = 0 at the start).
 * Both use the same matrices, refactor and solve; in exact arithmetic their
 * iterates are equal.  In FP64 the direct form's iterates carry the solve's
 * rounding of the ~600 V node voltages (last-stage noise floor up to ~1e-9 V
 * behind weak, reverse-biased modules), the correction form's only the
 * rounding of the branch currents (~1e-13..1e-12 V).  The stall rule of
 * pvglu_pv_newton_dropin / _batched uses noise level 1e-8 V (direct) or
 * 1e-11 V (correction).  No counterpart in the reference (its only form is
 * the direct one). */
enum { PVGLU_NEWTON_CORRECTION = 0, PVGLU_NEWTON_DIRECT = 1 };
int pvglu_pv_set_newton_form(pvglu_pv *pv, int form);
int pvglu_pv_get_newton_form(const pvglu_pv *pv);

/* Batched Newton (pvglu_pv_newton_batched, pvglu_pv_gradients): on = 1 (the
 * default) stops spending work on finished instances (converged, or stopped
 * at the noise floor).  Between polls (every 4 graph replays) the refactor +
 * solve skips every instance tile (the schedule's width of consecutive
 * instances) with no active instance; at a poll that finds at most three
 * quarters of the working batch active, the finished instances' results are
 * written to their columns of the caller's arrays and the active ones are
 * gathered into a smaller working batch, for which the iteration graphs are
 * captured again.  Per-instance arithmetic does not depend on the batch, so
 * results are bit-identical with 0, which runs every instance on every
 * iteration (for measurement).  No counterpart in the reference (one
 * instance per solve). */
int pvglu_pv_set_skip_frozen(pvglu_pv *pv, int on);

/* Batched Newton, correction form: on = 1 (the default) forms the companion
 * matrix's series-string entries inside the refactor (each from the two
 * module conductances it sums, in the stamping order, so bit-identical to
 * the stamped values) instead of stamping them into A and reading them
 * back; only the other entries are stamped.  Applies where the refactor's
 * LU levels run as level launches (HYBRID, LEVEL); 0 stamps every entry.
 * No counterpart in the reference (host assembly of every entry,
 * PV/002b_sistema_electrico_base_matriz.cpp:183-203). */
int pvglu_pv_set_stamp_on_load(pvglu_pv *pv, int on);

/* Host drop-in Newton solve of one plant, with the reference's call sequence
 * (PV/000_PVPP_simulator.cpp:144-235): symbolic setup once, then per
 * iteration host assembly, permutation, CSC value fill, GPU refactor, GPU
 * solve, un-permutation and the 011 update.  g: irradiance per module;
 * v_module: initial/final module voltages (in/out); x_out: node voltages (n).
 * tol: stop when max |delta| <= tol on the last homotopy stage (or after
 * 20 iterations with max |delta| at or below the form's noise level: the
 * FP64 noise floor);
 * ladder: 1 = the reference's conductance ladder, 0 = last stage only.
 * fixed_iters > 0 runs exactly that many iterations of the last stage.
 * stats[0..3]: iterations, host assembly ms, device ms, total ms. */
int pvglu_pv_newton_dropin(pvglu_pv *pv, const double *g, double *v_module,
                           double *x_out, double tol, int ladder,
                           int fixed_iters, double *stats);

/* Batched device Newton for `batch` scenarios of one plant (shared pattern):
 * device stamping, refactor+solve, 011 update, all stream-ordered and
 * replayed from a hipGraph per iteration.  g: [module][instance] host array;
 * v_module in/out [module][instance]; x_out [node][instance] (unpermuted);
 * power_out[instance] = sum over buses of V (G_N (V - 600)).
 * stats[0..3]: iterations, total ms, ms per iteration, max |delta|. */
int pvglu_pv_newton_batched(pvglu_pv *pv, int batch, const double *g,
                            double *v_module, double *x_out, double *power_out,
                            double tol, int max_iters, int ladder,
                            double *stats);

/* Per-instance iteration counts and final max |delta| (V) of the last
 * batched solve.  An instance whose final max |delta| exceeds tol stopped at
 * the FP64 noise floor (20 iterations with max |delta| at or below the
 * form's noise level). */
int pvglu_pv_newton_info(const pvglu_pv *pv, int batch, int *iters,
                         double *last_max_delta);

/* Adjoint gradients of total power at a converged batched solution:
 * dP/dG per module [module][instance] and dP/dg per cable [cable][instance].
 * Runs the Newton solve first (same arguments as pvglu_pv_newton_batched). */
int pvglu_pv_gradients(pvglu_pv *pv, int batch, const double *g,
                       double *v_module, double *power_out, double *dP_dG,
                       double *dP_dgc, double tol, int max_iters);

/* ------------------------------------------------ the reference's plant */

/* The reference's own PV plant (PV/003_create_electric_system.cpp:184-382,
 * PV/002_base_electric_system.cpp, PV/002b_sistema_electrico_base_matriz.cpp):
 * `blocks` inverters, each a block of panels_y strings of panels_x panels
 * (per panel an internal node, series and shunt resistances, a diode, the
 * photocurrent and a bypass diode when `bypass`), the strings optionally
 * cross-tied (interconnection), the block's DC side tied to ground and to an
 * inverter: a DC voltage source of panels_x * volts_per_panel whose power is
 * injected as an AC current (fixed from the previous iteration) between two
 * AC nodes; AC nodes are (real, imaginary) unknown pairs, consecutive blocks'
 * AC nodes are joined by wires and the first and last blocks' to the point
 * of common coupling, a 25 kV AC voltage source.  Voltage sources add their
 * currents as unknowns, with zero diagonals (hence the row swaps of
 * reorder_matrix), and the AC ring and the parallel strings give fill-in.
 * irradiance: W/m^2 for every panel, or per panel (panel_irradiance, block-
 * major, then string, then position; NULL = uniform); temperature in C. */
typedef struct pvglu_refplant pvglu_refplant;
enum {
  PVGLU_INTERCONNECTION_NOTHING = 0,
  PVGLU_INTERCONNECTION_FULL = 1,
  PVGLU_INTERCONNECTION_HONEYCOMB = 2,
  PVGLU_INTERCONNECTION_BRIDGE = 3
};
int pvglu_refplant_create(int blocks, int panels_x, int panels_y,
                          double volts_per_panel, int interconnection,
                          double irradiance, const double *panel_irradiance,
                          double temperature, int bypass,
                          pvglu_refplant **out);
void pvglu_refplant_destroy(pvglu_refplant *rp);
enum {
  PVGLU_REF_N = 0,         /* unknowns (nodes and voltage-source currents) */
  PVGLU_REF_NNZ_A,         /* nonzeros of the system before fill */
  PVGLU_REF_NONLINEAR,     /* nonlinear elements (diodes) */
  PVGLU_REF_INVERTERS,
  PVGLU_REF_VSOURCES,      /* voltage sources (DC inverter sides, PCC) */
  PVGLU_REF_PANELS,
  PVGLU_REF_PCC_NODE,      /* the PCC's AC node (1-based; imaginary = +1) */
  PVGLU_REF_PCC_CURRENT    /* the PCC source current's unknown (real part) */
};
int pvglu_refplant_info(const pvglu_refplant *rp, int what);
/* Pattern of the system at virtual conductance 1.0 (CSR, unpermuted). */
int pvglu_refplant_pattern(const pvglu_refplant *rp, int *rowptr, int *colidx);
/* prepare_matrices (PV/000_PVPP_simulator.cpp:10-93): reorder (AMD + the
 * row-only zero-diagonal swaps), GPU fill-in, plan; P and Prow as
 * reorder_matrix records them (n each).  Fails if Prow is not an
 * involution (the Newton loop's exchange_rows(Prow) needs one). */
int pvglu_refplant_ordering(pvglu_refplant *rp, int *P, int *Prow);
pvglu_plan *pvglu_refplant_plan(pvglu_refplant *rp);
/* Elements in stamping order (see pvglu_refplant.cpp for the layout);
 * returns their count (elem / val may be NULL to count). */
enum {
  PVGLU_ELEM_ADMITTANCE = 0,
  PVGLU_ELEM_AC_ADMITTANCE,
  PVGLU_ELEM_CURRENT_SOURCE,
  PVGLU_ELEM_VOLTAGE_SOURCE,
  PVGLU_ELEM_NONLINEAR,
  PVGLU_ELEM_INVERTER
};
int pvglu_refplant_elements(const pvglu_refplant *rp, int *elem, double *val);
/* Unknown of each node (node_x[node - 1]); returns the node count. */
int pvglu_refplant_node_unknowns(const pvglu_refplant *rp, int *node_x);
int pvglu_refplant_set_newton_form(pvglu_refplant *rp, int form);
/* The permuted system at the current iterate in the plan's CSC slot order
 * (the values and right-hand side the refactor and solve receive). */
int pvglu_refplant_system(pvglu_refplant *rp, double g_virtual, double *a_csc,
                          double *b_perm);
/* Iterate state: nonlinear element voltages and per inverter (V2 real,
 * V2 imaginary, I1) -- the reference's vector_voltages_nonlinear_...iterate
 * and list_values_inverter_dc_ac_iterate. */
int pvglu_refplant_get_state(const pvglu_refplant *rp, double *v, double *inv);
int pvglu_refplant_set_state(pvglu_refplant *rp, const double *v,
                             const double *inv);
/* The reference's Newton loop on this plant (PV/000_PVPP_simulator.cpp:
 * 144-253) from the current state: ladder (or the last stage only), host
 * assembly, permutations, GPU refactor and solve, the 011 update including
 * the inverter refresh (PV/011_calculate_new_iteration.cpp:88-109); stops
 * when the last stage's largest change (element voltages and inverter
 * states) is <= tol, or at the form's noise floor.  x_out (n, may be NULL):
 * unknowns in circuit order.  stats[0..5]: iterations, host ms, device ms,
 * total ms, real and imaginary part of the PCC power V conj(I). */
int pvglu_refplant_newton_dropin(pvglu_refplant *rp, double tol, int ladder,
                                 double *x_out, double *stats);

#ifdef __cplusplus
}
#endif
#endif
