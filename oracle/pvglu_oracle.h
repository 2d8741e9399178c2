/*
 * pvglu_oracle.h -- CPU oracle for the PV Newton-step GLU path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker.
 *
 * Plain C restatement of the reference algorithms (file:line citations are
 * relative to the reference repository, PV/ = PVPP_Plant_Simulator/).
 * Every array is 0-based int32 / double; CSR rows and CSC columns are sorted.
 *
 * Pins (see tests/test_oracle_*.py):
 *   orc_symbolic_fill ........ brute-force dense boolean elimination (n<=64),
 *                              closure invariant on random patterns
 *   orc_inclusive_scan ....... numpy.cumsum
 *   orc_deps_* / orc_levels .. reference units compiled into oracle/_ref
 *                              (006/007/008/009) on random + plant patterns,
 *                              plus partition / dependency-order invariants
 *   orc_csr_csc .............. CSR<->CSC round trip through both maps
 *   orc_glu_csc_sequential ... reference GLU_CSC_CPU_sequential in oracle/_ref
 *                              (bit-exact) and dense Doolittle LU (numpy)
 *   orc_substitute_lu_csc .... reference Substitute_LU_Vector_b_CSC_CPU in
 *                              oracle/_ref (bit-exact) and dense solve residual
 *   orc_newton_update ........ closed-form halving recurrence on hand cases
 *   orc_remove_zeros_diagonal  hand-worked examples (reference 004b needs its
 *                              util.cpp, which includes cuda.h: not buildable)
 *   orc_module_current ....... implicit single-diode equation solved by
 *                              bracketed Newton (independent of Lambert W)
 */
#ifndef PVGLU_ORACLE_H
#define PVGLU_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Symbolic no-pivot LU closure: (i,k) in L, (k,c) in U, k<min(i,c) => (i,c).
 * Same fixed point as PV/005a_obtain_fillin_matrix_CUDA.cu:18-236 (the
 * reference's pass loop with active-row masks reaches the unique closure).
 * Returns nnz of the filled pattern; *out_rp (n+1) and *out_ci are malloc'd. */
int orc_symbolic_fill(int n, const int *rp, const int *ci, int **out_rp,
                      int **out_ci);

/* In-place inclusive prefix sum, PV/005b_cumsum_CUDA.cu:3-31. */
void orc_inclusive_scan(int *a, int n);

/* Dependencies as CSR with ascending entries (std::set order).
 * kind 0: GLU3.0 rule, PV/006_get_dependencies_GLU.cpp:5-47
 *         deps(k) = { i<k : (i,k) in pattern }  (column k scanned twice)
 * kind 1: L rule, PV/006_get_dependencies_GLU.cpp:49-62  deps(i)={j<i:(i,j)}
 * kind 2: U rule, PV/006_get_dependencies_GLU.cpp:64-77  deps(i)={j>i:(i,j)}
 * Returns total dependency count; *dptr (n+1), *didx malloc'd. */
int orc_deps(int kind, int n, const int *rp, const int *ci, int **dptr,
             int **didx);

/* Levels with chains of columns, PV/007_convert_dependencies_in_levels.cpp
 * (basic levels :4-60, fusion :89-144, driver :146-210) flattened as in
 * PV/008_convert_chains_columns_levels_into_vector.cpp:4-30.
 * Outputs (malloc'd): start_levels[nlev+1], start_chains[nchains+1],
 * chains[n].  Returns nlev, or -1 if the dependency graph has a cycle. */
int orc_levels(int n, const int *dptr, const int *didx, int **start_levels,
               int **start_chains, int **chains, int *nchains);

/* Max filled-row length per level, PV/009_get_max_numbers_per_column.cpp. */
void orc_max_per_level(int nlev, const int *start_levels,
                       const int *start_chains, const int *chains,
                       const int *rp, int *out);

/* CSC pattern, value maps and diagonal pointers of a CSR pattern,
 * PV/010_create_CSR_CSC.cpp:15-128.  Arrays are caller-allocated:
 * cp[n+1], ri[nnz], csr2csc[nnz], csc2csr[nnz], csr_diag[n], csc_diag[n]. */
void orc_csr_csc(int n, const int *rp, const int *ci, int *cp, int *ri,
                 int *csr2csc, int *csc2csr, int *csr_diag, int *csc_diag);

/* In-place right-looking GLU on CSC values, sequential in column order,
 * PV/GLU_CSC_CPU_levels_strings.cpp:3-63,108-132.
 * Returns 0, or -1 (|pivot|<1e-8, :26-30) / -2 (no entry on or above the
 * diagonal, :15-20) with *bad_col set. */
int orc_glu_csc_sequential(int n, const int *rp, const int *ci,
                           const int *csr2csc, const int *cp, const int *ri,
                           double *vals, const int *csr_diag,
                           const int *csc_diag, int *bad_col);

/* Forward (unit L) then backward (U) column sweeps on CSC LU values,
 * PV/GLU_CSC_CPU_Triangular.cpp:3-43.  b is overwritten with x. */
void orc_substitute_lu_csc(int n, const int *cp, const int *ri,
                           const double *vals, const int *csc_diag, double *b);

/* Transposed solves for the adjoint: A^T y = g with A = L U, i.e. U^T z = g
 * (forward) then L^T y = z (backward).  No reference counterpart. */
void orc_solve_transpose_csc(int n, const int *cp, const int *ri,
                             const double *vals, const int *csc_diag,
                             double *g);

/* Damped voltage update of PV/011_calculate_new_iteration.cpp:6-17,37-86 for
 * ne elements: V_old in/out, V_new candidates (branch voltages).  *diff =
 * sum |delta| accumulated in long double in element order (:76-85).
 * Returns 0, or -1 when a damping factor falls below 1e-13 (:58-62). */
int orc_newton_update(int ne, double *v_old, const double *v_new,
                      long double *diff, double *max_abs_delta);

/* PV/util.cpp:61-77 */
void orc_apply_permutation(int n, double *vec, const int *perm);
void orc_invert_permutation(int n, const int *perm, int *inv);

/* Symmetric permutation new(k,l) = old(P[k],P[l]) of a CSR pattern with
 * values (PV/sparse_matrix.h:117-134 applied with Pinv), and row permutation
 * new[p[i]] = old[i] (PV/sparse_matrix.h:136-147).  Outputs caller-allocated
 * with the same nnz; rows come out sorted. */
void orc_permute_sym(int n, const int *rp, const int *ci, const double *v,
                     const int *P, int *orp, int *oci, double *ov);
void orc_permute_rows(int n, const int *rp, const int *ci, const double *v,
                      const int *p, int *orp, int *oci, double *ov);

/* Row-only swaps removing structurally zero diagonals,
 * PV/004b_remove_zeros_diagonal.cpp:12-100 with move_only_rows=true
 * (PV/004c_exchange_rows_columns.cpp:74-89).  The pattern (rp,ci,v) is
 * rewritten in place (same nnz); perm (identity on entry in the reference,
 * PV/004d_reorder_matrix.cpp:27-47) records the swaps.
 * Returns 0, or -1 when no swap partner exists (:82). */
int orc_remove_zeros_diagonal(int n, int *rp, int *ci, double *v, int *perm);

/* Single-diode module terminal current I(V) and dI/dV, dI/dIph from the
 * implicit equation I = Iph - I0 (exp((V+I Rs)/a) - 1) - (V + I Rs)/Rsh,
 * solved by safeguarded Newton on I. */
void orc_module_current(double V, double Iph, double I0, double a, double Rs,
                        double Rsh, double *I, double *dIdV, double *dIdIph);
/* orc_module_current over m modules (arrays of length m; dIdV / dIdIph may
 * be NULL). */
void orc_module_current_many(int m, const double *V, const double *Iph,
                             double I0, double a, double Rs, const double *Rsh,
                             double *I, double *dIdV, double *dIdIph);

/* The reference plant's diode models (PV/003_create_electric_system.cpp:
 * 7-85): kind 0 the panel diode I0 (exp(qkT V / na) - 1), kind 1 the bypass
 * diode 0.3e-6 (exp(qkT V) - 1), qkT = q / k / (273.15 + T); f and f' at V
 * for m elements. */
void orc_ref_diodes(int m, const int *kind, const double *V, double T,
                    double *f, double *df);

#ifdef __cplusplus
}
#endif
#endif
