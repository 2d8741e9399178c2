// ref_harness.cpp -- C entry points over the reference's own CPU units.
//
// TEST INFRASTRUCTURE ONLY.  oracle/build_ref.sh compiles this file together
// with reference sources read where they lie (AMD 1.1 in PV/AmdLib/*.c,
// PV/006..009, PV/GLU_CSC_CPU_levels_strings.cpp, PV/GLU_CSC_CPU_Triangular.cpp)
// into oracle/_ref/libpvref*.so.  Nothing here is a stand-in for a reference
// function: the harness only converts CSR arrays into the reference's
// std::vector / std::set containers, calls the reference functions, and
// flattens the results.  Units that need util.cpp (which includes cuda.h) are
// not built.
#include "main.h"
#include "sparse_matrix.h"

#include <cstdlib>
#include <cstring>

extern "C" {
#include "amd.h"
}

// Declarations exactly as the reference declares them (PV/matrix_functions.h,
// PV/006_get_dependencies_GLU.cpp, PV/009_get_max_numbers_per_column.cpp).
vector<set<int>> get_dependencies_GLU_3_0(int n,
                                          const vector<vector<int>> &fill);
vector<set<int>> get_dependencies_matrix_L(int n,
                                           const vector<vector<int>> &fill);
vector<set<int>> get_dependencies_matrix_U(int n,
                                           const vector<vector<int>> &fill);
vector<vector<list<int>>>
convert_dependence_list_in_chains_of_columns(const vector<set<int>> &deps);
void convert_chains_columns_levels_into_vector(
    const vector<vector<list<int>>> &levels, vector<int> &start_levels,
    vector<int> &start_chains, vector<int> &chains);
vector<int>
get_max_numbers_per_column(const vector<vector<list<int>>> &levels,
                           const vector<vector<int>> &nonzero_values);
void GLU_CSC_CPU_sequential(const int n, const int *CSR_start_rows,
                            const int *CSR_position_columns,
                            const int *CSR_corresponding_value_in_CSC,
                            const int *CSC_start_columns,
                            const int *CSC_position_rows,
                            FLOATING_TYPE *CSC_values,
                            const int *CSR_start_from_diagonal,
                            const int *CSC_start_from_diagonal);
void Substitute_LU_Vector_b_CSC_CPU(const unsigned int n,
                                    const int *CSC_start_columns,
                                    const int *CSC_position_rows,
                                    const FLOATING_TYPE *CSC_values,
                                    const int *CSC_start_from_diagonal,
                                    FLOATING_TYPE *b);

static vector<vector<int>> rows_of(int n, const int *rp, const int *ci) {
  vector<vector<int>> r(n);
  for (int i = 0; i < n; i++)
    r[i].assign(ci + rp[i], ci + rp[i + 1]);
  return r;
}

static int *dup_vec(const vector<int> &v) {
  int *p = (int *)malloc(sizeof(int) * (v.size() ? v.size() : 1));
  if (!v.empty())
    memcpy(p, v.data(), sizeof(int) * v.size());
  return p;
}

extern "C" {

void ref_free(void *p) { free(p); }

// PV/004a_algorithm_amd_function.cpp:13-42: amd_order on the row-major
// boolean pattern, default Control (NULL).  Returns the AMD status.
int ref_amd_order(int n, const int *rp, const int *ci, int *P) {
  return amd_order(n, rp, ci, P, NULL, NULL);
}

int ref_deps(int kind, int n, const int *rp, const int *ci, int **dptr,
             int **didx) {
  auto fill = rows_of(n, rp, ci);
  vector<set<int>> d = kind == 0   ? get_dependencies_GLU_3_0(n, fill)
                       : kind == 1 ? get_dependencies_matrix_L(n, fill)
                                   : get_dependencies_matrix_U(n, fill);
  vector<int> ptr(1, 0), idx;
  for (auto &s : d) {
    for (int x : s)
      idx.push_back(x);
    ptr.push_back((int)idx.size());
  }
  *dptr = dup_vec(ptr);
  *didx = dup_vec(idx);
  return (int)idx.size();
}

int ref_levels(int n, const int *dptr, const int *didx, int **start_levels,
               int **start_chains, int **chains, int *nchains) {
  vector<set<int>> d(n);
  for (int i = 0; i < n; i++)
    d[i].insert(didx + dptr[i], didx + dptr[i + 1]);
  auto lv = convert_dependence_list_in_chains_of_columns(d);
  vector<int> sl, sc, cc;
  convert_chains_columns_levels_into_vector(lv, sl, sc, cc);
  *start_levels = dup_vec(sl);
  *start_chains = dup_vec(sc);
  *chains = dup_vec(cc);
  *nchains = (int)sc.size() - 1;
  return (int)sl.size() - 1;
}

int ref_max_per_level(int n, const int *rp, const int *ci, const int *dptr,
                      const int *didx, int *out) {
  vector<set<int>> d(n);
  for (int i = 0; i < n; i++)
    d[i].insert(didx + dptr[i], didx + dptr[i + 1]);
  auto lv = convert_dependence_list_in_chains_of_columns(d);
  auto m = get_max_numbers_per_column(lv, rows_of(n, rp, ci));
  for (size_t l = 0; l < m.size(); l++)
    out[l] = m[l];
  return (int)m.size();
}

void ref_glu_csc_sequential(int n, const int *rp, const int *ci,
                            const int *csr2csc, const int *cp, const int *ri,
                            double *vals, const int *csr_diag,
                            const int *csc_diag) {
  GLU_CSC_CPU_sequential(n, rp, ci, csr2csc, cp, ri, vals, csr_diag,
                         csc_diag);
}

void ref_substitute_lu_csc(int n, const int *cp, const int *ri,
                           const double *vals, const int *csc_diag,
                           double *b) {
  Substitute_LU_Vector_b_CSC_CPU((unsigned)n, cp, ri, vals, csc_diag, b);
}

} // extern "C"
