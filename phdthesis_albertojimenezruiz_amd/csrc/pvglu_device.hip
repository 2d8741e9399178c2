// pvglu_device.hip -- gfx950 kernels and the device half of the C-ABI.
//
// Numeric path (the Newton-step refactor + solve of the reference,
// PV/factorize_electrical_system_CUDA.cu:61-152, PV/factorize_matrix_GLU_CUDA.cuh:51-211,
// PV/triangular_substitution.cu:61-271), written in gather form:
//   * a lane owns one (column, instance) and computes each entry of the column
//     as  a(i,c) - sum_j L(i,j) U(j,c)  with j ascending, reading operands
//     that earlier levels (or earlier columns of its chain) finished, so no
//     atomics are needed and every entry sees the exact operation order of the
//     sequential CPU GLU (bit-identical results);
//   * values are slot-major / instance-minor: with `width` lanes of one chain
//     on consecutive instances, a wavefront reads 64 x 8 B = 512 B contiguous;
//     when width == 64 all pattern metadata is wavefront-uniform and lives in
//     scalar registers (readfirstlane), so pattern reads are s_load traffic;
//   * LEVEL schedule: one launch per level (the reference's launch structure);
//     COMPONENT schedule: one launch; each workgroup owns one independent
//     component x instance tile and walks all levels with workgroup barriers.
#include <hip/hip_runtime.h>

#include "pvglu_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

using namespace pvglu;

#define HIP_TRY(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail(PVGLU_E_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                         \
  } while (0)

namespace pvglu {

// Device mirror of a level set in kernel order.
struct DevLevels {
  int *level_chain_ptr = nullptr, *chain_col_ptr = nullptr, *cols = nullptr;
  int *comp_ptr = nullptr, *comp_lv_begin = nullptr, *comp_lv_end = nullptr,
      *comp_lv_level = nullptr;
  int nlev = 0, max_comp_chains = 0;
  std::vector<int> h_level_chain_ptr; // host copy for launch sizing
  std::vector<int> h_level_comp_chains; // widest component per level
  // chain range [begin, end) of component group g at level lv, for 1, 2 and
  // 4 groups: h_grp[gi][(g * nlev + lv) * 2 + {0, 1}]
  std::vector<int> h_grp[3];
};

struct DevicePlan {
  int device = -1;
  int *cp = nullptr, *ri = nullptr, *op_ptr = nullptr, *op_l = nullptr,
      *op_u = nullptr, *diag_slot = nullptr, *lrow_ptr = nullptr,
      *lrow_slot = nullptr, *lrow_col = nullptr, *urow_ptr = nullptr,
      *urow_slot = nullptr, *urow_col = nullptr;
  DevLevels lu, l, u, lt;
  int *lu_rec = nullptr, *lu_seg = nullptr, *lu_run = nullptr,
      *lu_fast = nullptr, *u_rec = nullptr, *u_seg = nullptr, *u_run = nullptr,
      *u_fast = nullptr, *l_rec = nullptr, *l_seg = nullptr, *l_run = nullptr,
      *l_fast = nullptr, *ut_rec = nullptr, *ut_seg = nullptr, *ut_run = nullptr,
      *ut_fast = nullptr, *lt_rec = nullptr, *lt_seg = nullptr,
      *lt_run = nullptr, *lt_fast = nullptr, *drf = nullptr;
  int *status = nullptr; // [code, column, instance, pad]
  std::vector<void *> allocs;
  // component groups (first component of each group, 1 / 2 / 4 groups) and
  // the side streams and events of the grouped schedules
  std::vector<int> comp_bounds[3];
  bool side_ok = false;
  hipStream_t side[3] = {};
  hipEvent_t ev_fork = nullptr, ev_skew[3] = {}, ev_join[3] = {};
};

} // namespace pvglu

// ------------------------------------------------------------ kernel views
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

static Pat pat_of(const DevicePlan *d) {
  return Pat{d->cp,       d->ri,        d->op_ptr,    d->op_l,     d->op_u,
             d->diag_slot, d->lrow_ptr, d->lrow_slot, d->lrow_col, d->urow_ptr,
             d->urow_slot, d->urow_col, d->lu_rec,  d->lu_seg,    d->lu_run,
             d->lu_fast,   d->u_rec,    d->u_seg,  d->u_run,     d->u_fast,
             d->l_rec,     d->l_seg,    d->l_run,  d->l_fast,    d->ut_rec,
             d->ut_seg,    d->ut_run,   d->ut_fast, d->lt_rec,   d->lt_seg,
             d->lt_run,    d->lt_fast,  d->drf,    d->status};
}
static Lev lev_of(const DevLevels &L) {
  return Lev{L.level_chain_ptr, L.chain_col_ptr, L.cols,
             L.comp_ptr,        L.comp_lv_begin, L.comp_lv_end,
             L.comp_lv_level};
}

template <int W> __device__ __forceinline__ int uniform(int v) {
  if constexpr (W == 64)
    return __builtin_amdgcn_readfirstlane(v);
  else
    return v;
}

__device__ __forceinline__ void record_status(int *status, int code, int col,
                                              int inst) {
  if (atomicCAS(status, 0, code) == 0) {
    status[1] = col;
    status[2] = inst;
  }
}

// One column of the refactor for one instance (gather form), optionally
// followed by the forward substitution of row c.
// Operation order per entry = GLU_CSC_CPU_sequential's (PV/GLU_CSC_CPU_levels_strings.cpp:32-62):
// v - L*U for ascending j, then the division by the pivot for i > c.
// Operand loads are issued in chunks of kChunk independent loads, then the
// chunk's updates are applied in order: a wide column (the DC bus: 24 updates
// on the diagonal, a 24-term forward row) costs a few memory latencies instead
// of one per update.
constexpr int kChunk = 8;

// One chunk: CH operand pairs loaded unconditionally (indices past the end
// repeat the last pair), then the updates applied in order.  GET(k) returns
// the two operands of update k.
template <int CH, class GET>
__device__ __forceinline__ double chunk_sub(double v, int kb, int k1, GET get) {
  double l[CH], r[CH];
#pragma unroll
  for (int j = 0; j < CH; ++j)
    get(min(kb + j, k1 - 1), l[j], r[j]);
#pragma unroll
  for (int j = 0; j < CH; ++j)
    if (kb + j < k1)
      v = v - l[j] * r[j];
  return v;
}

// v - sum_k l_k r_k over [k0, k1) in ascending k: full chunks of kChunk
// independent loads, then one chunk sized to the remainder, so a sum costs
// ceil(n / kChunk) memory round trips and issues at most 2 n + kChunk loads.
template <class GET>
__device__ __forceinline__ double dot_sub(double v, int k0, int k1, GET get) {
  int kb = k0;
  for (; k1 - kb >= kChunk; kb += kChunk)
    v = chunk_sub<kChunk>(v, kb, k1, get);
  const int rest = k1 - kb;
  if (rest > kChunk / 2)
    v = chunk_sub<kChunk>(v, kb, k1, get);
  else if (rest > 2)
    v = chunk_sub<kChunk / 2>(v, kb, k1, get);
  else if (rest == 2)
    v = chunk_sub<2>(v, kb, k1, get);
  else if (rest == 1)
    v = chunk_sub<1>(v, kb, k1, get);
  return v;
}

// Two sums sharing their first operand: v - sum_k l_k r_k and
// y - sum_k l_k s_k, each in ascending k (the same operation order as two
// separate dot_sub calls), from one set of loads.
template <int CH, class GET>
__device__ __forceinline__ void chunk_sub2(double &v, double &y, int kb, int k1,
                                           GET get) {
  double l[CH], r[CH], t[CH];
#pragma unroll
  for (int j = 0; j < CH; ++j)
    get(min(kb + j, k1 - 1), l[j], r[j], t[j]);
#pragma unroll
  for (int j = 0; j < CH; ++j)
    if (kb + j < k1) {
      v = v - l[j] * r[j];
      y = y - l[j] * t[j];
    }
}

template <class GET>
__device__ __forceinline__ void dot_sub2(double &v, double &y, int k0, int k1,
                                         GET get) {
  int kb = k0;
  for (; k1 - kb >= kChunk; kb += kChunk)
    chunk_sub2<kChunk>(v, y, kb, k1, get);
  const int rest = k1 - kb;
  if (rest > kChunk / 2)
    chunk_sub2<kChunk>(v, y, kb, k1, get);
  else if (rest > 2)
    chunk_sub2<kChunk / 2>(v, y, kb, k1, get);
  else if (rest > 0)
    chunk_sub2<2>(v, y, kb, k1, get);
}

template <int W>
__device__ __forceinline__ double dot_sub_chunked(double v, const int *sl_arr,
                                                  const int *sr_arr,
                                                  const double *L,
                                                  const double *R, int k0,
                                                  int k1, size_t B, int inst) {
  return dot_sub(v, k0, k1, [&](int k, double &l, double &r) {
    l = L[(size_t)uniform<W>(sl_arr[k]) * B + inst];
    r = R[(size_t)uniform<W>(sr_arr[k]) * B + inst];
  });
}

// Same over consecutive CSC slots q of LU against y at the rows ri[q].
template <int W>
__device__ __forceinline__ double dot_sub_csc(double v, const int *ri,
                                              const double *LU,
                                              const double *y, int q0, int q1,
                                              size_t B, int inst) {
  return dot_sub(v, q0, q1, [&](int q, double &l, double &r) {
    l = LU[(size_t)q * B + inst];
    r = y[(size_t)uniform<W>(ri[q]) * B + inst];
  });
}

template <int W, bool FUSE>
__device__ __forceinline__ void refactor_column(const Pat &P, int c,
                                                const double *A, double *LU,
                                                const double *b, double *x,
                                                size_t B, int inst) {
  const int q0 = uniform<W>(P.cp[c]);
  const int q1 = uniform<W>(P.cp[c + 1]);
  const int qd = uniform<W>(P.diag[c]);
  double piv = 1.0;
  bool row_with_diag = false; // forward row computed with the diagonal
  double yrow = 0.0;
  if constexpr (FUSE)
    row_with_diag = uniform<W>(P.drf[c]) != 0;
  // the A values of CH entries are loaded together, then the entries are
  // updated in slot order
  auto entries = [&](auto ch, int qb) {
    constexpr int CH = decltype(ch)::value;
    const int cnt = min(CH, q1 - qb);
    double a[CH];
    int optr[CH + 1]; // the chunk's update ranges, loaded together
#pragma unroll
    for (int k = 0; k < CH; ++k)
      a[k] = A[(size_t)min(qb + k, q1 - 1) * B + inst];
#pragma unroll
    for (int k = 0; k <= CH; ++k)
      optr[k] = uniform<W>(P.op_ptr[min(qb + k, q1)]);
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      if (k >= cnt)
        break;
      const int q = qb + k;
      double v = a[k];
      const int o0 = optr[k], o1 = optr[k + 1];
      if (FUSE && row_with_diag && q == qd) {
        const int r0 = uniform<W>(P.lrow_ptr[c]);
        yrow = b[(size_t)c * B + inst];
        dot_sub2(v, yrow, o0, o1, [&](int o, double &l, double &u, double &yj) {
          l = LU[(size_t)uniform<W>(P.op_l[o]) * B + inst];
          u = LU[(size_t)uniform<W>(P.op_u[o]) * B + inst];
          yj = x[(size_t)uniform<W>(P.lrow_col[r0 + (o - o0)]) * B + inst];
        });
      } else {
        v = dot_sub_chunked<W>(v, P.op_l, P.op_u, LU, LU, o0, o1, B, inst);
      }
      if (q == qd) {
        piv = v;
        if (fabs(v) < 1e-8)
          record_status(P.status, PVGLU_E_SINGULAR, c, inst);
      } else if (q > qd) {
        v = v / piv;
      }
      LU[(size_t)q * B + inst] = v;
    }
  };
  for (int qb = q0; qb < q1; qb += kChunk) {
    if (q1 - qb > kChunk / 2)
      entries(std::integral_constant<int, kChunk>(), qb);
    else
      entries(std::integral_constant<int, kChunk / 2>(), qb);
  }
  if constexpr (FUSE) {
    if (row_with_diag)
      x[(size_t)c * B + inst] = yrow;
    else
      x[(size_t)c * B + inst] = dot_sub_chunked<W>(
          b[(size_t)c * B + inst], P.lrow_slot, P.lrow_col, LU, x,
          uniform<W>(P.lrow_ptr[c]), uniform<W>(P.lrow_ptr[c + 1]), B, inst);
  }
}

// Row i of the forward substitution: y_i = b_i - sum_j L(i,j) y_j (j asc),
// the per-target order of Process_Column_L_CSC_CPU sweeps (PV/GLU_CSC_CPU_Triangular.cpp:3-11,27-33).
template <int W>
__device__ __forceinline__ void lsolve_row(const Pat &P, int i,
                                           const double *LU, const double *b,
                                           double *x, size_t B, int inst) {
  x[(size_t)i * B + inst] = dot_sub_chunked<W>(
      b[(size_t)i * B + inst], P.lrow_slot, P.lrow_col, LU, x,
      uniform<W>(P.lrow_ptr[i]), uniform<W>(P.lrow_ptr[i + 1]), B, inst);
}

// Row i of the backward substitution: x_i = (y_i - sum_j U(i,j) x_j)/U(i,i)
// with j descending, the per-target order of Process_Column_U_CSC_CPU sweeps
// (PV/GLU_CSC_CPU_Triangular.cpp:13-25,35-39).
template <int W>
__device__ __forceinline__ void usolve_row(const Pat &P, int i,
                                           const double *LU, double *x,
                                           size_t B, int inst) {
  const int d = uniform<W>(P.diag[i]);
  const double dv = LU[(size_t)d * B + inst];
  const double s = dot_sub_chunked<W>(
      x[(size_t)i * B + inst], P.urow_slot, P.urow_col, LU, x,
      uniform<W>(P.urow_ptr[i]), uniform<W>(P.urow_ptr[i + 1]), B, inst);
  x[(size_t)i * B + inst] = s / dv;
}

// Adjoint, U^T z = g: z_j = (g_j - sum_{i<j} U(i,j) z_i) / U(j,j), the CSC
// column j above its diagonal (contiguous slots), i ascending.
template <int W>
__device__ __forceinline__ void utsolve_col(const Pat &P, int j,
                                            const double *LU, const double *g,
                                            double *y, size_t B, int inst) {
  const int q0 = uniform<W>(P.cp[j]);
  const int d = uniform<W>(P.diag[j]);
  const double s =
      dot_sub_csc<W>(g[(size_t)j * B + inst], P.ri, LU, y, q0, d, B, inst);
  y[(size_t)j * B + inst] = s / LU[(size_t)d * B + inst];
}

// Adjoint, L^T y = z: y_j = z_j - sum_{i>j} L(i,j) y_i, column j below its
// diagonal, i ascending.
template <int W>
__device__ __forceinline__ void ltsolve_col(const Pat &P, int j,
                                            const double *LU, double *y,
                                            size_t B, int inst) {
  const int d = uniform<W>(P.diag[j]);
  const int q1 = uniform<W>(P.cp[j + 1]);
  y[(size_t)j * B + inst] =
      dot_sub_csc<W>(y[(size_t)j * B + inst], P.ri, LU, y, d + 1, q1, B, inst);
}


// ------------------------------------------------ register-forwarding chains
// Runs of columns with <= 4 entries execute from the 16-int records built by
// compile_fast_programs: the column's values live in registers, operands that
// the previous column of the run produced are forwarded from registers, and
// the A and b values two columns ahead are loaded while the current column
// computes.  A strided segment shares one record: its columns' slots and rows
// are computed, so inside a segment no pattern data is loaded at all.
// Each entry still sees the generic path's operation sequence (bit-identical).

__device__ __forceinline__ double sel4(double v0, double v1, double v2,
                                       double v3, int i) {
  return i == 0 ? v0 : (i == 1 ? v1 : (i == 2 ? v2 : v3));
}

template <int W, int N>
__device__ __forceinline__ void load_record(const int *r, int (&f)[N]) {
  const int4 *r4 = reinterpret_cast<const int4 *>(r);
#pragma unroll
  for (int k = 0; k < N / 4; ++k) {
    const int4 v = r4[k];
    f[4 * k] = uniform<W>(v.x);
    f[4 * k + 1] = uniform<W>(v.y);
    f[4 * k + 2] = uniform<W>(v.z);
    f[4 * k + 3] = uniform<W>(v.w);
  }
}

// A and b values of one refactor column.
struct ColVals {
  double a0, a1, a2, a3, bb;
};

template <bool FUSE>
__device__ __forceinline__ void load_col(ColVals &c, int slot0, int col, int ne,
                                         const double *A, const double *b,
                                         size_t B, int inst) {
  c.a0 = A[(size_t)slot0 * B + inst];
  c.a1 = ne > 1 ? A[(size_t)(slot0 + 1) * B + inst] : 0.0;
  c.a2 = ne > 2 ? A[(size_t)(slot0 + 2) * B + inst] : 0.0;
  c.a3 = ne > 3 ? A[(size_t)(slot0 + 3) * B + inst] : 0.0;
  c.bb = FUSE ? b[(size_t)col * B + inst] : 0.0;
}

// A segment whose operands are all forwarded from the previous column runs
// as a three-stage pipeline with a fixed sequence of memory operations per
// column (the loads of column i+2 are issued before column i computes, stores
// are unconditional in NE, the loop is unrolled three times so no register
// copies of in-flight loads are needed): the compiler's counter-based waits
// then wait for the column being computed only, not for the prefetches.
template <int NE, bool FUSE>
__device__ __forceinline__ void pure_load_col(ColVals &c, int slot0, int col,
                                              const double *A,
                                              const double *b, size_t B,
                                              int inst) {
  c.a0 = A[(size_t)slot0 * B + inst];
  if constexpr (NE > 1)
    c.a1 = A[(size_t)(slot0 + 1) * B + inst];
  if constexpr (NE > 2)
    c.a2 = A[(size_t)(slot0 + 2) * B + inst];
  if constexpr (NE > 3)
    c.a3 = A[(size_t)(slot0 + 3) * B + inst];
  if constexpr (FUSE)
    c.bb = b[(size_t)col * B + inst];
}

template <int NE, bool FUSE>
__device__ __forceinline__ void pure_lu_col(const ColVals &c, int slot0,
                                            int col, const int (&f)[16],
                                            double *LU, double *x, size_t B,
                                            int inst, double &p0, double &p1,
                                            double &p2, double &p3,
                                            double &yprev, int &badcol) {
  const int meta = f[2];
  const int d = (meta >> 3) & 7, nops = (meta >> 6) & 7, nlr = (meta >> 9) & 3;
  double v0 = c.a0, v1 = NE > 1 ? c.a1 : 0.0, v2 = NE > 2 ? c.a2 : 0.0,
         v3 = NE > 3 ? c.a3 : 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (k < nops) {
      const int tu = f[5 + 2 * k];
      const double l = sel4(p0, p1, p2, p3, -1 - f[4 + 2 * k]);
      const double prod = l * sel4(v0, v1, v2, v3, tu >> 2);
      switch (tu & 3) {
      case 0: v0 = v0 - prod; break;
      case 1: v1 = v1 - prod; break;
      case 2: v2 = v2 - prod; break;
      default: v3 = v3 - prod; break;
      }
    }
  }
  const double piv = sel4(v0, v1, v2, v3, d);
  if (fabs(piv) < 1e-8 && badcol < 0)
    badcol = col;
  if (NE > 1 && d < 1)
    v1 = v1 / piv;
  if (NE > 2 && d < 2)
    v2 = v2 / piv;
  if (NE > 3 && d < 3)
    v3 = v3 / piv;
  LU[(size_t)slot0 * B + inst] = v0;
  if constexpr (NE > 1)
    LU[(size_t)(slot0 + 1) * B + inst] = v1;
  if constexpr (NE > 2)
    LU[(size_t)(slot0 + 2) * B + inst] = v2;
  if constexpr (NE > 3)
    LU[(size_t)(slot0 + 3) * B + inst] = v3;
  if constexpr (FUSE) {
    double y = c.bb;
    if (nlr > 0)
      y = y - sel4(p0, p1, p2, p3, -1 - f[12]) * yprev;
    x[(size_t)col * B + inst] = y;
    yprev = y;
  }
  p0 = v0;
  p1 = v1;
  p2 = v2;
  p3 = v3;
}

// Software pipeline over the steps i of a segment: the loads of step i + D
// are issued before step i computes, into a ring of D + 1 register buffers
// indexed by unrolled constants (no copies of registers with loads in
// flight); LD clamps its step to the segment.
template <int D, class T, class LD, class RUN>
__device__ __forceinline__ void pipeline(int len, LD ld, RUN run) {
  T buf[D + 1];
#pragma unroll
  for (int k = 0; k < D; ++k)
    ld(buf[k], k);
  for (int i = 0;;) {
#pragma unroll
    for (int u = 0; u <= D; ++u) {
      ld(buf[(u + D) % (D + 1)], i + D);
      run(buf[u], i);
      if (++i >= len)
        return;
    }
  }
}

// Prefetch depth of the refactor and substitution register pipelines.
constexpr int kLuDepth = 2, kRowDepth = 2;

template <int NE, bool FUSE>
__device__ __forceinline__ void pure_lu_segment(
    const int (&f)[16], int len, int dslot, int dcol, const double *A,
    double *LU, const double *b, double *x, size_t B, int inst, double &p0,
    double &p1, double &p2, double &p3, double &yprev, int &badcol) {
  pipeline<kLuDepth, ColVals>(
      len,
      [&](ColVals &c, int i) {
        i = min(i, len - 1);
        pure_load_col<NE, FUSE>(c, f[0] + i * dslot, f[1] + i * dcol, A, b, B,
                                inst);
      },
      [&](const ColVals &c, int i) {
        pure_lu_col<NE, FUSE>(c, f[0] + i * dslot, f[1] + i * dcol, f, LU, x,
                              B, inst, p0, p1, p2, p3, yprev, badcol);
      });
}

// Backward-substitution values of one row: y, U(i,i) and the two U terms.
struct RowVals {
  double y, dv, u0, u1;
};

// Row sweeps share one register program: x_i = (src_i - sum_j T(i,j) x_j)
// [/ T(i,i) when DIV], T = U (backward, src = x in place, DIV) or unit L
// (forward, src = b).  Records hold slots relative to the row's diagonal slot.
template <int NT, bool DIV>
__device__ __forceinline__ void pure_row_segment(const int (&f)[8], int len,
                                                 int drow, int ddiag,
                                                 const double *LU,
                                                 const double *src, double *x,
                                                 size_t B, int inst,
                                                 double &xprev) {
  pipeline<kRowDepth, RowVals>(
      len,
      [&](RowVals &r, int i) {
        i = min(i, len - 1);
        const int row = f[0] + i * drow, dslot = f[1] + i * ddiag;
        r.y = src[(size_t)row * B + inst];
        if constexpr (DIV)
          r.dv = LU[(size_t)dslot * B + inst];
        if constexpr (NT > 0)
          r.u0 = LU[(size_t)(dslot + f[4]) * B + inst];
      },
      [&](const RowVals &r, int i) {
        double s = r.y;
        if constexpr (NT > 0)
          s = s - r.u0 * xprev;
        const double xi = DIV ? s / r.dv : s;
        x[(size_t)(f[0] + i * drow) * B + inst] = xi;
        xprev = xi;
      });
}

template <int W, bool FUSE>
__device__ __forceinline__ void fast_lu_chain(const Pat &P, int t0, int t1,
                                              const double *A, double *LU,
                                              const double *b, double *x,
                                              size_t B, int inst) {
  double p0 = 0, p1 = 0, p2 = 0, p3 = 0, yprev = 0;
  for (int t = t0; t < t1;) {
    const int *sg = P.lu_seg + (size_t)t * 4;
    const int len = uniform<W>(sg[0]);
    const int dslot = uniform<W>(sg[1]);
    const int dcol = uniform<W>(sg[2]);
    int f[16];
    load_record<W>(P.lu_rec + (size_t)t * 16, f);
    const int meta = f[2];
    const int ne = meta & 7, d = (meta >> 3) & 7, nops = (meta >> 6) & 7,
              nlr = (meta >> 9) & 3;
    if (uniform<W>(sg[3]) & (FUSE ? 1 : 2)) {
      int badcol = -1;
      switch (ne) {
      case 1:
        pure_lu_segment<1, FUSE>(f, len, dslot, dcol, A, LU, b, x, B, inst, p0, p1,
                           p2, p3, yprev, badcol);
        break;
      case 2:
        pure_lu_segment<2, FUSE>(f, len, dslot, dcol, A, LU, b, x, B, inst, p0, p1,
                           p2, p3, yprev, badcol);
        break;
      case 3:
        pure_lu_segment<3, FUSE>(f, len, dslot, dcol, A, LU, b, x, B, inst, p0, p1,
                           p2, p3, yprev, badcol);
        break;
      default:
        pure_lu_segment<4, FUSE>(f, len, dslot, dcol, A, LU, b, x, B, inst, p0, p1,
                           p2, p3, yprev, badcol);
        break;
      }
      if (badcol >= 0)
        record_status(P.status, PVGLU_E_SINGULAR, badcol, inst);
      t += len;
      continue;
    }
    ColVals c0, c1, c2;
    load_col<FUSE>(c0, f[0], f[1], ne, A, b, B, inst);
    if (len > 1)
      load_col<FUSE>(c1, f[0] + dslot, f[1] + dcol, ne, A, b, B, inst);
    for (int i = 0; i < len; ++i) {
      if (i + 2 < len)
        load_col<FUSE>(c2, f[0] + (i + 2) * dslot, f[1] + (i + 2) * dcol, ne,
                       A, b, B, inst);
      const int slot0 = f[0] + i * dslot, col = f[1] + i * dcol;
      double v0 = c0.a0, v1 = c0.a1, v2 = c0.a2, v3 = c0.a3;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (k < nops) {
          const int lsrc = f[4 + 2 * k], tu = f[5 + 2 * k];
          const double l = lsrc >= 0 ? LU[(size_t)lsrc * B + inst]
                                     : sel4(p0, p1, p2, p3, -1 - lsrc);
          const double prod = l * sel4(v0, v1, v2, v3, tu >> 2);
          switch (tu & 3) {
          case 0: v0 = v0 - prod; break;
          case 1: v1 = v1 - prod; break;
          case 2: v2 = v2 - prod; break;
          default: v3 = v3 - prod; break;
          }
        }
      }
      const double piv = sel4(v0, v1, v2, v3, d);
      if (fabs(piv) < 1e-8)
        record_status(P.status, PVGLU_E_SINGULAR, col, inst);
      if (d < 1 && ne > 1)
        v1 = v1 / piv;
      if (d < 2 && ne > 2)
        v2 = v2 / piv;
      if (d < 3 && ne > 3)
        v3 = v3 / piv;
      LU[(size_t)slot0 * B + inst] = v0;
      if (ne > 1)
        LU[(size_t)(slot0 + 1) * B + inst] = v1;
      if (ne > 2)
        LU[(size_t)(slot0 + 2) * B + inst] = v2;
      if (ne > 3)
        LU[(size_t)(slot0 + 3) * B + inst] = v3;
      if constexpr (FUSE) { // forward substitution of row col
        double y = c0.bb;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          if (k < nlr) {
            const int lsrc = f[12 + 2 * k], ysrc = f[13 + 2 * k];
            const double l = lsrc >= 0 ? LU[(size_t)lsrc * B + inst]
                                       : sel4(p0, p1, p2, p3, -1 - lsrc);
            const double yj = ysrc >= 0 ? x[(size_t)ysrc * B + inst] : yprev;
            y = y - l * yj;
          }
        }
        x[(size_t)col * B + inst] = y;
        yprev = y;
      }
      p0 = v0;
      p1 = v1;
      p2 = v2;
      p3 = v3;
      c0 = c1;
      c1 = c2;
    }
    t += len;
  }
}

template <bool DIV>
__device__ __forceinline__ void load_row(RowVals &v, int row, int dslot, int nt,
                                         int rel0, int rel1, const double *LU,
                                         const double *src, size_t B,
                                         int inst) {
  v.y = src[(size_t)row * B + inst];
  v.dv = DIV ? LU[(size_t)dslot * B + inst] : 1.0;
  v.u0 = nt > 0 ? LU[(size_t)(dslot + rel0) * B + inst] : 0.0;
  v.u1 = nt > 1 ? LU[(size_t)(dslot + rel1) * B + inst] : 0.0;
}

template <int W, bool DIV>
__device__ __forceinline__ void fast_row_chain(const int *rec, const int *seg,
                                               int t0, int t1,
                                               const double *LU,
                                               const double *src, double *x,
                                               size_t B, int inst) {
  double xprev = 0;
  for (int t = t0; t < t1;) {
    const int *sg = seg + (size_t)t * 4;
    const int len = uniform<W>(sg[0]);
    const int drow = uniform<W>(sg[1]);
    const int ddiag = uniform<W>(sg[2]);
    int f[8];
    load_record<W>(rec + (size_t)t * 8, f);
    const int nt = f[2];
    if (uniform<W>(sg[3])) {
      if (nt == 0)
        pure_row_segment<0, DIV>(f, len, drow, ddiag, LU, src, x, B, inst,
                                 xprev);
      else
        pure_row_segment<1, DIV>(f, len, drow, ddiag, LU, src, x, B, inst,
                                 xprev);
      t += len;
      continue;
    }
    RowVals r0, r1, r2;
    load_row<DIV>(r0, f[0], f[1], nt, f[4], f[6], LU, src, B, inst);
    if (len > 1)
      load_row<DIV>(r1, f[0] + drow, f[1] + ddiag, nt, f[4], f[6], LU, src, B,
                    inst);
    for (int i = 0; i < len; ++i) {
      if (i + 2 < len)
        load_row<DIV>(r2, f[0] + (i + 2) * drow, f[1] + (i + 2) * ddiag, nt,
                      f[4], f[6], LU, src, B, inst);
      const int row = f[0] + i * drow;
      double s = r0.y;
      if (nt > 0)
        s = s - r0.u0 * (f[5] >= 0 ? x[(size_t)f[5] * B + inst] : xprev);
      if (nt > 1)
        s = s - r0.u1 * (f[7] >= 0 ? x[(size_t)f[7] * B + inst] : xprev);
      const double xi = DIV ? s / r0.dv : s;
      x[(size_t)row * B + inst] = xi;
      xprev = xi;
      r0 = r1;
      r1 = r2;
    }
    t += len;
  }
}

enum Op { OP_REFACTOR = 0, OP_REFACTOR_FUSED, OP_LSOLVE, OP_USOLVE, OP_UT, OP_LT };

template <int W, int OP>
__device__ __forceinline__ void run_chain(const Pat &P, const Lev &L, int chain,
                                          const double *A, double *LU,
                                          const double *b, double *x, size_t B,
                                          int inst) {
  const int t0 = uniform<W>(L.chain_col_ptr[chain]);
  const int t1 = uniform<W>(L.chain_col_ptr[chain + 1]);
  if constexpr (OP == OP_REFACTOR_FUSED || OP == OP_REFACTOR) {
    constexpr bool FUSE = OP == OP_REFACTOR_FUSED;
    if (uniform<W>(P.lu_fast[chain])) {
      fast_lu_chain<W, FUSE>(P, t0, t1, A, LU, b, x, B, inst);
      return;
    }
    // mixed chain: register runs between generic columns
    for (int t = t0; t < t1;) {
      const int e = uniform<W>(P.lu_run[t]);
      if (e > t) {
        fast_lu_chain<W, FUSE>(P, t, e, A, LU, b, x, B, inst);
        t = e;
      } else {
        refactor_column<W, FUSE>(P, uniform<W>(L.cols[t]), A, LU, b, x, B,
                                 inst);
        ++t;
      }
    }
    return;
  }
  if constexpr (OP == OP_USOLVE || OP == OP_LSOLVE || OP == OP_UT ||
                OP == OP_LT) {
    // row programs: backward (U), forward (L), adjoint U^T (divides) and L^T
    constexpr bool DIV = OP == OP_USOLVE || OP == OP_UT;
    const int *rec, *seg, *run, *whole;
    if constexpr (OP == OP_USOLVE) {
      rec = P.u_rec, seg = P.u_seg, run = P.u_run, whole = P.u_fast;
    } else if constexpr (OP == OP_LSOLVE) {
      rec = P.l_rec, seg = P.l_seg, run = P.l_run, whole = P.l_fast;
    } else if constexpr (OP == OP_UT) {
      rec = P.ut_rec, seg = P.ut_seg, run = P.ut_run, whole = P.ut_fast;
    } else {
      rec = P.lt_rec, seg = P.lt_seg, run = P.lt_run, whole = P.lt_fast;
    }
    // source of the right-hand side: forward b, U^T g; in place otherwise
    const double *src = (OP == OP_LSOLVE || OP == OP_UT) ? b : x;
    if (uniform<W>(whole[chain])) {
      fast_row_chain<W, DIV>(rec, seg, t0, t1, LU, src, x, B, inst);
      return;
    }
    for (int t = t0; t < t1;) {
      const int e = uniform<W>(run[t]);
      if (e > t) {
        fast_row_chain<W, DIV>(rec, seg, t, e, LU, src, x, B, inst);
        t = e;
      } else {
        const int c = uniform<W>(L.cols[t]);
        if constexpr (OP == OP_USOLVE)
          usolve_row<W>(P, c, LU, x, B, inst);
        else if constexpr (OP == OP_LSOLVE)
          lsolve_row<W>(P, c, LU, b, x, B, inst);
        else if constexpr (OP == OP_UT)
          utsolve_col<W>(P, c, LU, b, x, B, inst);
        else
          ltsolve_col<W>(P, c, LU, x, B, inst);
        ++t;
      }
    }
    return;
  }
  for (int t = t0; t < t1; ++t) {
    const int c = uniform<W>(L.cols[t]);
    if constexpr (OP == OP_REFACTOR)
      refactor_column<W, false>(P, c, A, LU, b, x, B, inst);
    else if constexpr (OP == OP_REFACTOR_FUSED)
      refactor_column<W, true>(P, c, A, LU, b, x, B, inst);
    else if constexpr (OP == OP_LSOLVE)
      lsolve_row<W>(P, c, LU, b, x, B, inst);
    else if constexpr (OP == OP_USOLVE)
      usolve_row<W>(P, c, LU, x, B, inst);
    else if constexpr (OP == OP_UT)
      utsolve_col<W>(P, c, LU, b, x, B, inst);
    else
      ltsolve_col<W>(P, c, LU, x, B, inst);
  }
}

// LEVEL schedule: one launch per level.  Thread -> (chain group, lane);
// group -> (instance tile, chain of the level).
template <int W, int OP>
__global__ void __launch_bounds__(256)
    k_level(Pat P, Lev L, int first_chain, int nchains, int tiles,
            const double *A, double *LU, const double *b, double *x, int B) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int group = uniform<W>(gid / W);
  const int lane = gid % W;
  if (group >= nchains * tiles)
    return;
  // consecutive wavefronts take consecutive instance tiles of one chain, so
  // the tiles of a slot row are streamed together
  const int tile = group % tiles;
  const int chain = first_chain + group / tiles;
  const int inst = tile * W + lane;
  if (inst >= B)
    return;
  run_chain<W, OP>(P, L, chain, A, LU, b, x, (size_t)B, inst);
}

// COMPONENT schedule: one launch for the whole refactor + solve.  Block ->
// (component, instance tile); each level of the component is split over the
// block's chain slots; barriers separate levels.  Every thread reaches every
// barrier (loop bounds are block-uniform; inactive lanes only skip work).
template <int W, bool FUSE>
__device__ __forceinline__ void component_levels(const Pat &P, const Lev &L,
                                                 int comp, int slot, int slots,
                                                 bool active, const double *A,
                                                 double *LU, const double *b,
                                                 double *x, size_t B, int inst,
                                                 int op, int lv_lo = 0,
                                                 int lv_hi = 1 << 30) {
  const int e0 = L.comp_ptr[comp], e1 = L.comp_ptr[comp + 1];
  for (int e = e0; e < e1; ++e) {
    const int lv = L.comp_lv_level[e];
    if (lv < lv_lo || lv >= lv_hi)
      continue; // block-uniform: no work, no barrier
    const int c0 = L.comp_lv_begin[e], c1 = L.comp_lv_end[e];
    if (active)
      for (int ch = c0 + slot; ch < c1; ch += slots) {
        const int chain = uniform<W>(ch);
        if (op == OP_REFACTOR_FUSED)
          run_chain<W, OP_REFACTOR_FUSED>(P, L, chain, A, LU, b, x, B, inst);
        else if (op == OP_REFACTOR)
          run_chain<W, OP_REFACTOR>(P, L, chain, A, LU, b, x, B, inst);
        else if (op == OP_LSOLVE)
          run_chain<W, OP_LSOLVE>(P, L, chain, A, LU, b, x, B, inst);
        else if (op == OP_USOLVE)
          run_chain<W, OP_USOLVE>(P, L, chain, A, LU, b, x, B, inst);
        else if (op == OP_UT)
          run_chain<W, OP_UT>(P, L, chain, A, LU, b, x, B, inst);
        else
          run_chain<W, OP_LT>(P, L, chain, A, LU, b, x, B, inst);
      }
    __syncthreads();
  }
}

// mode 0: refactor only; 1: refactor + solve; 2: solve only; 3: transposed solve
template <int W>
__device__ __forceinline__ void component_body(const Pat &P, const Lev &Llu,
                                               const Lev &Ll, const Lev &Lu,
                                               const Lev &Llt, int mode,
                                               bool fuse, int comp, int tile,
                                               const double *A, double *LU,
                                               const double *b, double *x,
                                               int B) {
  const int slots = blockDim.x / W;
  const int slot = uniform<W>((int)threadIdx.x / W);
  const int lane = threadIdx.x % W;
  const int inst = tile * W + lane;
  const bool active = inst < B;
  const size_t Bs = (size_t)B;
  if (mode == 3) {
    component_levels<W, false>(P, Llu, comp, slot, slots, active, A, LU, b, x,
                               Bs, inst, OP_UT);
    component_levels<W, false>(P, Llt, comp, slot, slots, active, A, LU, b, x,
                               Bs, inst, OP_LT);
    return;
  }
  if (mode <= 1)
    component_levels<W, false>(P, Llu, comp, slot, slots, active, A, LU, b, x,
                               Bs, inst,
                               (mode == 1 && fuse) ? OP_REFACTOR_FUSED
                                                   : OP_REFACTOR);
  if (mode == 0)
    return;
  if (mode == 2 || !fuse)
    component_levels<W, false>(P, Ll, comp, slot, slots, active, A, LU, b, x,
                               Bs, inst, OP_LSOLVE);
  component_levels<W, false>(P, Lu, comp, slot, slots, active, A, LU, b, x, Bs,
                             inst, OP_USOLVE);
}

template <int W>
__global__ void __launch_bounds__(1024)
    k_component(Pat P, Lev Llu, Lev Ll, Lev Lu, Lev Llt, int mode, bool fuse,
                int tiles, const double *A, double *LU, const double *b,
                double *x, int B) {
  component_body<W>(P, Llu, Ll, Lu, Llt, mode, fuse, blockIdx.x / tiles,
                    blockIdx.x % tiles, A, LU, b, x, B);
}

// HYBRID schedule: a run of narrow levels (few chains) in one launch.  Block
// -> (component, tile); the four level sets are walked in order
// (refactor, forward, backward, transposed-L) over the given level ranges,
// each with its operation, barriers between levels.
struct LvRange {
  int lo[4], hi[4], op[4];
};

template <int W>
__global__ void __launch_bounds__(1024)
    k_hybrid(Pat P, Lev Llu, Lev Ll, Lev Lu, Lev Llt, LvRange R, int comp0,
             int tiles, const double *A, double *LU, const double *b,
             double *x, int B) {
  const int comp = comp0 + blockIdx.x / tiles, tile = blockIdx.x % tiles;
  const int slots = blockDim.x / W;
  const int slot = uniform<W>((int)threadIdx.x / W);
  const int inst = tile * W + (int)threadIdx.x % W;
  const bool active = inst < B;
  const Lev *sets[4] = {&Llu, &Ll, &Lu, &Llt};
  for (int k = 0; k < 4; ++k)
    if (R.lo[k] < R.hi[k])
      component_levels<W, false>(P, *sets[k], comp, slot, slots, active, A, LU,
                                 b, x, (size_t)B, inst, R.op[k], R.lo[k],
                                 R.hi[k]);
}

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

template <int W>
__global__ void __launch_bounds__(1024) k_multi(MultiArgs m, int mode) {
  int g = 0;
  while (g + 1 < m.ngroups && (int)blockIdx.x >= m.block_start[g + 1])
    ++g;
  const int local = blockIdx.x - m.block_start[g];
  component_body<W>(m.pat[g], m.lu[g], m.l[g], m.u[g], m.lt[g], mode,
                    m.fuse[g] != 0, local / m.tiles[g], local % m.tiles[g],
                    m.a[g], m.luv[g], m.b[g], m.x[g], m.B[g]);
}

__global__ void k_fill_int(int *p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = v;
}

// ----------------------------------------------------------- 011 update
// One thread per (element chunk, instance); Σ|Δ| is accumulated with atomics
// (order-dependent in the last bits; the maximum is exact).  Per element the
// damping factor
// halves until |candidate - old| <= 1 V and nothing is NaN
// (is_value_change_justified, PV/011_calculate_new_iteration.cpp:6-17); the
// reference's loop halves each unjustified element's factor independently and
// aborts once any factor is below 1e-13 (:44-74).
__global__ void __launch_bounds__(256)
    k_newton_update(int ne, int B, int per_thread, const int *__restrict__ pos,
                    const int *__restrict__ neg, const double *__restrict__ x,
                    double *__restrict__ v, double *delta_sum,
                    unsigned long long *delta_max, const int *active,
                    int *status) {
  // block -> (element chunk, tile of up to 256 instances)
  const int chunk = (int)blockIdx.x;
  const int in = (int)(blockIdx.y * blockDim.x + threadIdx.x);
  if (in >= B)
    return;
  if (active && !active[in])
    return; // converged or stalled instance: frozen
  const int e0 = chunk * per_thread;
  const int e1 = min(ne, e0 + per_thread);
  double sum = 0.0, mx = 0.0;
  for (int e = e0; e < e1; ++e) {
    const int p = pos[e], m_ = neg[e];
    const double vp = p >= 0 ? x[(size_t)p * B + in] : 0.0;
    const double vn = m_ >= 0 ? x[(size_t)m_ * B + in] : 0.0;
    const double nw = vp - vn; // voltage_drop_dc (PV/002_base_electric_system.cpp:21-33)
    const double old = v[(size_t)e * B + in];
    double m = 1.0, cand = old;
    bool ok = false;
    for (int k = 0; k < 64; ++k) {
      if (m < 1e-13)
        break;
      cand = old * (1.0 - m) + nw * m;
      if (!isnan(old) && !isnan(cand) && !(fabs(old - cand) > 1.0)) {
        ok = true;
        break;
      }
      m = m / 2.0;
    }
    if (!ok) {
      record_status(status, PVGLU_E_NOCONV, e, in);
      cand = old;
    }
    const double d = fabs(cand - old);
    sum += d;
    mx = fmax(mx, d);
    v[(size_t)e * B + in] = cand;
  }
  atomicAdd(&delta_sum[in], sum);
  atomicMax(&delta_max[in], (unsigned long long)__double_as_longlong(mx));
}

// -------------------------------------------------------------- host side

namespace pvglu {

template <typename T>
static int dev_upload(DevicePlan *d, T **dst, const std::vector<T> &src) {
  size_t bytes = sizeof(T) * std::max<size_t>(src.size(), 1);
  void *p = nullptr;
  HIP_TRY(hipMalloc(&p, bytes));
  d->allocs.push_back(p);
  if (!src.empty())
    HIP_TRY(hipMemcpy(p, src.data(), sizeof(T) * src.size(),
                      hipMemcpyHostToDevice));
  *dst = (T *)p;
  return PVGLU_OK;
}

static int upload_levels(DevicePlan *d, const KernelLevels &k, DevLevels &L) {
  int st;
  if ((st = dev_upload(d, &L.level_chain_ptr, k.level_chain_ptr)) ||
      (st = dev_upload(d, &L.chain_col_ptr, k.chain_col_ptr)) ||
      (st = dev_upload(d, &L.cols, k.cols)) ||
      (st = dev_upload(d, &L.comp_ptr, k.comp_ptr)) ||
      (st = dev_upload(d, &L.comp_lv_begin, k.comp_lv_begin)) ||
      (st = dev_upload(d, &L.comp_lv_end, k.comp_lv_end)) ||
      (st = dev_upload(d, &L.comp_lv_level, k.comp_lv_level)))
    return st;
  L.nlev = (int)k.level_chain_ptr.size() - 1;
  L.max_comp_chains = k.max_comp_chains;
  L.h_level_chain_ptr = k.level_chain_ptr;
  L.h_level_comp_chains.assign(std::max(L.nlev, 0), 0);
  for (size_t e = 0; e < k.comp_lv_level.size(); ++e) {
    int &m = L.h_level_comp_chains[k.comp_lv_level[e]];
    m = std::max(m, k.comp_lv_end[e] - k.comp_lv_begin[e]);
  }
  return PVGLU_OK;
}

void device_plan_destroy(pvglu_plan *plan) {
  if (!plan->dev)
    return;
  int cur = 0;
  bool switched = false;
  if (hipGetDevice(&cur) == hipSuccess && cur != plan->dev->device) {
    switched = hipSetDevice(plan->dev->device) == hipSuccess;
  }
  for (void *p : plan->dev->allocs)
    (void)hipFree(p);
  DevicePlan *d = plan->dev;
  for (int g = 0; g < 3; ++g) {
    if (d->side[g])
      (void)hipStreamDestroy(d->side[g]);
    if (d->ev_skew[g])
      (void)hipEventDestroy(d->ev_skew[g]);
    if (d->ev_join[g])
      (void)hipEventDestroy(d->ev_join[g]);
  }
  if (d->ev_fork)
    (void)hipEventDestroy(d->ev_fork);
  if (switched)
    (void)hipSetDevice(cur);
  delete plan->dev;
  plan->dev = nullptr;
}

} // namespace pvglu

static int require_device(pvglu_plan *P) {
  if (!P)
    return fail(PVGLU_E_ARG, "null plan");
  if (!P->dev) {
    int st = pvglu_plan_upload(P);
    if (st)
      return st;
  }
  return PVGLU_OK;
}

extern "C" int pvglu_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess)
    return 0;
  return n;
}

extern "C" const char *pvglu_version(void) { return "pvglu 1.0 gfx950"; }

// Component groups: G runs of consecutive components with about n / G
// columns each; bounds[g] = first component of group g.
static std::vector<int> component_bounds(const pvglu_plan *P, int G) {
  std::vector<int> cols(P->ncomp, 0);
  for (int c : P->comp_of)
    cols[c]++;
  std::vector<int> b(G + 1, P->ncomp);
  b[0] = 0;
  long acc = 0;
  int g = 1;
  for (int c = 0; c < P->ncomp && g < G; ++c) {
    acc += cols[c];
    while (g < G && acc * G >= (long)P->n * g)
      b[g++] = c + 1;
  }
  for (int k = 1; k <= G; ++k)
    b[k] = std::max(b[k], b[k - 1]);
  return b;
}

// Chain ranges of each group at each level.  Chains of a level are ordered
// by component, so a group of consecutive components owns a contiguous range.
static void group_ranges(const KernelLevels &k, const std::vector<int> &bounds,
                         int nlev, std::vector<int> &out) {
  const int G = (int)bounds.size() - 1;
  out.assign((size_t)G * std::max(nlev, 0) * 2, 0);
  for (int g = 0; g < G; ++g)
    for (int lv = 0; lv < nlev; ++lv) {
      int lo = 1 << 30, hi = -1;
      for (int c = bounds[g]; c < bounds[g + 1]; ++c)
        for (int e = k.comp_ptr[c]; e < k.comp_ptr[c + 1]; ++e)
          if (k.comp_lv_level[e] == lv) {
            lo = std::min(lo, k.comp_lv_begin[e]);
            hi = std::max(hi, k.comp_lv_end[e]);
          }
      if (hi > lo) {
        out[((size_t)g * nlev + lv) * 2] = lo;
        out[((size_t)g * nlev + lv) * 2 + 1] = hi;
      }
    }
}

static int setup_groups(pvglu_plan *P) {
  DevicePlan *d = P->dev;
  const KernelLevels *ks[4] = {&P->k_lu, &P->k_l, &P->k_u, &P->k_lt};
  DevLevels *ls[4] = {&d->lu, &d->l, &d->u, &d->lt};
  for (int gi = 0; gi < 3; ++gi) {
    d->comp_bounds[gi] = component_bounds(P, 1 << gi);
    for (int k = 0; k < 4; ++k)
      group_ranges(*ks[k], d->comp_bounds[gi], ls[k]->nlev, ls[k]->h_grp[gi]);
  }
  if (P->ncomp < 2)
    return PVGLU_OK;
  for (int g = 0; g < 3; ++g)
    HIP_TRY(hipStreamCreateWithFlags(&d->side[g], hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&d->ev_fork, hipEventDisableTiming));
  for (int g = 0; g < 3; ++g) {
    HIP_TRY(hipEventCreateWithFlags(&d->ev_skew[g], hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&d->ev_join[g], hipEventDisableTiming));
  }
  d->side_ok = true;
  return PVGLU_OK;
}

extern "C" int pvglu_plan_upload(pvglu_plan *P) {
  clear_error();
  if (!P)
    return fail(PVGLU_E_ARG, "plan_upload: null plan");
  if (P->dev)
    return PVGLU_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(PVGLU_E_HIP, "plan_upload: no HIP device");
  auto *d = new DevicePlan();
  HIP_TRY(hipGetDevice(&d->device));
  P->dev = d;
  int st;
  if ((st = dev_upload(d, &d->cp, P->cp)) || (st = dev_upload(d, &d->ri, P->ri)) ||
      (st = dev_upload(d, &d->op_ptr, P->op_ptr)) ||
      (st = dev_upload(d, &d->op_l, P->op_l)) ||
      (st = dev_upload(d, &d->op_u, P->op_u)) ||
      (st = dev_upload(d, &d->diag_slot, P->diag_slot)) ||
      (st = dev_upload(d, &d->lrow_ptr, P->lrow_ptr)) ||
      (st = dev_upload(d, &d->lrow_slot, P->lrow_slot)) ||
      (st = dev_upload(d, &d->lrow_col, P->lrow_col)) ||
      (st = dev_upload(d, &d->urow_ptr, P->urow_ptr)) ||
      (st = dev_upload(d, &d->urow_slot, P->urow_slot)) ||
      (st = dev_upload(d, &d->urow_col, P->urow_col)) ||
      (st = upload_levels(d, P->k_lu, d->lu)) ||
      (st = upload_levels(d, P->k_l, d->l)) ||
      (st = upload_levels(d, P->k_u, d->u)) ||
      (st = upload_levels(d, P->k_lt, d->lt)) ||
      (st = dev_upload(d, &d->lu_rec, P->fast_lu_rec)) ||
      (st = dev_upload(d, &d->lu_fast, P->fast_lu_chain)) ||
      (st = dev_upload(d, &d->lu_run, P->fast_lu_run)) ||
      (st = dev_upload(d, &d->lu_seg, P->fast_lu_seg)) ||
      (st = dev_upload(d, &d->u_seg, P->fast_u_seg)) ||
      (st = dev_upload(d, &d->u_run, P->fast_u_run)) ||
      (st = dev_upload(d, &d->u_rec, P->fast_u_rec)) ||
      (st = dev_upload(d, &d->u_fast, P->fast_u_chain)) ||
      (st = dev_upload(d, &d->l_rec, P->fast_l_rec)) ||
      (st = dev_upload(d, &d->l_seg, P->fast_l_seg)) ||
      (st = dev_upload(d, &d->l_run, P->fast_l_run)) ||
      (st = dev_upload(d, &d->l_fast, P->fast_l_chain)) ||
      (st = dev_upload(d, &d->ut_rec, P->fast_ut_rec)) ||
      (st = dev_upload(d, &d->ut_seg, P->fast_ut_seg)) ||
      (st = dev_upload(d, &d->ut_run, P->fast_ut_run)) ||
      (st = dev_upload(d, &d->ut_fast, P->fast_ut_chain)) ||
      (st = dev_upload(d, &d->lt_rec, P->fast_lt_rec)) ||
      (st = dev_upload(d, &d->lt_seg, P->fast_lt_seg)) ||
      (st = dev_upload(d, &d->lt_run, P->fast_lt_run)) ||
      (st = dev_upload(d, &d->lt_fast, P->fast_lt_chain)) ||
      (st = dev_upload(d, &d->drf, P->diag_row_fused)) ||
      (st = dev_upload(d, &d->status, std::vector<int>(4, 0))) ||
      (st = setup_groups(P))) {
    device_plan_destroy(P);
    return st;
  }
  return PVGLU_OK;
}

// ------------------------------------------------------------- launching

static bool valid_width(int w) {
  return w == 1 || w == 2 || w == 4 || w == 8 || w == 16 || w == 32 || w == 64;
}

static void pick_default_schedule(pvglu_plan *P, int B, int &variant, int &W,
                                  int &threads) {
  variant = P->sched_variant;
  W = P->sched_width;
  if (!valid_width(W))
    W = B >= 64 ? 64 : (B >= 16 ? 16 : (B >= 4 ? 4 : 1));
  if (variant == PVGLU_SCHED_AUTO)
    variant = (P->ncomp >= 8 && B < 256) ? PVGLU_SCHED_COMPONENT
                                         : PVGLU_SCHED_HYBRID;
  // COMPONENT: chain slots per block = widest component level, capped so the
  // block has at most 1024 threads.
  int maxc = std::max({P->k_lu.max_comp_chains, P->k_u.max_comp_chains,
                       P->k_l.max_comp_chains, P->k_lt.max_comp_chains, 1});
  int slots = 1;
  while (slots < maxc && slots * 2 * W <= 1024)
    slots *= 2;
  threads = std::max(64, slots * W);
  if (threads > 1024)
    threads = 1024;
}

// One level (or the chain range of one component group of it) with the
// operation chosen at run time.
template <int W>
static void launch_one_level(pvglu_plan *P, const DevLevels &L, int first,
                             int nch, int op, const double *A, double *LU,
                             const double *b, double *x, int B,
                             hipStream_t s) {
  const int tiles = (B + W - 1) / W;
  const long threads = (long)nch * tiles * W;
  const int blocks = (int)((threads + 255) / 256);
  if (blocks == 0)
    return;
  const Pat pat = pat_of(P->dev);
  const Lev lev = lev_of(L);
#define PVGLU_LVL(OPC)                                                         \
  hipLaunchKernelGGL((k_level<W, OPC>), dim3(blocks), dim3(256), 0, s, pat,    \
                     lev, first, nch, tiles, A, LU, b, x, B)
  switch (op) {
  case OP_REFACTOR: PVGLU_LVL(OP_REFACTOR); break;
  case OP_REFACTOR_FUSED: PVGLU_LVL(OP_REFACTOR_FUSED); break;
  case OP_LSOLVE: PVGLU_LVL(OP_LSOLVE); break;
  case OP_USOLVE: PVGLU_LVL(OP_USOLVE); break;
  case OP_UT: PVGLU_LVL(OP_UT); break;
  default: PVGLU_LVL(OP_LT); break;
  }
#undef PVGLU_LVL
}

// Chains per component at or below which a level counts as narrow.
constexpr int kNarrowChains = 4;

static int group_index(int G) { return G >= 4 ? 2 : (G == 2 ? 1 : 0); }

// LEVEL and HYBRID launches.  The independent components are split into G
// groups of consecutive components (G = 1, 2, 4); group g's launches go to
// stream g (the caller's stream for g = 0, the plan's side streams
// otherwise), forked from and joined back into the caller's stream by
// events.  Group g+1 starts after group g's first level, so one group's
// latency-bound narrow levels run while the next group's wide levels use
// the memory bandwidth.  HYBRID merges each run of two or more narrow steps
// of a group into one component-style launch.
template <int W>
static int launch_steps_w(pvglu_plan *P, int mode, bool hybrid, int G,
                          const double *A, double *LU, const double *b,
                          double *x, int B, hipStream_t s) {
  DevicePlan *d = P->dev;
  if (G > 1 && !d->side_ok)
    G = 1;
  const int gi = group_index(G);
  G = 1 << gi;
  const DevLevels *sets[4] = {&d->lu, &d->l, &d->u, &d->lt};
  int ops[4] = {-1, -1, -1, -1};
  const bool fused = mode == 1 && P->fuse_lsolve;
  if (mode == 3) {
    ops[0] = OP_UT;
    ops[3] = OP_LT;
  } else {
    if (mode <= 1)
      ops[0] = fused ? OP_REFACTOR_FUSED : OP_REFACTOR;
    if (mode == 2 || (mode == 1 && !fused))
      ops[1] = OP_LSOLVE;
    if (mode >= 1)
      ops[2] = OP_USOLVE;
  }
  struct Step {
    int set, lv, width;
  };
  std::vector<Step> steps;
  for (int k = 0; k < 4; ++k)
    if (ops[k] >= 0)
      for (int lv = 0; lv < sets[k]->nlev; ++lv)
        steps.push_back({k, lv, sets[k]->h_level_comp_chains[lv]});
  const int nsteps = (int)steps.size();
  const int tiles = (B + W - 1) / W;
  const Pat pat = pat_of(d);
  if (G > 1) {
    HIP_TRY(hipEventRecord(d->ev_fork, s));
    for (int g = 1; g < G; ++g)
      HIP_TRY(hipStreamWaitEvent(d->side[g - 1], d->ev_fork, 0));
  }
  for (int g = 0; g < G; ++g) {
    hipStream_t st = g == 0 ? s : d->side[g - 1];
    if (g > 0)
      HIP_TRY(hipStreamWaitEvent(st, d->ev_skew[g - 1], 0));
    const int c0 = d->comp_bounds[gi][g], c1 = d->comp_bounds[gi][g + 1];
    bool skewed = false;
    for (int i = 0; i < nsteps;) {
      int j = i, widest = 1;
      if (hybrid)
        while (j < nsteps && steps[j].width <= kNarrowChains) {
          widest = std::max(widest, steps[j].width);
          ++j;
        }
      if (j - i < 2) {
        const Step &sp = steps[i];
        const DevLevels &L = *sets[sp.set];
        const int *r = &L.h_grp[gi][((size_t)g * L.nlev + sp.lv) * 2];
        launch_one_level<W>(P, L, r[0], r[1] - r[0], ops[sp.set], A, LU, b, x,
                            B, st);
        ++i;
      } else {
        LvRange R;
        for (int k = 0; k < 4; ++k) {
          R.lo[k] = 0;
          R.hi[k] = 0;
          R.op[k] = ops[k] < 0 ? 0 : ops[k];
        }
        for (int t = i; t < j; ++t) {
          const Step &sp = steps[t];
          if (R.lo[sp.set] == R.hi[sp.set])
            R.lo[sp.set] = sp.lv;
          R.hi[sp.set] = sp.lv + 1;
        }
        int slots = 1;
        while (slots < widest)
          slots *= 2;
        const int threads = std::max(64, slots * W);
        if (c1 > c0)
          hipLaunchKernelGGL((k_hybrid<W>), dim3((c1 - c0) * tiles),
                             dim3(threads), 0, st, pat, lev_of(d->lu),
                             lev_of(d->l), lev_of(d->u), lev_of(d->lt), R, c0,
                             tiles, A, LU, b, x, B);
        i = j;
      }
      if (g + 1 < G && !skewed) {
        // the next group starts once this group's first step is done
        HIP_TRY(hipEventRecord(d->ev_skew[g], st));
        skewed = true;
      }
    }
  }
  HIP_TRY(hipGetLastError());
  if (G > 1)
    for (int g = 1; g < G; ++g) {
      HIP_TRY(hipEventRecord(d->ev_join[g - 1], d->side[g - 1]));
      HIP_TRY(hipStreamWaitEvent(s, d->ev_join[g - 1], 0));
    }
  return PVGLU_OK;
}

static int launch_steps(pvglu_plan *P, int W, int mode, bool hybrid, int G,
                        const double *A, double *LU, const double *b, double *x,
                        int B, hipStream_t s) {
  switch (W) {
  case 1: return launch_steps_w<1>(P, mode, hybrid, G, A, LU, b, x, B, s);
  case 2: return launch_steps_w<2>(P, mode, hybrid, G, A, LU, b, x, B, s);
  case 4: return launch_steps_w<4>(P, mode, hybrid, G, A, LU, b, x, B, s);
  case 8: return launch_steps_w<8>(P, mode, hybrid, G, A, LU, b, x, B, s);
  case 16: return launch_steps_w<16>(P, mode, hybrid, G, A, LU, b, x, B, s);
  case 32: return launch_steps_w<32>(P, mode, hybrid, G, A, LU, b, x, B, s);
  default: return launch_steps_w<64>(P, mode, hybrid, G, A, LU, b, x, B, s);
  }
}

template <int W>
static int launch_component_w(pvglu_plan *P, int mode, int threads,
                              const double *A, double *LU, const double *b,
                              double *x, int B, hipStream_t s) {
  const int tiles = (B + W - 1) / W;
  const DevicePlan *d = P->dev;
  hipLaunchKernelGGL((k_component<W>), dim3(P->ncomp * tiles), dim3(threads),
                     0, s, pat_of(d), lev_of(d->lu), lev_of(d->l),
                     lev_of(d->u), lev_of(d->lt), mode, P->fuse_lsolve, tiles,
                     A, LU, b, x, B);
  HIP_TRY(hipGetLastError());
  return PVGLU_OK;
}

static int launch_component(pvglu_plan *P, int W, int mode, int threads,
                            const double *A, double *LU, const double *b,
                            double *x, int B, hipStream_t s) {
  switch (W) {
  case 1: return launch_component_w<1>(P, mode, threads, A, LU, b, x, B, s);
  case 2: return launch_component_w<2>(P, mode, threads, A, LU, b, x, B, s);
  case 4: return launch_component_w<4>(P, mode, threads, A, LU, b, x, B, s);
  case 8: return launch_component_w<8>(P, mode, threads, A, LU, b, x, B, s);
  case 16: return launch_component_w<16>(P, mode, threads, A, LU, b, x, B, s);
  case 32: return launch_component_w<32>(P, mode, threads, A, LU, b, x, B, s);
  default: return launch_component_w<64>(P, mode, threads, A, LU, b, x, B, s);
  }
}

// mode: 0 refactor, 1 refactor+solve, 2 solve, 3 transposed solve
static int run_numeric(pvglu_plan *P, int mode, const double *A, double *LU,
                       const double *b, double *x, int B, void *stream) {
  clear_error();
  int st = require_device(P);
  if (st)
    return st;
  if (B <= 0)
    return fail(PVGLU_E_ARG, "batch must be positive");
  if ((mode != 2 && mode != 3 && (!A || !LU)) ||
      ((mode == 2 || mode == 3) && !LU) || (mode >= 1 && (!b || !x)))
    return fail(PVGLU_E_ARG, "null device pointer");
  hipStream_t s = (hipStream_t)stream;
  int variant, W, threads;
  pick_default_schedule(P, B, variant, W, threads);
  if (variant == PVGLU_SCHED_COMPONENT)
    return launch_component(P, W, mode, threads, A, LU, b, x, B, s);
  return launch_steps(P, W, mode, variant == PVGLU_SCHED_HYBRID,
                      P->sched_groups, A, LU, b, x, B, s);
}

extern "C" int pvglu_refactor(pvglu_plan *P, const double *a, double *lu,
                              int batch, void *stream) {
  return run_numeric(P, 0, a, lu, nullptr, nullptr, batch, stream);
}

extern "C" int pvglu_solve(pvglu_plan *P, const double *lu, const double *b,
                           double *x, int batch, void *stream) {
  return run_numeric(P, 2, nullptr, (double *)lu, b, x, batch, stream);
}

extern "C" int pvglu_refactor_solve(pvglu_plan *P, const double *a, double *lu,
                                    const double *b, double *x, int batch,
                                    void *stream) {
  return run_numeric(P, 1, a, lu, b, x, batch, stream);
}

extern "C" int pvglu_solve_transpose(pvglu_plan *P, const double *lu,
                                     const double *g, double *y, int batch,
                                     void *stream) {
  return run_numeric(P, 3, nullptr, (double *)lu, g, y, batch, stream);
}

extern "C" int pvglu_status(pvglu_plan *P, int *code, int *column,
                            int *instance) {
  clear_error();
  int st = require_device(P);
  if (st)
    return st;
  int h[4] = {0, 0, 0, 0};
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(h, P->dev->status, sizeof h, hipMemcpyDeviceToHost));
  if (code)
    *code = h[0];
  if (column)
    *column = h[1];
  if (instance)
    *instance = h[2];
  return PVGLU_OK;
}

extern "C" int pvglu_status_reset(pvglu_plan *P, void *stream) {
  clear_error();
  int st = require_device(P);
  if (st)
    return st;
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, (hipStream_t)stream,
                     P->dev->status, 4, 0);
  HIP_TRY(hipGetLastError());
  return PVGLU_OK;
}

extern "C" int pvglu_set_schedule(pvglu_plan *P, int variant, int width) {
  clear_error();
  if (!P)
    return fail(PVGLU_E_ARG, "set_schedule: null plan");
  if (variant < PVGLU_SCHED_AUTO || variant > PVGLU_SCHED_HYBRID)
    return fail(PVGLU_E_ARG, "set_schedule: unknown variant %d", variant);
  if (width != 0 && !valid_width(width))
    return fail(PVGLU_E_ARG, "set_schedule: width %d not a power of two <= 64",
                width);
  P->sched_variant = variant;
  P->sched_width = width;
  return PVGLU_OK;
}

extern "C" int pvglu_set_stream_groups(pvglu_plan *P, int groups) {
  clear_error();
  if (!P)
    return fail(PVGLU_E_ARG, "set_stream_groups: null plan");
  if (groups != 1 && groups != 2 && groups != 4)
    return fail(PVGLU_E_ARG, "set_stream_groups: %d not 1, 2 or 4", groups);
  P->sched_groups = groups;
  return PVGLU_OK;
}

extern "C" int pvglu_get_stream_groups(const pvglu_plan *P) {
  if (!P)
    return fail(PVGLU_E_ARG, "get_stream_groups: null plan");
  return P->sched_groups;
}

extern "C" int pvglu_get_schedule(const pvglu_plan *P, int *variant,
                                  int *width) {
  if (!P)
    return fail(PVGLU_E_ARG, "get_schedule: null plan");
  if (variant)
    *variant = P->sched_variant;
  if (width)
    *width = P->sched_width;
  return PVGLU_OK;
}

extern "C" int pvglu_newton_update(pvglu_plan *P, int ne, int batch,
                                   const int *pos, const int *neg,
                                   const double *x, double *v_elem,
                                   double *delta_sum, double *delta_max,
                                   const int *active, void *stream) {
  clear_error();
  int st = require_device(P);
  if (st)
    return st;
  if (ne <= 0 || batch <= 0 || !pos || !neg || !x || !v_elem || !delta_sum ||
      !delta_max)
    return fail(PVGLU_E_ARG, "newton_update: bad arguments");
  const int per = 32;
  const int chunks = (ne + per - 1) / per;
  const int bs = batch >= 256 ? 256 : ((batch + 63) / 64) * 64;
  // zero the per-instance reductions (fill kernel: capture-safe)
  hipLaunchKernelGGL(k_fill_int, dim3((2 * batch + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, (int *)delta_sum, 2 * batch, 0);
  hipLaunchKernelGGL(k_fill_int, dim3((2 * batch + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, (int *)delta_max, 2 * batch, 0);
  hipLaunchKernelGGL(k_newton_update, dim3(chunks, (batch + bs - 1) / bs),
                     dim3(bs), 0,
                     (hipStream_t)stream, ne, batch, per, pos, neg, x, v_elem,
                     delta_sum, (unsigned long long *)delta_max, active,
                     P->dev->status);
  HIP_TRY(hipGetLastError());
  return PVGLU_OK;
}

// ------------------------------------------------------------- autotuner
// Times every (variant, width) candidate of refactor+solve for one batch size
// on library-allocated scratch values (diagonal 1e6, off-diagonal 1: a
// diagonally dominant stand-in; timing does not depend on the values) and
// keeps the fastest.  Results are cached per (pattern hash, batch).

__global__ void k_tune_values(int phase, const int *cp, const int *diag, int n,
                              double *a, double *b, int B) {
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (phase == 0) {
    if (gid < (long)cp[n] * B)
      a[gid] = 1.0;
    if (gid < (long)n * B)
      b[gid] = 1.0;
  } else if (gid < (long)n * B) {
    const int col = (int)(gid / B), inst = (int)(gid % B);
    a[(size_t)diag[col] * B + inst] = 1e6;
  }
}

static std::mutex g_tune_mu;
struct TuneChoice {
  int variant, width, groups;
};
static std::map<std::pair<uint64_t, int>, TuneChoice> g_tune_cache;

extern "C" int pvglu_autotune(pvglu_plan *P, int batch, int reps,
                              double *best_ms) {
  clear_error();
  int st = require_device(P);
  if (st)
    return st;
  if (batch <= 0)
    return fail(PVGLU_E_ARG, "autotune: batch must be positive");
  if (reps <= 0)
    reps = 5;
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune_cache.find({P->hash, batch});
    if (it != g_tune_cache.end() && !best_ms) {
      P->sched_variant = it->second.variant;
      P->sched_width = it->second.width;
      P->sched_groups = it->second.groups;
      P->tuned_batch = batch;
      return PVGLU_OK;
    }
  }
  const size_t nv = (size_t)P->nnz * batch, nx = (size_t)P->n * batch;
  double *a = nullptr, *lu = nullptr, *b = nullptr, *x = nullptr;
  auto release = [&]() {
    (void)hipFree(a);
    (void)hipFree(lu);
    (void)hipFree(b);
    (void)hipFree(x);
  };
  if (hipMalloc(&a, nv * 8) != hipSuccess || hipMalloc(&lu, nv * 8) != hipSuccess ||
      hipMalloc(&b, nx * 8) != hipSuccess || hipMalloc(&x, nx * 8) != hipSuccess) {
    release();
    return fail(PVGLU_E_HIP, "autotune: scratch allocation failed");
  }
  hipStream_t s;
  if (hipStreamCreate(&s) != hipSuccess) {
    release();
    return fail(PVGLU_E_HIP, "autotune: stream creation failed");
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  for (int phase = 0; phase < 2; ++phase)
    hipLaunchKernelGGL(k_tune_values, dim3((unsigned)((nv + 255) / 256)),
                       dim3(256), 0, s, phase, P->dev->cp, P->dev->diag_slot,
                       P->n, a, b, batch);
  const int save_v = P->sched_variant, save_w = P->sched_width,
            save_g = P->sched_groups;
  double best = 1e300;
  int bv = PVGLU_SCHED_LEVEL, bw = 1, bg = 1;
  const int widths[] = {1, 2, 4, 8, 16, 32, 64};
  for (int variant = PVGLU_SCHED_LEVEL; variant <= PVGLU_SCHED_HYBRID && !st;
       ++variant) {
    for (int w : widths) {
      if (w > 1 && w / 2 >= batch)
        continue; // wider tiles only idle lanes
      for (int groups = 1; groups <= 4; groups *= 2) {
        if (groups > 1 && (variant == PVGLU_SCHED_COMPONENT ||
                           !P->dev->side_ok || groups > P->ncomp))
          continue;
        P->sched_variant = variant;
        P->sched_width = w;
        P->sched_groups = groups;
        st = pvglu_refactor_solve(P, a, lu, b, x, batch, s);
        if (st)
          break;
        std::vector<float> t;
        for (int r = 0; r < reps; ++r) {
          (void)hipEventRecord(e0, s);
          st = pvglu_refactor_solve(P, a, lu, b, x, batch, s);
          (void)hipEventRecord(e1, s);
          (void)hipEventSynchronize(e1);
          float ms = 0;
          (void)hipEventElapsedTime(&ms, e0, e1);
          t.push_back(ms);
        }
        std::sort(t.begin(), t.end());
        const double med = t[t.size() / 2];
        if (med < best) {
          best = med;
          bv = variant;
          bw = w;
          bg = groups;
        }
      }
      if (st)
        break;
    }
  }
  (void)hipStreamSynchronize(s);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  release();
  if (st) {
    P->sched_variant = save_v;
    P->sched_width = save_w;
    P->sched_groups = save_g;
    return st;
  }
  // clear any status the stand-in values may have recorded
  int zero[4] = {0, 0, 0, 0};
  HIP_TRY(hipMemcpy(P->dev->status, zero, sizeof zero, hipMemcpyHostToDevice));
  P->sched_variant = bv;
  P->sched_width = bw;
  P->sched_groups = bg;
  P->tuned_batch = batch;
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune_cache[{P->hash, batch}] = {bv, bw, bg};
  }
  if (best_ms)
    *best_ms = best;
  return PVGLU_OK;
}

// ------------------------------------------------------ multi-pattern batch

struct pvglu_multi {
  int ngroups = 0;
  std::vector<pvglu_plan *> plans;
  std::vector<int> batches;
  int width = 16, threads = 256;
};

extern "C" int pvglu_multi_create(int ngroups, pvglu_plan *const *plans,
                                  const int *batches, pvglu_multi **out) {
  clear_error();
  if (!out || ngroups <= 0 || ngroups > kMaxGroups || !plans || !batches)
    return fail(PVGLU_E_ARG, "multi_create: 1..%d groups required", kMaxGroups);
  auto *m = new pvglu_multi();
  m->ngroups = ngroups;
  int minB = 1 << 30, maxc = 1;
  for (int g = 0; g < ngroups; g++) {
    if (!plans[g] || batches[g] <= 0) {
      delete m;
      return fail(PVGLU_E_ARG, "multi_create: bad group %d", g);
    }
    int st = require_device(plans[g]);
    if (st) {
      delete m;
      return st;
    }
    m->plans.push_back(plans[g]);
    m->batches.push_back(batches[g]);
    minB = std::min(minB, batches[g]);
    const pvglu_plan *P = plans[g];
    maxc = std::max({maxc, P->k_lu.max_comp_chains, P->k_u.max_comp_chains,
                     P->k_l.max_comp_chains, P->k_lt.max_comp_chains});
  }
  m->width = minB >= 64 ? 16 : (minB >= 16 ? 8 : (minB >= 4 ? 4 : 1));
  int slots = 1;
  while (slots < maxc && slots * 2 * m->width <= 1024)
    slots *= 2;
  m->threads = std::min(1024, std::max(64, slots * m->width));
  *out = m;
  return PVGLU_OK;
}

extern "C" void pvglu_multi_destroy(pvglu_multi *m) { delete m; }

template <int W>
static int launch_multi_w(pvglu_multi *m, int mode, const double *const *a,
                          double *const *lu, const double *const *b,
                          double *const *x, hipStream_t s) {
  MultiArgs args{};
  args.ngroups = m->ngroups;
  int blocks = 0;
  for (int g = 0; g < m->ngroups; g++) {
    pvglu_plan *P = m->plans[g];
    const DevicePlan *d = P->dev;
    const int tiles = (m->batches[g] + W - 1) / W;
    args.block_start[g] = blocks;
    blocks += P->ncomp * tiles;
    args.tiles[g] = tiles;
    args.B[g] = m->batches[g];
    args.fuse[g] = P->fuse_lsolve ? 1 : 0;
    args.pat[g] = pat_of(d);
    args.lu[g] = lev_of(d->lu);
    args.l[g] = lev_of(d->l);
    args.u[g] = lev_of(d->u);
    args.lt[g] = lev_of(d->lt);
    args.a[g] = a ? a[g] : nullptr;
    args.luv[g] = lu[g];
    args.b[g] = b[g];
    args.x[g] = x[g];
  }
  args.block_start[m->ngroups] = blocks;
  hipLaunchKernelGGL((k_multi<W>), dim3(blocks), dim3(m->threads), 0, s, args,
                     mode);
  HIP_TRY(hipGetLastError());
  return PVGLU_OK;
}

static int launch_multi(pvglu_multi *m, int mode, const double *const *a,
                        double *const *lu, const double *const *b,
                        double *const *x, void *stream) {
  clear_error();
  if (!m || !lu || !b || !x || (mode != 3 && !a))
    return fail(PVGLU_E_ARG, "multi: null pointer arrays");
  for (int g = 0; g < m->ngroups; g++)
    if (!lu[g] || !b[g] || !x[g] || (mode != 3 && !a[g]))
      return fail(PVGLU_E_ARG, "multi: null device pointer in group %d", g);
  hipStream_t s = (hipStream_t)stream;
  switch (m->width) {
  case 1: return launch_multi_w<1>(m, mode, a, lu, b, x, s);
  case 4: return launch_multi_w<4>(m, mode, a, lu, b, x, s);
  case 8: return launch_multi_w<8>(m, mode, a, lu, b, x, s);
  default: return launch_multi_w<16>(m, mode, a, lu, b, x, s);
  }
}

extern "C" int pvglu_multi_refactor_solve(pvglu_multi *m,
                                          const double *const *a,
                                          double *const *lu,
                                          const double *const *b,
                                          double *const *x, void *stream) {
  return launch_multi(m, 1, a, lu, b, x, stream);
}

extern "C" int pvglu_multi_solve_transpose(pvglu_multi *m,
                                           const double *const *lu,
                                           const double *const *g,
                                           double *const *y, void *stream) {
  return launch_multi(m, 3, nullptr, const_cast<double *const *>(lu), g, y,
                      stream);
}
