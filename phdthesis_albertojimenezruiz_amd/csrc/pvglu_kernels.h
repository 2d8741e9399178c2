// pvglu_kernels.h -- gfx950 kernels of the numeric path (device code).
//
// Included only by the per-width translation units pvglu_k<W>.hip, which
// instantiate Launch<W>; pvglu_device.hip holds the host side.
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
#pragma once

#include <hip/hip_runtime.h>

#include "pvglu_internal.h"
#include "pvglu_launch.h"

#include <cmath>
#include <type_traits>

namespace pvglu {

template <int W> __device__ __forceinline__ int uniform(int v) {
  if constexpr (W == 64)
    return __builtin_amdgcn_readfirstlane(v);
  else
    return v;
}

// True when no instance of the calling lane's W-lane tile is active (act
// non-null).  The W lanes of a tile are aligned consecutive lanes of one
// wavefront (64 % W == 0); `valid` is false for lanes past the batch.  In
// the block-per-tile kernels every group of W lanes holds the same tile, so
// the answer is the same in every wavefront of the block.
template <int W>
__device__ __forceinline__ bool tile_idle(const int *act, int inst,
                                          bool valid) {
  const unsigned long long m = __ballot(valid && act[inst] != 0);
  if constexpr (W == 64) {
    return m == 0;
  } else {
    const int first = (int)(threadIdx.x & 63) / W * W;
    return ((m >> first) & ((1ull << W) - 1)) == 0;
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
template <int CH = kChunk, class GET>
__device__ __forceinline__ double dot_sub(double v, int k0, int k1, GET get) {
  int kb = k0;
  for (; k1 - kb >= CH; kb += CH)
    v = chunk_sub<CH>(v, kb, k1, get);
  const int rest = k1 - kb;
  if (rest > CH / 2)
    v = chunk_sub<CH>(v, kb, k1, get);
  else if (rest > 2)
    v = chunk_sub<(CH / 2 > 2 ? CH / 2 : 2)>(v, kb, k1, get);
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

// On return p0..p3 hold the column's last four entries (entries ne-4 .. ne-1)
// and yprev its forward-substitution value: a register run that follows the
// column takes operands from them (records encoded against the window).
template <int W, bool FUSE>
__device__ __forceinline__ void refactor_column(const Pat &P, int c,
                                                const double *A, double *LU,
                                                const double *b, double *x,
                                                size_t B, int inst, double &p0,
                                                double &p1, double &p2,
                                                double &p3, double &yprev) {
  const int q0 = uniform<W>(P.cp[c]);
  const int q1 = uniform<W>(P.cp[c + 1]);
  const int qd = uniform<W>(P.diag[c]);
  double piv = 1.0;
  bool row_with_diag = false; // forward row computed with the diagonal
  double yrow = 0.0;
  if constexpr (FUSE)
    row_with_diag = (uniform<W>(P.drf[c]) & 1) != 0;
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
      p0 = p1;
      p1 = p2;
      p2 = p3;
      p3 = v;
    }
  };
  for (int qb = q0; qb < q1; qb += kChunk) {
    if (q1 - qb > kChunk / 2)
      entries(std::integral_constant<int, kChunk>(), qb);
    else
      entries(std::integral_constant<int, kChunk / 2>(), qb);
  }
  if constexpr (FUSE) {
    if (!row_with_diag)
      yrow = dot_sub_chunked<W>(b[(size_t)c * B + inst], P.lrow_slot,
                                P.lrow_col, LU, x, uniform<W>(P.lrow_ptr[c]),
                                uniform<W>(P.lrow_ptr[c + 1]), B, inst);
    x[(size_t)c * B + inst] = yrow;
    yprev = yrow;
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
  double xn; // stamped right-hand side on homotopy stages: next node voltage
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

// Streaming cache policy (STREAM = true; widths 32 and 64, chosen per launch by
// the host when the batch's values far exceed the 256 MB Infinity Cache).
// A and b are read once per launch, and the strings' L entries and forward
// results are not read again by the wide backward level before they would
// be evicted, so those loads and stores are non-temporal; the backward
// rows' loads are too.  The U entries and pivots keep the default policy:
// the backward level re-reads them, partly from the Infinity Cache.  Measured
// on 512 x 40x24x28 against the default policy: LU level 0 163.7 -> 145.6
// us, the narrow run 39.2 -> 35.1 us, U level 2 82.6 -> 76.5 us
// (non-temporal U stores slowed level 0 by 13 us for 3 us in level 2).  On
// small working sets (one instance, the 128-instance mixed batch) the same
// hints lose the cache reuse and cost time, so there they are off.
template <bool STREAM>
__device__ __forceinline__ double ldv(const double *p) {
  if constexpr (STREAM)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}
template <bool STREAM>
__device__ __forceinline__ void stv(double *p, double v) {
  if constexpr (STREAM)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// A segment whose operands are all forwarded from the previous column runs
// as a three-stage pipeline with a fixed sequence of memory operations per
// column (the loads of column i+2 are issued before column i computes, stores
// are unconditional in NE, the loop is unrolled three times so no register
// copies of in-flight loads are needed): the compiler's counter-based waits
// then wait for the column being computed only, not for the prefetches.
template <int NE, bool FUSE, bool STREAM = false>
__device__ __forceinline__ void pure_load_col(ColVals &c, int slot0, int col,
                                              const double *A,
                                              const double *b, size_t B,
                                              int inst) {
  c.a0 = ldv<STREAM>(&A[(size_t)slot0 * B + inst]);
  if constexpr (NE > 1)
    c.a1 = ldv<STREAM>(&A[(size_t)(slot0 + 1) * B + inst]);
  if constexpr (NE > 2)
    c.a2 = ldv<STREAM>(&A[(size_t)(slot0 + 2) * B + inst]);
  if constexpr (NE > 3)
    c.a3 = ldv<STREAM>(&A[(size_t)(slot0 + 3) * B + inst]);
  if constexpr (FUSE)
    c.bb = ldv<STREAM>(&b[(size_t)col * B + inst]);
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
//
// The steps run in whole rounds of D + 1 (one exit per round, none inside),
// the first round peeled, then at most D steps whose loads were issued by
// the last round.  The compiler derives each step's memory-counter wait
// from the operations issued since its loads; at a loop header it merges
// the entry and back-edge states, so (a) an exit after every step
// (register copies merged at the back edge) and (b) a loop entered from the
// D-load prologue (fewer operations in flight than a round leaves) each made
// it wait for nearly every prefetch once per round (vmcnt(4) and vmcnt(35)
// at the head of the tridiagonal loop at D = 8); with the peeled round the
// loop is entered in its steady state and every step waits only for its own
// loads (vmcnt(62)).  The scheduling barriers keep a step's prefetch ahead of
// its arithmetic: without them the scheduler moved a round's arithmetic
// above the next round's loads, draining the pipeline once per round.
template <int D, class T, class LD, class RUN>
__device__ __forceinline__ void pipeline(int len, LD ld, RUN run) {
  T buf[D + 1];
#pragma unroll
  for (int k = 0; k < D; ++k)
    ld(buf[k], k);
  auto round = [&](int i) {
#pragma unroll
    for (int u = 0; u <= D; ++u) {
      ld(buf[(u + D) % (D + 1)], i + u + D);
      __builtin_amdgcn_sched_barrier(0);
      run(buf[u], i + u);
      __builtin_amdgcn_sched_barrier(0);
    }
  };
  int i = 0;
  if (len > D) {
    round(0);
    for (i = D + 1; i + D < len; i += D + 1)
      round(i);
  }
  // buf[u] holds step i + u (loaded by the last round, or by the prologue)
  const int r = len - i; // 0 .. D
#pragma unroll
  for (int u = 0; u < D; ++u)
    if (u < r)
      run(buf[u], i + u);
}

// Prefetch depth of the refactor and substitution register pipelines: 2 in
// launches whose occupancy matters (wide levels), kNarrow*Depth in the
// narrow-run kernel, where few wavefronts run and each step of a chain would
// otherwise wait for a memory round trip issued only two steps earlier.
constexpr int kLuDepth = 2, kRowDepth = 2;
constexpr int kNarrowLuDepth = 8, kNarrowRowDepth = 12;

template <int NE, bool FUSE, int D = kLuDepth, bool STREAM = false>
__device__ __forceinline__ void pure_lu_segment(
    const int (&f)[16], int len, int dslot, int dcol, const double *A,
    double *LU, const double *b, double *x, size_t B, int inst, double &p0,
    double &p1, double &p2, double &p3, double &yprev, int &badcol) {
  pipeline<D, ColVals>(
      len,
      [&](ColVals &c, int i) {
        i = min(i, len - 1);
        pure_load_col<NE, FUSE, STREAM>(c, f[0] + i * dslot, f[1] + i * dcol,
                                        A, b, B, inst);
      },
      [&](const ColVals &c, int i) {
        pure_lu_col<NE, FUSE>(c, f[0] + i * dslot, f[1] + i * dcol, f, LU, x,
                              B, inst, p0, p1, p2, p3, yprev, badcol);
      });
}

// Tridiagonal step of a series string (segment flags bits 2/3): entries
// (U above, diagonal, L below); the diagonal minus the previous column's
// entry k times the U entry is the pivot, the L entry is divided by it, and
// the fused forward row subtracts the same L operand times the previous y.
// The same operations as pure_lu_col on such a record, without its
// run-time record decoding (selects and branches per update); KC = 2 fixes
// the forwarded entry at compile time (the string columns), and small
// pivots are looked for once per segment (running minimum, then a rescan
// of the stored pivots in the rare case).
template <bool FUSE, int D, int KC = -1, bool STREAM = false>
__device__ __forceinline__ void tri_lu_segment(
    int slot0, int col0, int len, int dslot, int dcol, int k, const double *A,
    double *LU, const double *b, double *x, size_t B, int inst, double &p0,
    double &p1, double &p2, double &p3, double &yprev, int &badcol) {
  if (KC >= 0)
    k = KC;
  double pl = sel4(p0, p1, p2, p3, k);
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, minpiv = 1.0;
  pipeline<D, ColVals>(
      len,
      [&](ColVals &c, int i) {
        i = min(i, len - 1);
        pure_load_col<3, FUSE, STREAM>(c, slot0 + i * dslot, col0 + i * dcol,
                                       A, b, B, inst);
      },
      [&](const ColVals &c, int i) {
        const size_t s = (size_t)(slot0 + i * dslot) * B + inst;
        const double u = c.a0;
        const double piv = c.a1 - pl * u;
        minpiv = fmin(minpiv, fabs(piv));
        const double lo = c.a2 / piv;
        LU[s] = u;
        LU[s + B] = piv;
        stv<STREAM>(&LU[s + 2 * B], lo);
        if constexpr (FUSE) {
          const double y = c.bb - pl * yprev;
          stv<STREAM>(&x[(size_t)(col0 + i * dcol) * B + inst], y);
          yprev = y;
        }
        v0 = u;
        v1 = piv;
        v2 = lo;
        if constexpr (KC == 2)
          pl = lo;
        else
          pl = k == 0 ? u : (k == 1 ? piv : (k == 2 ? lo : 0.0));
      });
  if (minpiv < 1e-8) // rare: find the first small pivot of the segment
    for (int i = 0; i < len; ++i)
      if (fabs(LU[(size_t)(slot0 + i * dslot + 1) * B + inst]) < 1e-8) {
        badcol = col0 + i * dcol;
        break;
      }
  p0 = v0;
  p1 = v1;
  p2 = v2;
  p3 = 0.0;
}

// tri_lu_segment with the fused forward row, its A values stamped on load:
// per column i the U-above entry is the conductance g_u of one module, the
// L-below entry g_l of another and the diagonal their sum, each term signed
// and, on homotopy stages (gv > 0), followed by the same sign times gv, in
// the stamping map's order -- the operations of the stamping gather
// (k_stamp_a) with its zero linear part, so the values are bit-identical to
// the stamped matrix's.  One load per module conductance replaces the three
// loads of A per column (and the stamping kernel's write of them).
//
// RES (record sr[6] bit 0): the correction-form right-hand side of each row
// is formed here too, from the branch currents i of the row's two modules
// -- the residual kernel's operations: i + gv (x_pos - x_neg) on a homotopy
// stage (gv > 0), then 0 - s0 i0 - s1 i1 in the row's branch order; zero for
// a frozen instance -- instead of being read from b, which the residual
// kernel then does not compute for these rows.  The modules' other
// terminals are the segment's previous and next nodes (checked at plan
// time); XH: some lane of the wavefront is on a homotopy stage, so the node
// voltages xk of the previous, this and the next node are carried along.
// A RES segment is a chained string (checked at plan time: column i's
// L-below module is column i+1's U-above one), so each column loads the
// conductance and branch current of one module, its L-below one, and takes
// the U-above module's from the previous column's registers.
template <int D, int KC = -1, bool STREAM = false, bool RES = false,
          bool XH = false>
__device__ __forceinline__ void tri_lu_segment_stamped(
    const int (&sr)[kStampRec], int slot0, int col0, int len, int dslot,
    int dcol, int k, const double *geq, const double *ieq, const double *xk,
    bool live, double gv, double *LU, const double *b, double *x, size_t B,
    int inst, double &p0, double &p1, double &p2, double &p3, double &yprev,
    int &badcol) {
  if (KC >= 0)
    k = KC;
  const int mu0 = sr[1], dmu = sr[2], ml0 = sr[3], dml = sr[4], sb = sr[5],
            rb = sr[6];
  const double su = (sb & 1) ? -1.0 : 1.0, sl = (sb & 2) ? -1.0 : 1.0,
               s1 = (sb & 4) ? -1.0 : 1.0, s2 = (sb & 8) ? -1.0 : 1.0;
  const bool first_u = (sb & 16) != 0;
  const bool rfirst_u = (rb & 2) != 0;
  const double r0s = (rb & 4) ? -1.0 : 1.0, r1s = (rb & 8) ? -1.0 : 1.0;
  double xprv = 0.0, xcur = 0.0; // node voltages of the previous / this row
  if constexpr (RES && XH) {
    xprv = ldv<STREAM>(&xk[(size_t)(col0 - dcol) * B + inst]);
    xcur = ldv<STREAM>(&xk[(size_t)col0 * B + inst]);
  }
  double gup = 0.0, iup = 0.0; // chained: the U-above module's g and i
  if constexpr (RES) {
    gup = ldv<STREAM>(&geq[(size_t)mu0 * B + inst]);
    iup = ldv<STREAM>(&ieq[(size_t)mu0 * B + inst]);
  }
  double pl = sel4(p0, p1, p2, p3, k);
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, minpiv = 1.0;
  auto term = [&](double acc, double sg, double g) {
    acc = acc + sg * g;
    if (gv > 0)
      acc = acc + sg * gv;
    return acc;
  };
  pipeline<D, ColVals>(
      len,
      [&](ColVals &c, int i) {
        i = min(i, len - 1);
        // without RES a string's next column reads one of the two module
        // values again (cached loads); chained (RES), each is read once and
        // not again before the update overwrites it (streaming loads)
        if constexpr (!RES)
          c.a0 = geq[(size_t)(mu0 + i * dmu) * B + inst];
        if constexpr (RES)
          c.a2 = ldv<STREAM>(&geq[(size_t)(ml0 + i * dml) * B + inst]);
        else
          c.a2 = geq[(size_t)(ml0 + i * dml) * B + inst];
        if constexpr (RES) { // the L-below module's branch current
          c.a3 = ldv<STREAM>(&ieq[(size_t)(ml0 + i * dml) * B + inst]);
          if constexpr (XH) // the next node (of the last column: sr[7])
            c.xn = ldv<STREAM>(
                &xk[(size_t)(i + 1 < len ? col0 + (i + 1) * dcol : sr[7]) * B +
                    inst]);
        } else {
          c.bb = ldv<STREAM>(&b[(size_t)(col0 + i * dcol) * B + inst]);
        }
      },
      [&](const ColVals &c, int i) {
        const size_t s = (size_t)(slot0 + i * dslot) * B + inst;
        const double zero = 0.0;
        const double gu = RES ? gup : c.a0;
        const double au = term(zero, su, gu);
        const double al = term(zero, sl, c.a2);
        const double ad = first_u ? term(term(zero, s1, gu), s2, c.a2)
                                  : term(term(zero, s1, c.a2), s2, gu);
        const double u = au;
        const double piv = ad - pl * u;
        minpiv = fmin(minpiv, fabs(piv));
        const double lo = al / piv;
        LU[s] = u;
        LU[s + B] = piv;
        stv<STREAM>(&LU[s + 2 * B], lo);
        double rhs;
        if constexpr (RES) {
          double ia = rfirst_u ? iup : c.a3, ib = rfirst_u ? c.a3 : iup;
          if constexpr (XH) {
            // x_pos - x_neg of each module: this node and the other terminal
            // (the previous node for the U-above module, the next for the
            // L-below one); s > 0: this node is the module's positive one
            const double xa = rfirst_u ? xprv : c.xn, xb = rfirst_u ? c.xn : xprv;
            if (gv > 0) {
              ia = ia + gv * (r0s > 0 ? xcur - xa : xa - xcur);
              ib = ib + gv * (r1s > 0 ? xcur - xb : xb - xcur);
            }
            xprv = xcur;
            xcur = c.xn;
          }
          double acc = 0.0;
          acc = acc - r0s * ia;
          acc = acc - r1s * ib;
          rhs = live ? acc : 0.0;
          gup = c.a2; // the next column's U-above module
          iup = c.a3;
        } else {
          rhs = c.bb;
        }
        const double y = rhs - pl * yprev;
        stv<STREAM>(&x[(size_t)(col0 + i * dcol) * B + inst], y);
        yprev = y;
        v0 = u;
        v1 = piv;
        v2 = lo;
        if constexpr (KC == 2)
          pl = lo;
        else
          pl = k == 0 ? u : (k == 1 ? piv : (k == 2 ? lo : 0.0));
      });
  if (minpiv < 1e-8)
    for (int i = 0; i < len; ++i)
      if (fabs(LU[(size_t)(slot0 + i * dslot + 1) * B + inst]) < 1e-8) {
        badcol = col0 + i * dcol;
        break;
      }
  p0 = v0;
  p1 = v1;
  p2 = v2;
  p3 = 0.0;
}

// Star column (column flag bit 1; the DC bus): only the diagonal has
// updates, so the entries above the diagonal are their A values, every U
// operand of the diagonal is such an A value, and the entries below the
// diagonal are A / pivot; the operands' slots and rows are arithmetic
// progressions, addressed from their first two values (no per-operand index
// loads).  Used by the narrow-run kernel, where the column is on a
// latency-bound chain.  Operation order per entry = refactor_column's.
constexpr int kStarMax = kStarMaxUpdates;
constexpr int kStarBelow = 4; // entries below the diagonal held in registers

// The star column's 16-int record (built at plan time, at the column's
// position in kernel order): q0, q1, diagonal slot, column flags, forward row
// range r0 r1, number of updates, then (first, step) of the L slots, the U
// slots and the forward rows.  One scalar load burst replaces the chain of
// dependent metadata loads (column pointers -> update ranges -> operands).
//
// Every load of the column is issued before anything is computed or stored:
// the operands of the diagonal's sum (and of the fused forward row), the
// diagonal, the right-hand side, up to kStarBelow entries below the diagonal
// and the last four entries above it (the register window of the run that
// follows).  When the U operands are the entries above the diagonal in slot
// order (the plants' bus columns: u0 = q0, step 1, one update per entry),
// those entries are stored from the operand registers.  The column then
// costs one or two memory round trips (over 63 loads in flight the issue
// waits) instead of one per 8-entry chunk of the copy and one for the
// entries below (measured on the 40x24x28 plant's bus column: 9.3 us).  The
// sums are unrolled to NS >= the update count (8, 16, 24 or 32), so few
// clamped repeat loads fill the counter.
template <int NS, int W, bool FUSE>
__device__ __forceinline__ void star_body(const Pat &P, const int (&f)[16],
                                          int c, const double *A, double *LU,
                                          const double *b, double *x, size_t B,
                                          int inst, double &p0, double &p1,
                                          double &p2, double &p3,
                                          double &yprev) {
  const int q0 = f[0], q1 = f[1], qd = f[2];
  const bool row_with_diag = FUSE && (f[3] & 1) != 0;
  const int nd = f[6]; // 1 .. NS
  const int l0 = f[7], dl = f[8], u0 = f[9], du = f[10], y0 = f[11],
            dy = f[12];
  const int na = qd - q0, nb = q1 - qd - 1; // entries above / below
  const bool above_is_u = u0 == q0 && du == 1 && nd == na;
  double l[NS], u[NS], yv[NS], bl[kStarBelow], wa[4];
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    const int k = min(j, nd - 1);
    u[j] = A[(size_t)(u0 + k * du) * B + inst];
    l[j] = LU[(size_t)(l0 + k * dl) * B + inst];
    if (row_with_diag)
      yv[j] = x[(size_t)(y0 + k * dy) * B + inst];
  }
  double v = A[(size_t)qd * B + inst];
  double y = row_with_diag ? b[(size_t)c * B + inst] : 0.0;
#pragma unroll
  for (int k = 0; k < kStarBelow; ++k)
    bl[k] = A[(size_t)min(qd + 1 + k, q1 - 1) * B + inst];
#pragma unroll
  for (int k = 0; k < 4; ++k) // entry na - 4 + k above the diagonal, or 0
    wa[k] = na - 4 + k >= 0 ? A[(size_t)(q0 + max(na - 4 + k, 0)) * B + inst]
                            : 0.0;
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    const double pv = v - l[j] * u[j];
    v = j < nd ? pv : v;
    if (row_with_diag) {
      const double py = y - l[j] * yv[j];
      y = j < nd ? py : y;
    }
  }
  if (fabs(v) < 1e-8)
    record_status(P.status, PVGLU_E_SINGULAR, c, inst);
  // entries above the diagonal: no updates (LU = A)
  if (above_is_u) {
#pragma unroll
    for (int j = 0; j < NS; ++j)
      if (j < nd)
        LU[(size_t)(q0 + j) * B + inst] = u[j];
  } else {
    for (int qb = q0; qb < qd; qb += 8) {
      double tv[8];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        tv[k] = A[(size_t)min(qb + k, qd - 1) * B + inst];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (qb + k < qd)
          LU[(size_t)(qb + k) * B + inst] = tv[k];
    }
  }
  LU[(size_t)qd * B + inst] = v;
  // entries below the diagonal: no updates, divided by the pivot
#pragma unroll
  for (int k = 0; k < kStarBelow; ++k) {
    bl[k] = bl[k] / v;
    if (k < nb)
      LU[(size_t)(qd + 1 + k) * B + inst] = bl[k];
  }
  // the window: the column's last four entries in slot order
  double w[4];
  if (nb <= kStarBelow) {
    // entries na-4 .. na-1 above, the diagonal, then the nb below
    const int m = nb;
    w[0] = m == 0 ? wa[1] : m == 1 ? wa[2] : m == 2 ? wa[3] : m == 3 ? v : bl[0];
    w[1] = m == 0 ? wa[2] : m == 1 ? wa[3] : m == 2 ? v : m == 3 ? bl[0] : bl[1];
    w[2] = m == 0 ? wa[3] : m == 1 ? v : m == 2 ? bl[0] : m == 3 ? bl[1] : bl[2];
    w[3] = m == 0 ? v : m == 1 ? bl[0] : m == 2 ? bl[1] : m == 3 ? bl[2] : bl[3];
  } else {
    w[0] = bl[0];
    w[1] = bl[1];
    w[2] = bl[2];
    w[3] = bl[3];
    for (int qb = qd + 1 + kStarBelow; qb < q1; qb += 4) {
      double tv[4];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        tv[k] = A[(size_t)min(qb + k, q1 - 1) * B + inst];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (qb + k < q1) {
          const double lv = tv[k] / v;
          LU[(size_t)(qb + k) * B + inst] = lv;
          w[0] = w[1];
          w[1] = w[2];
          w[2] = w[3];
          w[3] = lv;
        }
    }
  }
  p0 = w[0];
  p1 = w[1];
  p2 = w[2];
  p3 = w[3];
  if constexpr (FUSE) {
    if (!row_with_diag)
      y = dot_sub_chunked<W>(b[(size_t)c * B + inst], P.lrow_slot, P.lrow_col,
                             LU, x, f[4], f[5], B, inst);
    x[(size_t)c * B + inst] = y;
    yprev = y;
  }
}

template <int W, bool FUSE>
__device__ __forceinline__ void star_column(const Pat &P, int t, int c,
                                            const double *A, double *LU,
                                            const double *b, double *x,
                                            size_t B, int inst, double &p0,
                                            double &p1, double &p2,
                                            double &p3, double &yprev) {
  int f[16];
  load_record<W>(P.lu_rec + (size_t)t * 16, f);
  const int nd = f[6];
  if (nd <= 8)
    star_body<8, W, FUSE>(P, f, c, A, LU, b, x, B, inst, p0, p1, p2, p3, yprev);
  else if (nd <= 16)
    star_body<16, W, FUSE>(P, f, c, A, LU, b, x, B, inst, p0, p1, p2, p3, yprev);
  else if (nd <= 24)
    star_body<24, W, FUSE>(P, f, c, A, LU, b, x, B, inst, p0, p1, p2, p3, yprev);
  else
    star_body<kStarMax, W, FUSE>(P, f, c, A, LU, b, x, B, inst, p0, p1, p2, p3,
                                 yprev);
}

// Backward-substitution values of one row: y, U(i,i) and the two U terms.
struct RowVals {
  double y, dv, u0, u1;
};

// Row sweeps share one register program: x_i = (src_i - sum_j T(i,j) x_j)
// [/ T(i,i) when DIV], T = U (backward, src = x in place, DIV) or unit L
// (forward, src = b).  Records hold slots relative to the row's diagonal slot.
template <int NT, bool DIV, int D = kRowDepth, bool STREAM = false>
__device__ __forceinline__ void pure_row_segment(const int (&f)[8], int len,
                                                 int drow, int ddiag,
                                                 const double *LU,
                                                 const double *src, double *x,
                                                 size_t B, int inst,
                                                 double &xprev) {
  pipeline<D, RowVals>(
      len,
      [&](RowVals &r, int i) {
        i = min(i, len - 1);
        const int row = f[0] + i * drow, dslot = f[1] + i * ddiag;
        r.y = ldv<STREAM>(&src[(size_t)row * B + inst]);
        if constexpr (DIV)
          r.dv = ldv<STREAM>(&LU[(size_t)dslot * B + inst]);
        if constexpr (NT > 0)
          r.u0 = ldv<STREAM>(&LU[(size_t)(dslot + f[4]) * B + inst]);
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

// p0..p3 / yprev: the previous column's window (zero at a chain's start),
// updated to this run's last column.
// SF: the launch stamps the A values of the string segments that have a
// stamping record (OP_REFACTOR_STAMPED, FUSE).
template <int W, bool FUSE, int D = kLuDepth, bool STREAM = false,
          bool SF = false>
__device__ __forceinline__ void fast_lu_chain(const Pat &P, int t0, int t1,
                                              const double *A, double *LU,
                                              const double *b, double *x,
                                              size_t B, int inst, double &p0,
                                              double &p1, double &p2,
                                              double &p3, double &yprev) {
  double gv = 0.0;
  if constexpr (SF)
    gv = P.gvs[inst];
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
    const int sflag = uniform<W>(sg[3]);
    if (sflag & (FUSE ? 8 : 4)) {
      int badcol = -1;
      if constexpr (SF) {
        int sr[kStampRec];
        load_record<W>(P.stamp_seg + (size_t)t * kStampRec, sr);
        if (sr[0]) {
          const bool live = P.live ? P.live[inst] != 0 : true;
          const int kc = (sflag >> 4) & 3;
#define PVGLU_STAMPED(KC_, RES_, XH_)                                          \
  tri_lu_segment_stamped<D, KC_, STREAM, RES_, XH_>(                           \
      sr, f[0], f[1], len, dslot, dcol, kc, P.geq, P.ieq, P.xk, live, gv, LU,  \
      b, x, B, inst, p0, p1, p2, p3, yprev, badcol)
          if ((sr[6] & 1) && P.ieq && kc == 2) {
            // node voltages are read only when a lane of the wavefront is on
            // a homotopy stage (uniform choice of the variant)
            if (__ballot(live && gv > 0) != 0)
              PVGLU_STAMPED(2, true, true);
            else
              PVGLU_STAMPED(2, true, false);
          } else if (kc == 2) {
            PVGLU_STAMPED(2, false, true);
          } else {
            PVGLU_STAMPED(-1, false, true);
          }
#undef PVGLU_STAMPED
          if (badcol >= 0)
            record_status(P.status, PVGLU_E_SINGULAR, badcol, inst);
          t += len;
          continue;
        }
      }
      if (((sflag >> 4) & 3) == 2)
        tri_lu_segment<FUSE, D, 2, STREAM>(f[0], f[1], len, dslot, dcol, 2, A,
                                           LU, b, x, B, inst, p0, p1, p2, p3,
                                           yprev, badcol);
      else
        tri_lu_segment<FUSE, D, -1, STREAM>(f[0], f[1], len, dslot, dcol,
                                            (sflag >> 4) & 3, A, LU, b, x, B,
                                            inst, p0, p1, p2, p3, yprev,
                                            badcol);
      if (badcol >= 0)
        record_status(P.status, PVGLU_E_SINGULAR, badcol, inst);
      t += len;
      continue;
    }
    if (sflag & (FUSE ? 1 : 2)) {
      int badcol = -1;
      switch (ne) {
      case 1:
        pure_lu_segment<1, FUSE, D, STREAM>(f, len, dslot, dcol, A, LU, b, x, B, inst, p0, p1,
                           p2, p3, yprev, badcol);
        break;
      case 2:
        pure_lu_segment<2, FUSE, D, STREAM>(f, len, dslot, dcol, A, LU, b, x, B, inst, p0, p1,
                           p2, p3, yprev, badcol);
        break;
      case 3:
        pure_lu_segment<3, FUSE, D, STREAM>(f, len, dslot, dcol, A, LU, b, x, B, inst, p0, p1,
                           p2, p3, yprev, badcol);
        break;
      default:
        pure_lu_segment<4, FUSE, D, STREAM>(f, len, dslot, dcol, A, LU, b, x, B, inst, p0, p1,
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

template <int W, bool DIV, int D = kRowDepth, bool STREAM = false>
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
        pure_row_segment<0, DIV, D, STREAM>(f, len, drow, ddiag, LU, src, x, B, inst,
                                 xprev);
      else
        pure_row_segment<1, DIV, D, STREAM>(f, len, drow, ddiag, LU, src, x, B, inst,
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


// NAR = 1: the narrow-run kernel's variant (deep register pipelines, star
// columns loaded in one round trip).
template <int W, int OP, int NAR = 0, bool STREAM = false>
__device__ __forceinline__ void run_chain(const Pat &P, const Lev &L, int chain,
                                          const double *A, double *LU,
                                          const double *b, double *x, size_t B,
                                          int inst) {
  constexpr int DL = NAR ? kNarrowLuDepth : kLuDepth;
  constexpr int DR = NAR ? kNarrowRowDepth : kRowDepth;
  const int t0 = uniform<W>(L.chain_col_ptr[chain]);
  const int t1 = uniform<W>(L.chain_col_ptr[chain + 1]);
  if constexpr (OP == OP_REFACTOR_FUSED || OP == OP_REFACTOR ||
                OP == OP_REFACTOR_STAMPED) {
    constexpr bool FUSE = OP != OP_REFACTOR;
    constexpr bool SF = OP == OP_REFACTOR_STAMPED;
    double p0 = 0, p1 = 0, p2 = 0, p3 = 0, yprev = 0;
    if (uniform<W>(P.lu_fast[chain])) {
      fast_lu_chain<W, FUSE, DL, STREAM, SF>(P, t0, t1, A, LU, b, x, B, inst,
                                             p0, p1, p2, p3, yprev);
      return;
    }
    // mixed chain: register runs between generic columns; a run's first
    // column may take operands from the generic column before it (window)
    for (int t = t0; t < t1;) {
      const int e = uniform<W>(P.lu_run[t]);
      if (e > t) {
        fast_lu_chain<W, FUSE, DL, STREAM, SF>(P, t, e, A, LU, b, x, B, inst,
                                               p0, p1, p2, p3, yprev);
        t = e;
      } else {
        const int c = uniform<W>(L.cols[t]);
        bool star = false;
        if constexpr (NAR != 0)
          star = (uniform<W>(P.drf[c]) & 2) != 0;
        if constexpr (NAR != 0) {
          if (star)
            star_column<W, FUSE>(P, t, c, A, LU, b, x, B, inst, p0, p1, p2,
                                 p3, yprev);
        }
        if (!star)
          refactor_column<W, FUSE>(P, c, A, LU, b, x, B, inst, p0, p1, p2, p3,
                                   yprev);
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
      fast_row_chain<W, DIV, DR, STREAM>(rec, seg, t0, t1, LU, src, x, B,
                                         inst);
      return;
    }
    for (int t = t0; t < t1;) {
      const int e = uniform<W>(run[t]);
      if (e > t) {
        fast_row_chain<W, DIV, DR, STREAM>(rec, seg, t, e, LU, src, x, B,
                                           inst);
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
    if constexpr (OP == OP_LSOLVE)
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
// group -> (instance tile, chain of the level).  The backward sweeps (U and
// L^T) take the groups in descending order: a wide backward level re-reads
// the factors and forward results that the wide forward level wrote shortly
// before, and in reverse order its first blocks read what was written last,
// which the 256 MB Infinity Cache still holds (measured: 10 us less per
// 512-instance refactor+solve of the 40x24x28 plant).
template <int W, int OP, bool STREAM = false>
__global__ void __launch_bounds__(256)
    k_level(Pat P, Lev L, int first_chain, int nchains, int tiles,
            const double *A, double *LU, const double *b, double *x, int B) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int g0 = uniform<W>(gid / W);
  const int lane = gid % W;
  if (g0 >= nchains * tiles)
    return;
  constexpr bool backward = OP == OP_USOLVE || OP == OP_LT;
  const int group = backward ? nchains * tiles - 1 - g0 : g0;
  // consecutive wavefronts take consecutive instance tiles of one chain, so
  // the tiles of a slot row are streamed together
  const int tile = group % tiles;
  const int chain = first_chain + group / tiles;
  const int inst = tile * W + lane;
  if (inst >= B)
    return;
  if (P.act && tile_idle<W>(P.act, inst, true))
    return;
  run_chain<W, OP, 0, STREAM>(P, L, chain, A, LU, b, x, (size_t)B, inst);
}

// COMPONENT schedule: one launch for the whole refactor + solve.  Block ->
// (component, instance tile); each level of the component is split over the
// block's chain slots; barriers separate levels.  Every thread reaches every
// barrier (loop bounds are block-uniform; inactive lanes only skip work).
template <int W, int NAR, bool STREAM = false>
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
          run_chain<W, OP_REFACTOR_FUSED, NAR, STREAM>(P, L, chain, A, LU, b, x, B,
                                              inst);
        else if (op == OP_REFACTOR)
          run_chain<W, OP_REFACTOR, NAR, STREAM>(P, L, chain, A, LU, b, x, B,
                                              inst);
        else if (op == OP_LSOLVE)
          run_chain<W, OP_LSOLVE, NAR, STREAM>(P, L, chain, A, LU, b, x, B,
                                              inst);
        else if (op == OP_USOLVE)
          run_chain<W, OP_USOLVE, NAR, STREAM>(P, L, chain, A, LU, b, x, B,
                                              inst);
        else if (op == OP_UT)
          run_chain<W, OP_UT, NAR, STREAM>(P, L, chain, A, LU, b, x, B,
                                              inst);
        else
          run_chain<W, OP_LT, NAR, STREAM>(P, L, chain, A, LU, b, x, B,
                                              inst);
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
  if (P.act && tile_idle<W>(P.act, inst, active))
    return; // the whole block: one tile, before any barrier
  if (mode == 3) {
    component_levels<W, 0>(P, Llu, comp, slot, slots, active, A, LU, b, x,
                               Bs, inst, OP_UT);
    component_levels<W, 0>(P, Llt, comp, slot, slots, active, A, LU, b, x,
                               Bs, inst, OP_LT);
    return;
  }
  if (mode <= 1)
    component_levels<W, 0>(P, Llu, comp, slot, slots, active, A, LU, b, x,
                               Bs, inst,
                               (mode == 1 && fuse) ? OP_REFACTOR_FUSED
                                                   : OP_REFACTOR);
  if (mode == 0)
    return;
  if (mode == 2 || !fuse)
    component_levels<W, 0>(P, Ll, comp, slot, slots, active, A, LU, b, x,
                               Bs, inst, OP_LSOLVE);
  component_levels<W, 0>(P, Lu, comp, slot, slots, active, A, LU, b, x, Bs,
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
// each with its operation, barriers between levels.  Few wavefronts run, so
// the chains use the deep register pipelines and star columns take the
// one-round-trip path.  At most 4 chain slots of W <= 64 lanes (a narrow
// level has at most 4 chains per component), so at most 256 threads.
template <int W, bool STREAM = false>
__global__ void __launch_bounds__(256)
    k_hybrid(Pat P, Lev Llu, Lev Ll, Lev Lu, Lev Llt, LvRange R, int comp0,
             int tiles, const double *A, double *LU, const double *b,
             double *x, int B) {
  const int comp = comp0 + blockIdx.x / tiles, tile = blockIdx.x % tiles;
  const int slots = blockDim.x / W;
  const int slot = uniform<W>((int)threadIdx.x / W);
  const int inst = tile * W + (int)threadIdx.x % W;
  const bool active = inst < B;
  if (P.act && tile_idle<W>(P.act, inst, active))
    return; // the whole block: one tile, before any barrier
  // level set k selected field by field (no indexed array of kernel
  // arguments, which would live in scratch memory)
  for (int k = 0; k < 4; ++k) {
    const int lo = k == 0 ? R.lo[0] : k == 1 ? R.lo[1] : k == 2 ? R.lo[2] : R.lo[3];
    const int hi = k == 0 ? R.hi[0] : k == 1 ? R.hi[1] : k == 2 ? R.hi[2] : R.hi[3];
    const int op = k == 0 ? R.op[0] : k == 1 ? R.op[1] : k == 2 ? R.op[2] : R.op[3];
    if (lo >= hi)
      continue;
    const Lev Lk = k == 0 ? Llu : k == 1 ? Ll : k == 2 ? Lu : Llt;
    component_levels<W, 1, STREAM>(P, Lk, comp, slot, slots, active, A, LU,
                                   b, x, (size_t)B, inst, op, lo, hi);
  }
}

// Multi-pattern LEVEL step: group g's blocks take its chains of the level
// as k_level does (backward sweeps in descending order).
template <int W, int OP>
__global__ void __launch_bounds__(256) k_level_multi(MultiLevel m) {
  int g = 0;
  while (g + 1 < m.ngroups && (int)blockIdx.x >= m.block_start[g + 1])
    ++g;
  const int gid = ((int)blockIdx.x - m.block_start[g]) * blockDim.x + threadIdx.x;
  const int g0 = uniform<W>(gid / W);
  const int lane = gid % W;
  const int tiles = m.tiles[g], total = m.nch[g] * tiles;
  if (g0 >= total)
    return;
  constexpr bool backward = OP == OP_USOLVE || OP == OP_LT;
  const int group = backward ? total - 1 - g0 : g0;
  const int tile = group % tiles;
  const int chain = m.first[g] + group / tiles;
  const int inst = tile * W + lane;
  if (inst >= m.B[g])
    return;
  run_chain<W, OP>(m.pat[g], m.lev[g], chain, m.a[g], m.luv[g], m.b[g],
                   m.x[g], (size_t)m.B[g], inst);
}

// Multi-pattern narrow run: block -> (group, component, tile), the body of
// k_hybrid with the group's plan and level ranges.
template <int W>
__global__ void __launch_bounds__(256) k_run_multi(MultiRun m) {
  int g = 0;
  while (g + 1 < m.ngroups && (int)blockIdx.x >= m.block_start[g + 1])
    ++g;
  const int local = (int)blockIdx.x - m.block_start[g];
  const int tiles = m.tiles[g];
  const int comp = local / tiles, tile = local % tiles;
  const int slots = blockDim.x / W;
  const int slot = uniform<W>((int)threadIdx.x / W);
  const int inst = tile * W + (int)threadIdx.x % W;
  const int B = m.B[g];
  const bool active = inst < B;
  const LvRange &R = m.R[g];
  for (int k = 0; k < 4; ++k) {
    const int lo = k == 0 ? R.lo[0] : k == 1 ? R.lo[1] : k == 2 ? R.lo[2] : R.lo[3];
    const int hi = k == 0 ? R.hi[0] : k == 1 ? R.hi[1] : k == 2 ? R.hi[2] : R.hi[3];
    const int op = k == 0 ? R.op[0] : k == 1 ? R.op[1] : k == 2 ? R.op[2] : R.op[3];
    if (lo >= hi)
      continue;
    const Lev Lk = k == 0 ? m.lu[g] : k == 1 ? m.l[g] : k == 2 ? m.u[g] : m.lt[g];
    component_levels<W, 1>(m.pat[g], Lk, comp, slot, slots, active, m.a[g],
                           m.luv[g], m.b[g], m.x[g], (size_t)B, inst, op, lo,
                           hi);
  }
}

// Multi-pattern work list: groups of instances with their own plans in one
// launch; block -> (group, component, instance tile).

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

template <int W>
void Launch<W>::level(const Pat &pat, const Lev &lev, int op, int first,
                      int nch, int tiles, const double *A, double *LU,
                      const double *b, double *x, int B, hipStream_t s,
                      bool stream) {
  const long threads = (long)nch * tiles * W;
  const int blocks = (int)((threads + 255) / 256);
  if (blocks == 0)
    return;
  constexpr bool kStreamable = W >= 32; // streaming policy: widths 32, 64
#define PVGLU_LVL(OPC)                                                         \
  do {                                                                         \
    if (kStreamable && stream)                                                 \
      hipLaunchKernelGGL((k_level<W, OPC, kStreamable>), dim3(blocks),         \
                         dim3(256), 0, s, pat, lev, first, nch, tiles, A, LU,  \
                         b, x, B);                                             \
    else                                                                       \
      hipLaunchKernelGGL((k_level<W, OPC>), dim3(blocks), dim3(256), 0, s,     \
                         pat, lev, first, nch, tiles, A, LU, b, x, B);         \
  } while (0)
  switch (op) {
  case OP_REFACTOR: PVGLU_LVL(OP_REFACTOR); break;
  case OP_REFACTOR_FUSED: PVGLU_LVL(OP_REFACTOR_FUSED); break;
  case OP_REFACTOR_STAMPED: PVGLU_LVL(OP_REFACTOR_STAMPED); break;
  case OP_LSOLVE: PVGLU_LVL(OP_LSOLVE); break;
  case OP_USOLVE: PVGLU_LVL(OP_USOLVE); break;
  case OP_UT: PVGLU_LVL(OP_UT); break;
  default: PVGLU_LVL(OP_LT); break;
  }
#undef PVGLU_LVL
}

template <int W>
void Launch<W>::hybrid(const Pat &pat, const Lev &lu, const Lev &l,
                       const Lev &u, const Lev &lt, const LvRange &R,
                       int comp0, int ncomp, int tiles, int threads,
                       const double *A, double *LU, const double *b, double *x,
                       int B, hipStream_t s, bool stream) {
  if (ncomp <= 0)
    return;
  constexpr bool kStreamable = W >= 32;
  if (kStreamable && stream)
    hipLaunchKernelGGL((k_hybrid<W, kStreamable>), dim3(ncomp * tiles),
                       dim3(threads), 0, s, pat, lu, l, u, lt, R, comp0,
                       tiles, A, LU, b, x, B);
  else
    hipLaunchKernelGGL((k_hybrid<W>), dim3(ncomp * tiles), dim3(threads), 0, s,
                       pat, lu, l, u, lt, R, comp0, tiles, A, LU, b, x, B);
}

template <int W>
void Launch<W>::component(const Pat &pat, const Lev &lu, const Lev &l,
                          const Lev &u, const Lev &lt, int mode, bool fuse,
                          int ncomp, int tiles, int threads, const double *A,
                          double *LU, const double *b, double *x, int B,
                          hipStream_t s) {
  hipLaunchKernelGGL((k_component<W>), dim3(ncomp * tiles), dim3(threads), 0,
                     s, pat, lu, l, u, lt, mode, fuse, tiles, A, LU, b, x, B);
}

template <int W>
void Launch<W>::multi(const MultiArgs &m, int mode, int blocks, int threads,
                      hipStream_t s) {
  hipLaunchKernelGGL((k_multi<W>), dim3(blocks), dim3(threads), 0, s, m, mode);
}

template <int W>
void Launch<W>::level_multi(const MultiLevel &m, int op, int blocks,
                            hipStream_t s) {
  if (blocks == 0)
    return;
#define PVGLU_LVLM(OPC)                                                        \
  hipLaunchKernelGGL((k_level_multi<W, OPC>), dim3(blocks), dim3(256), 0, s, m)
  switch (op) {
  case OP_REFACTOR: PVGLU_LVLM(OP_REFACTOR); break;
  case OP_REFACTOR_FUSED: PVGLU_LVLM(OP_REFACTOR_FUSED); break;
  case OP_LSOLVE: PVGLU_LVLM(OP_LSOLVE); break;
  case OP_USOLVE: PVGLU_LVLM(OP_USOLVE); break;
  case OP_UT: PVGLU_LVLM(OP_UT); break;
  default: PVGLU_LVLM(OP_LT); break;
  }
#undef PVGLU_LVLM
}

template <int W>
void Launch<W>::run_multi(const MultiRun &m, int blocks, int threads,
                          hipStream_t s) {
  if (blocks > 0)
    hipLaunchKernelGGL((k_run_multi<W>), dim3(blocks), dim3(threads), 0, s, m);
}

} // namespace pvglu
