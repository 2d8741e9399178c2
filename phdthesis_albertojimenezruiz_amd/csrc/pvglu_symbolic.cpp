// pvglu_symbolic.cpp -- host half of the symbolic plan.
//
// Restates, for a filled pattern, the one-time analysis of prepare_matrices
// (PV/000_PVPP_simulator.cpp:33-90): dependencies (PV/006), levels and chains
// (PV/007), their flattening (PV/008), per-level row widths (PV/009) and the
// CSR/CSC layout with value maps (PV/010).  Then compiles the gather schedule
// the HIP kernels execute: for every CSC slot the ordered list of (L, U)
// operand slots of its updates, the substitution row lists, and a component
// partition for the single-launch schedule.
#include "pvglu_internal.h"

#include <algorithm>
#include <array>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>

namespace pvglu {

static thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

void clear_error() { g_err.clear(); }

const char *last_error() { return g_err.c_str(); }

// ------------------------------------------------------------------ 006
void dependencies(DepKind kind, int n, const std::vector<int> &rp,
                  const std::vector<int> &ci, std::vector<int> &dptr,
                  std::vector<int> &didx) {
  // GLU (006:5-47): deps(k) = {i<k : (i,k)} -- the reference scans column k
  // in both of its loops, and its "column i of L is not empty" test always
  // passes on the diagonal.  L (006:49-62): {j<i : (i,j)}.  U (006:64-77):
  // {j>i : (i,j)}.  LT (adjoint, L^T solve): deps(j) = {i>j : (i,j)}.
  auto wants = [kind](int i, int j, int &owner, int &dep) {
    switch (kind) {
    case DEP_GLU: owner = j; dep = i; return i < j;
    case DEP_L: owner = i; dep = j; return j < i;
    case DEP_U: owner = i; dep = j; return j > i;
    default: owner = j; dep = i; return i > j;
    }
  };
  dptr.assign(n + 1, 0);
  for (int i = 0; i < n; i++)
    for (int p = rp[i]; p < rp[i + 1]; p++) {
      int o, d;
      if (wants(i, ci[p], o, d))
        dptr[o + 1]++;
    }
  for (int i = 0; i < n; i++)
    dptr[i + 1] += dptr[i];
  didx.assign(dptr[n], 0);
  std::vector<int> fill(dptr.begin(), dptr.end() - 1);
  // Rows are visited in ascending order, so every list is ascending for the
  // column-owned kinds; the row-owned kinds inherit the sorted row order.
  for (int i = 0; i < n; i++)
    for (int p = rp[i]; p < rp[i + 1]; p++) {
      int o, d;
      if (wants(i, ci[p], o, d))
        didx[fill[o]++] = d;
    }
}

// --------------------------------------------------------------- 007/008
bool levels_with_chains(int n, const std::vector<int> &dptr,
                        const std::vector<int> &didx, LevelSet &out) {
  // Basic levels (007:4-60) are the longest-path levels of the dependency
  // DAG with ascending column order inside a level; computed here with one
  // topological sweep instead of repeated scans.
  std::vector<int> indeg(n), level(n, 0), rptr(n + 1, 0);
  for (int i = 0; i < n; i++) {
    indeg[i] = dptr[i + 1] - dptr[i];
    for (int p = dptr[i]; p < dptr[i + 1]; p++)
      rptr[didx[p] + 1]++;
  }
  std::partial_sum(rptr.begin(), rptr.end(), rptr.begin());
  std::vector<int> radj(rptr[n]), rfill(rptr.begin(), rptr.end() - 1);
  for (int i = 0; i < n; i++)
    for (int p = dptr[i]; p < dptr[i + 1]; p++)
      radj[rfill[didx[p]]++] = i;
  std::vector<int> order;
  order.reserve(n);
  for (int i = 0; i < n; i++)
    if (!indeg[i])
      order.push_back(i);
  for (size_t h = 0; h < order.size(); h++) {
    int i = order[h];
    for (int p = rptr[i]; p < rptr[i + 1]; p++) {
      int d = radj[p];
      level[d] = std::max(level[d], level[i] + 1);
      if (--indeg[d] == 0)
        order.push_back(d);
    }
  }
  if ((int)order.size() != n)
    return false;
  int nbasic = 0;
  for (int i = 0; i < n; i++)
    nbasic = std::max(nbasic, level[i] + 1);
  std::vector<int> bstart(nbasic + 1, 0), bcols(n);
  for (int i = 0; i < n; i++)
    bstart[level[i] + 1]++;
  std::partial_sum(bstart.begin(), bstart.end(), bstart.begin());
  {
    std::vector<int> bf(bstart.begin(), bstart.end() - 1);
    for (int i = 0; i < n; i++)
      bcols[bf[level[i]]++] = i;
  }

  // Fusion of basic levels into chains (007:89-144,146-210): a basic level
  // joins the current level when each of its columns has all positioned
  // dependencies in exactly one chain, and no two columns pick one chain.
  std::vector<int> pos(n, -1), next(n, -1), head, tail, lvl_first, hook(n),
      used(n, -1);
  int cur_first = 0;
  for (int bl = 0; bl < nbasic; bl++) {
    const int b0 = bstart[bl], b1 = bstart[bl + 1];
    bool fuse = bl > 0;
    for (int t = b0; fuse && t < b1; t++) {
      int col = bcols[t], found = -1;
      for (int p = dptr[col]; p < dptr[col + 1]; p++) {
        int pp = pos[didx[p]];
        if (pp == -1 || pp == found)
          continue;
        if (found != -1) {
          found = -2;
          break;
        }
        found = pp;
      }
      if (found < 0 || used[found] == bl) {
        fuse = false;
        break;
      }
      used[found] = bl;
      hook[t - b0] = found;
    }
    if (fuse) {
      for (int t = b0; t < b1; t++) {
        int col = bcols[t], ch = cur_first + hook[t - b0];
        next[tail[ch]] = col;
        tail[ch] = col;
        pos[col] = hook[t - b0];
      }
    } else {
      std::fill(pos.begin(), pos.end(), -1);
      cur_first = (int)head.size();
      lvl_first.push_back(cur_first);
      for (int t = b0; t < b1; t++) {
        pos[bcols[t]] = t - b0;
        head.push_back(bcols[t]);
        tail.push_back(bcols[t]);
      }
    }
  }
  const int nch = (int)head.size();
  out.start_levels = lvl_first;
  out.start_levels.push_back(nch);
  out.start_chains.assign(nch + 1, 0);
  out.chains.clear();
  out.chains.reserve(n);
  for (int c = 0; c < nch; c++) {
    out.start_chains[c] = (int)out.chains.size();
    for (int col = head[c]; col != -1; col = next[col])
      out.chains.push_back(col);
  }
  out.start_chains[nch] = (int)out.chains.size();
  return true;
}

void build_kernel_levels(const LevelSet &ls, const std::vector<int> &comp_of,
                         int ncomp, KernelLevels &out) {
  const int nlev = ls.nlev();
  out.level_chain_ptr = ls.start_levels;
  out.chain_col_ptr.assign(1, 0);
  out.cols.clear();
  out.max_comp_chains = 0;
  // per component: list of (level, begin, end)
  std::vector<std::vector<std::array<int, 3>>> per(ncomp);
  std::vector<int> idx;
  for (int L = 0; L < nlev; L++) {
    int c0 = ls.start_levels[L], c1 = ls.start_levels[L + 1];
    idx.resize(c1 - c0);
    std::iota(idx.begin(), idx.end(), c0);
    auto comp = [&](int c) { return comp_of[ls.chains[ls.start_chains[c]]]; };
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int b) { return comp(a) < comp(b); });
    int base = (int)out.chain_col_ptr.size() - 1;
    for (size_t k = 0; k < idx.size(); k++) {
      int c = idx[k];
      for (int t = ls.start_chains[c]; t < ls.start_chains[c + 1]; t++)
        out.cols.push_back(ls.chains[t]);
      out.chain_col_ptr.push_back((int)out.cols.size());
    }
    for (size_t k = 0; k < idx.size();) {
      size_t e = k;
      int cp = comp(idx[k]);
      while (e < idx.size() && comp(idx[e]) == cp)
        e++;
      per[cp].push_back({L, base + (int)k, base + (int)e});
      out.max_comp_chains = std::max(out.max_comp_chains, (int)(e - k));
      k = e;
    }
  }
  out.comp_ptr.assign(ncomp + 1, 0);
  out.comp_lv_begin.clear();
  out.comp_lv_end.clear();
  out.comp_lv_level.clear();
  for (int c = 0; c < ncomp; c++) {
    for (auto &e : per[c]) {
      out.comp_lv_begin.push_back(e[1]);
      out.comp_lv_end.push_back(e[2]);
      out.comp_lv_level.push_back(e[0]);
    }
    out.comp_ptr[c + 1] = (int)out.comp_lv_begin.size();
  }
}

// Register programs.  A column qualifies when it has at most 4 entries, at
// most 4 updates and at most 2 forward-substitution terms (the series strings
// of a PV plant: tridiagonal eliminations).  A maximal run of qualifying
// consecutive columns of a chain runs on the register path; the other columns
// of the chain (a DC bus, say) run the generic column code between runs.
// Refactor record (16 ints) for the column at kernel position t:
//   [0] first CSC slot  [1] column  [2] ne | diag<<3 | nops<<6 | nlr<<9
//   [3] unused          [4+2k] L operand of update k: >=0 CSC slot, <0 entry
//   -1-e of the previous column of the run (a register)
//   [5+2k] target entry | U entry<<2 (U operands are entries of this column)
//   [12+2k] L operand of forward term k (same encoding)
//   [13+2k] y operand: >=0 row, -1 the previous column's y (a register)
// Row record (8 ints) for the row at kernel position t of the U levels
// (backward) or of the L levels (forward, without the division):
//   [0] row  [1] diagonal slot  [2] nterms  [3] unused
//   [4+2k] U (L) slot of term k minus the diagonal slot
//   [5+2k] x operand: >=0 row, -1 previous row's x
// Strided segments: consecutive positions of a run whose records agree in
// every field but [0] and [1], which advance by constant steps (a series
// string: the same tridiagonal elimination step column after column), share
// one record; seg[4t..4t+2] = (length, step of [0], step of [1]) at each
// segment start, so the kernel loads one record per segment and computes the
// slots and rows of the following columns; seg[4t+3] flags segments whose
// operands all come from the previous column's registers.
// run[t] = end position of the register run that contains t (0 if t does not
// qualify); chain flag = 1 when one run covers the whole chain.
static constexpr int kFastMaxE = 4, kFastMaxOps = 4, kFastMaxLr = 2;

// With the fused forward substitution the record also carries the L row of
// the column (at most kFastMaxLr terms); without it the row is left out.
static bool lu_column_qualifies(const pvglu_plan *P, int c) {
  const int q0 = P->cp[c], ne = P->cp[c + 1] - q0;
  if (ne > kFastMaxE)
    return false;
  if (P->fuse_lsolve && P->lrow_ptr[c + 1] - P->lrow_ptr[c] > kFastMaxLr)
    return false;
  return P->op_ptr[q0 + ne] - P->op_ptr[q0] <= kFastMaxOps;
}

// Operands in the previous column's register window are encoded as
// -1 - (position in the window); the window is the previous column's last
// four entries (all of them for a register column, which has at most four).
static void encode_lu_record(const pvglu_plan *P, int c, int prev, int *r) {
  const int q0 = P->cp[c], ne = P->cp[c + 1] - q0;
  const int d = P->diag_slot[c] - q0;
  const int w0 = prev >= 0 ? std::max(P->cp[prev], P->cp[prev + 1] - 4) : 0;
  auto lsrc_of = [&](int sl) {
    if (prev >= 0 && sl >= w0 && sl < P->cp[prev + 1])
      return -1 - (sl - w0);
    return sl;
  };
  int nops = 0;
  for (int q = q0; q < q0 + ne; q++)
    for (int o = P->op_ptr[q]; o < P->op_ptr[q + 1]; o++, nops++) {
      r[4 + 2 * nops] = lsrc_of(P->op_l[o]);
      r[5 + 2 * nops] = (q - q0) | ((P->op_u[o] - q0) << 2);
    }
  const int nlr = P->fuse_lsolve ? P->lrow_ptr[c + 1] - P->lrow_ptr[c] : 0;
  for (int k = 0; k < nlr; k++) {
    const int j = P->lrow_col[P->lrow_ptr[c] + k];
    r[12 + 2 * k] = lsrc_of(P->lrow_slot[P->lrow_ptr[c] + k]);
    r[13 + 2 * k] = (j == prev) ? -1 : j;
  }
  r[0] = q0;
  r[1] = c;
  r[2] = ne | (d << 3) | (nops << 6) | (nlr << 9);
}

static void encode_l_record(const pvglu_plan *P, int i, int prev, int *r) {
  const int nt = P->lrow_ptr[i + 1] - P->lrow_ptr[i];
  r[0] = i;
  r[1] = P->diag_slot[i];
  r[2] = nt;
  for (int k = 0; k < nt; k++) {
    const int j = P->lrow_col[P->lrow_ptr[i] + k];
    r[4 + 2 * k] = P->lrow_slot[P->lrow_ptr[i] + k] - P->diag_slot[i];
    r[5 + 2 * k] = (j == prev) ? -1 : j;
  }
}

// Transposed sweeps (adjoint): column j of U above its diagonal (U^T, with
// the division) or of L below it (L^T), as row records over CSC slots.
static void encode_t_record(const pvglu_plan *P, int j, int prev, bool upper,
                            int *r) {
  const int d = P->diag_slot[j];
  const int q0 = upper ? P->cp[j] : d + 1, q1 = upper ? d : P->cp[j + 1];
  r[0] = j;
  r[1] = d;
  r[2] = q1 - q0;
  for (int k = 0; k < q1 - q0; k++) {
    const int i = P->ri[q0 + k];
    r[4 + 2 * k] = q0 + k - d;
    r[5 + 2 * k] = (i == prev) ? -1 : i;
  }
}

static void encode_u_record(const pvglu_plan *P, int i, int prev, int *r) {
  const int nt = P->urow_ptr[i + 1] - P->urow_ptr[i];
  r[0] = i;
  r[1] = P->diag_slot[i];
  r[2] = nt;
  for (int k = 0; k < nt; k++) {
    const int j = P->urow_col[P->urow_ptr[i] + k];
    r[4 + 2 * k] = P->urow_slot[P->urow_ptr[i] + k] - P->diag_slot[i];
    r[5 + 2 * k] = (j == prev) ? -1 : j;
  }
}

// Shared run splitting: positions of each chain are grouped into maximal runs
// of qualifying columns; records are encoded with the previous column of the
// run.  A run's first column is encoded against lead(previous column of the
// chain) -- the column itself when the kernel leaves its register window
// behind (a star column), else -1 (no forwarding).
template <class Q, class E, class LEAD>
static int compile_runs(const KernelLevels &K, int n, int rec_ints, Q qualifies,
                        E encode, std::vector<int> &rec, std::vector<int> &run,
                        std::vector<int> &chain_flag, LEAD lead) {
  const int nch = (int)K.chain_col_ptr.size() - 1;
  rec.assign((size_t)std::max(n, 1) * rec_ints, 0);
  run.assign(std::max(n, 1), 0);
  chain_flag.assign(std::max(nch, 1), 0);
  int whole = 0;
  for (int ch = 0; ch < nch; ch++) {
    const int t0 = K.chain_col_ptr[ch], t1 = K.chain_col_ptr[ch + 1];
    int t = t0;
    while (t < t1) {
      if (!qualifies(K.cols[t])) {
        ++t;
        continue;
      }
      int e = t;
      while (e < t1 && qualifies(K.cols[e]))
        ++e;
      for (int u = t; u < e; u++) {
        encode(K.cols[u],
               u > t ? K.cols[u - 1] : (u > t0 ? lead(K.cols[u - 1]) : -1),
               &rec[(size_t)u * rec_ints]);
        run[u] = e;
      }
      if (t == t0 && e == t1) {
        chain_flag[ch] = 1;
        ++whole;
      }
      t = e;
    }
  }
  return whole;
}

template <class F>
static void compile_segments(const std::vector<int> &rec, int rec_ints,
                             const std::vector<int> &run, int n,
                             std::vector<int> &seg, F forwarded) {
  seg.assign((size_t)std::max(n, 1) * 4, 0);
  auto same = [&](int a, int b) {
    for (int k = 2; k < rec_ints; k++)
      if (rec[(size_t)a * rec_ints + k] != rec[(size_t)b * rec_ints + k])
        return false;
    return true;
  };
  auto step = [&](int a, int k) {
    return rec[(size_t)(a + 1) * rec_ints + k] - rec[(size_t)a * rec_ints + k];
  };
  for (int t = 0; t < n;) {
    const int e = run[t];
    if (e <= t) {
      ++t;
      continue;
    }
    // t starts a run [t, e)
    for (int u = t; u < e;) {
      int len = 1, d0 = 0, d1 = 0;
      if (u + 1 < e && same(u, u + 1)) {
        d0 = step(u, 0);
        d1 = step(u, 1);
        len = 2;
        while (u + len < e && same(u + len - 1, u + len) &&
               step(u + len - 1, 0) == d0 && step(u + len - 1, 1) == d1)
          ++len;
      }
      seg[(size_t)u * 4] = len;
      seg[(size_t)u * 4 + 1] = d0;
      seg[(size_t)u * 4 + 2] = d1;
      seg[(size_t)u * 4 + 3] = forwarded(&rec[(size_t)u * rec_ints]);
      u += len;
    }
    t = e;
  }
}

void compile_fast_programs(pvglu_plan *P) {
  auto none = [](int) { return -1; };
  // refactor: a run after a star column forwards from its window (the star
  // path and the generic gather both leave it in registers)
  P->fast_lu_chains = compile_runs(
      P->k_lu, P->n, 16, [&](int c) { return lu_column_qualifies(P, c); },
      [&](int c, int prev, int *r) { encode_lu_record(P, c, prev, r); },
      P->fast_lu_rec, P->fast_lu_run, P->fast_lu_chain,
      [&](int c) {
        // the kernels leave a generic column's last four entries in the
        // window; such a column has at least four (a column failing the
        // register path with fewer would need more than two fused forward
        // terms, i.e. at least three U entries), and the guard keeps the
        // encoding consistent whatever the pattern
        return ((P->col_flags[c] & 2) && P->cp[c + 1] - P->cp[c] >= 4) ? c
                                                                       : -1;
      });
  // star columns (never on the register path: more than four entries or
  // forward terms) get their 16-int record at their position: q0, q1,
  // diagonal slot, flags, forward row range, update count, then (first,
  // step) of the L slots, U slots and forward rows
  for (int t = 0; t < P->n; t++) {
    const int c = P->k_lu.cols[t];
    if (P->fast_lu_run[t] > t || !(P->col_flags[c] & 2))
      continue;
    const int qd = P->diag_slot[c];
    const int o0 = P->op_ptr[qd], nd = P->op_ptr[qd + 1] - o0;
    const int o1 = o0 + std::min(1, nd - 1);
    const int r0 = P->lrow_ptr[c], r1 = P->lrow_ptr[c + 1];
    int *r = &P->fast_lu_rec[(size_t)t * 16];
    r[0] = P->cp[c];
    r[1] = P->cp[c + 1];
    r[2] = qd;
    r[3] = P->col_flags[c];
    r[4] = r0;
    r[5] = r1;
    r[6] = nd;
    r[7] = P->op_l[o0];
    r[8] = P->op_l[o1] - P->op_l[o0];
    r[9] = P->op_u[o0];
    r[10] = P->op_u[o1] - P->op_u[o0];
    if (P->col_flags[c] & 1) {
      r[11] = P->lrow_col[r0];
      r[12] = P->lrow_col[r0 + (o1 - o0)] - P->lrow_col[r0];
    }
  }
  P->fast_u_chains = compile_runs(
      P->k_u, P->n, 8,
      [&](int i) {
        return P->urow_ptr[i + 1] - P->urow_ptr[i] <= kFastMaxLr;
      },
      [&](int i, int prev, int *r) { encode_u_record(P, i, prev, r); },
      P->fast_u_rec, P->fast_u_run, P->fast_u_chain, none);
  // seg[4t+3] bit 1: every update operand of the record comes from the
  // previous column's registers; bit 0: so do the forward-substitution
  // operands.  Such a segment issues no pattern-addressed loads: the kernel
  // runs it as a static three-stage load pipeline (bit 0 with the fused
  // forward substitution, bit 1 for the refactor alone).
  // Bits 2 and 3 mark the tridiagonal step of a series string (the common
  // record): three entries (U above, diagonal, L below), one update of the
  // diagonal by the previous column's entry k (bits 4-5) times the U entry,
  // and (bit 3) one forward term with the same L operand times the previous
  // y.  The kernel runs such segments from a branch-free specialization.
  compile_segments(P->fast_lu_rec, 16, P->fast_lu_run, P->n, P->fast_lu_seg,
                   [](const int *r) {
                     const int ne = r[2] & 7, d = (r[2] >> 3) & 7;
                     const int nops = (r[2] >> 6) & 7, nlr = (r[2] >> 9) & 3;
                     for (int k = 0; k < nops; k++)
                       if (r[4 + 2 * k] >= 0)
                         return 0;
                     const bool lrow_fwd =
                         nlr == 0 || (nlr == 1 && r[12] < 0 && r[13] == -1);
                     int flags = lrow_fwd ? 3 : 2;
                     if (ne == 3 && d == 1 && nops == 1 && r[5] == 1) {
                       flags |= 4 | ((-1 - r[4]) << 4);
                       if (nlr == 1 && r[12] == r[4] && r[13] == -1)
                         flags |= 8;
                     }
                     return flags;
                   });
  auto row_forwarded = [](const int *r) {
    return (r[2] == 0 || (r[2] == 1 && r[5] == -1)) ? 1 : 0;
  };
  compile_segments(P->fast_u_rec, 8, P->fast_u_run, P->n, P->fast_u_seg,
                   row_forwarded);
  P->fast_l_chains = compile_runs(
      P->k_l, P->n, 8,
      [&](int i) {
        return P->lrow_ptr[i + 1] - P->lrow_ptr[i] <= kFastMaxLr;
      },
      [&](int i, int prev, int *r) { encode_l_record(P, i, prev, r); },
      P->fast_l_rec, P->fast_l_run, P->fast_l_chain, none);
  compile_segments(P->fast_l_rec, 8, P->fast_l_run, P->n, P->fast_l_seg,
                   row_forwarded);
  P->fast_ut_chains = compile_runs(
      P->k_lu, P->n, 8,
      [&](int j) { return P->diag_slot[j] - P->cp[j] <= kFastMaxLr; },
      [&](int j, int prev, int *r) { encode_t_record(P, j, prev, true, r); },
      P->fast_ut_rec, P->fast_ut_run, P->fast_ut_chain, none);
  compile_segments(P->fast_ut_rec, 8, P->fast_ut_run, P->n, P->fast_ut_seg,
                   row_forwarded);
  P->fast_lt_chains = compile_runs(
      P->k_lt, P->n, 8,
      [&](int j) { return P->cp[j + 1] - P->diag_slot[j] - 1 <= kFastMaxLr; },
      [&](int j, int prev, int *r) { encode_t_record(P, j, prev, false, r); },
      P->fast_lt_rec, P->fast_lt_run, P->fast_lt_chain, none);
  compile_segments(P->fast_lt_rec, 8, P->fast_lt_run, P->n, P->fast_lt_seg,
                   row_forwarded);
  P->fast_lu_segs = P->fast_u_segs = 0;
  for (int t = 0; t < P->n; t++) {
    P->fast_lu_segs += P->fast_lu_seg[(size_t)t * 4] > 0;
    P->fast_u_segs += P->fast_u_seg[(size_t)t * 4] > 0;
  }
}

} // namespace pvglu

using namespace pvglu;

// --------------------------------------------------------------- C-ABI

extern "C" const char *pvglu_last_error(void) { return pvglu::last_error(); }
extern "C" void pvglu_free(void *p) { free(p); }

static uint64_t fnv1a(uint64_t h, const void *data, size_t bytes) {
  const unsigned char *p = (const unsigned char *)data;
  for (size_t i = 0; i < bytes; i++) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

extern "C" int pvglu_plan_create(int n, const int *rowptr, const int *colidx,
                                 pvglu_plan **out) {
  clear_error();
  if (!out || n <= 0 || !rowptr || !colidx)
    return fail(PVGLU_E_ARG, "plan_create: bad arguments");
  *out = nullptr;
  if (rowptr[0] != 0)
    return fail(PVGLU_E_ARG, "plan_create: rowptr[0] != 0");
  for (int i = 0; i < n; i++) {
    if (rowptr[i + 1] < rowptr[i])
      return fail(PVGLU_E_ARG, "plan_create: rowptr not monotone at %d", i);
    bool diag = false;
    for (int p = rowptr[i]; p < rowptr[i + 1]; p++) {
      int c = colidx[p];
      if (c < 0 || c >= n || (p > rowptr[i] && colidx[p - 1] >= c))
        return fail(PVGLU_E_ARG, "plan_create: row %d not sorted/in range", i);
      diag |= c == i;
    }
    if (!diag)
      return fail(PVGLU_E_NODIAG, "plan_create: no diagonal in row %d", i);
  }
  auto *P = new pvglu_plan();
  P->n = n;
  P->nnz = rowptr[n];
  P->rp.assign(rowptr, rowptr + n + 1);
  P->ci.assign(colidx, colidx + P->nnz);
  const int nnz = P->nnz;

  // 010: CSC view, value maps, first-past-diagonal pointers (010:15-128)
  P->cp.assign(n + 1, 0);
  for (int p = 0; p < nnz; p++)
    P->cp[P->ci[p] + 1]++;
  std::partial_sum(P->cp.begin(), P->cp.end(), P->cp.begin());
  P->ri.resize(nnz);
  P->csr2csc.resize(nnz);
  P->csc2csr.resize(nnz);
  {
    std::vector<int> fill(P->cp.begin(), P->cp.end() - 1);
    for (int i = 0; i < n; i++)
      for (int p = P->rp[i]; p < P->rp[i + 1]; p++) {
        int q = fill[P->ci[p]]++;
        P->ri[q] = i;
        P->csr2csc[p] = q;
        P->csc2csr[q] = p;
      }
  }
  P->csr_diag.resize(n);
  P->csc_diag.resize(n);
  P->diag_slot.resize(n);
  for (int i = 0; i < n; i++) {
    int p = P->rp[i];
    while (p < P->rp[i + 1] && P->ci[p] <= i)
      p++;
    P->csr_diag[i] = p;
    int q = P->cp[i];
    while (q < P->cp[i + 1] && P->ri[q] <= i)
      q++;
    P->csc_diag[i] = q;
    P->diag_slot[i] = q - 1; // row sorted and diagonal present => ri[q-1]==i
  }

  // 006 + 007 + 008 for the refactor, the two substitutions and the adjoint
  std::vector<int> dptr, didx;
  const DepKind kinds[4] = {DEP_GLU, DEP_L, DEP_U, DEP_LT};
  LevelSet *sets[4] = {&P->lu, &P->l, &P->u, &P->lt};
  std::vector<int> glu_ptr, glu_idx;
  for (int k = 0; k < 4; k++) {
    dependencies(kinds[k], n, P->rp, P->ci, dptr, didx);
    if (!levels_with_chains(n, dptr, didx, *sets[k])) {
      delete P;
      return fail(PVGLU_E_ARG, "plan_create: dependency cycle");
    }
    if (k == 0) {
      glu_ptr = dptr;
      glu_idx = didx;
    }
  }
  // 009: widest filled row per refactor level (assumes symmetry, 009:16)
  P->max_per_level.assign(P->lu.nlev(), 0);
  for (int L = 0; L < P->lu.nlev(); L++)
    for (int c = P->lu.start_levels[L]; c < P->lu.start_levels[L + 1]; c++)
      for (int t = P->lu.start_chains[c]; t < P->lu.start_chains[c + 1]; t++) {
        int col = P->lu.chains[t];
        P->max_per_level[L] =
            std::max(P->max_per_level[L], P->rp[col + 1] - P->rp[col]);
      }

  // Gather schedule of the refactor.  Column c, U entry (j,c), j<c ascending,
  // times L entries (i,j), i>j, updates target (i,c) -- exactly the updates
  // GLU_CSC_CPU_sequential (PV/GLU_CSC_CPU_levels_strings.cpp:37-62) applies
  // to column c, in the same order.  A missing target means the pattern is
  // not closed under elimination.
  {
    std::vector<int> slot_of_row(n, -1);
    std::vector<std::vector<std::pair<int, int>>> ops_of(0);
    std::vector<int> cnt(nnz + 1, 0);
    std::vector<int> tl, tu, tq; // (target slot, L slot, U slot)
    for (int c = 0; c < n; c++) {
      for (int q = P->cp[c]; q < P->cp[c + 1]; q++)
        slot_of_row[P->ri[q]] = q;
      for (int qu = P->cp[c]; qu < P->cp[c + 1] && P->ri[qu] < c; qu++) {
        int j = P->ri[qu];
        for (int ql = P->csc_diag[j]; ql < P->cp[j + 1]; ql++) {
          int i = P->ri[ql];
          int tgt = slot_of_row[i];
          if (tgt < 0) {
            delete P;
            return fail(PVGLU_E_ARG,
                        "plan_create: pattern not closed: (%d,%d) missing", i, c);
          }
          tq.push_back(tgt);
          tl.push_back(ql);
          tu.push_back(qu);
          cnt[tgt + 1]++;
        }
      }
      for (int q = P->cp[c]; q < P->cp[c + 1]; q++)
        slot_of_row[P->ri[q]] = -1;
    }
    std::partial_sum(cnt.begin(), cnt.end(), cnt.begin());
    P->op_ptr = cnt;
    P->op_l.resize(tq.size());
    P->op_u.resize(tq.size());
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (size_t k = 0; k < tq.size(); k++) { // stable: keeps ascending j
      int o = fill[tq[k]]++;
      P->op_l[o] = tl[k];
      P->op_u[o] = tu[k];
    }
  }
  // Substitution row lists
  P->lrow_ptr.assign(n + 1, 0);
  P->urow_ptr.assign(n + 1, 0);
  for (int i = 0; i < n; i++) {
    for (int p = P->rp[i]; p < P->rp[i + 1]; p++) {
      int j = P->ci[p];
      if (j < i) {
        P->lrow_slot.push_back(P->csr2csc[p]);
        P->lrow_col.push_back(j);
      }
    }
    P->lrow_ptr[i + 1] = (int)P->lrow_slot.size();
    for (int p = P->rp[i + 1] - 1; p >= P->rp[i]; p--) {
      int j = P->ci[p];
      if (j > i) {
        P->urow_slot.push_back(P->csr2csc[p]);
        P->urow_col.push_back(j);
      }
    }
    P->urow_ptr[i + 1] = (int)P->urow_slot.size();
  }
  // The forward substitution can ride in the refactor sweep when every L(c,j)
  // of row c is also a refactor dependency of column c, i.e. (j,c) present.
  {
    bool ok = true;
    std::vector<int> mark(n, -1);
    for (int c = 0; c < n && ok; c++) {
      for (int p = glu_ptr[c]; p < glu_ptr[c + 1]; p++)
        mark[glu_idx[p]] = c;
      for (int k = P->lrow_ptr[c]; k < P->lrow_ptr[c + 1]; k++)
        if (mark[P->lrow_col[k]] != c) {
          ok = false;
          break;
        }
    }
    P->fuse_lsolve = ok;
  }
  // Column flags.  Bit 0: the diagonal updates and the forward-substitution
  // row use the same L operands in the same order (both sums in one pass).
  // Bit 1: a star column -- only the diagonal has updates (1 to kStarMaxUpdates
  // of them), so every U operand is an entry above the diagonal without
  // updates, i.e. its A value, and the update operands are arithmetic
  // progressions: L slots, U slots and (fused row) forward rows each with a
  // constant step (the DC bus column: one update per string eliminated before
  // it, the strings' combiner columns numbered consecutively).  The kernel
  // addresses them from the first two of each.
  P->col_flags.assign(n, 0);
  for (int c = 0; c < n; c++) {
    const int qd = P->diag_slot[c];
    const int o0 = P->op_ptr[qd], o1 = P->op_ptr[qd + 1];
    if (P->fuse_lsolve) {
      const int r0 = P->lrow_ptr[c], r1 = P->lrow_ptr[c + 1];
      bool same = o1 - o0 == r1 - r0;
      for (int k = 0; same && k < o1 - o0; k++)
        same = P->op_l[o0 + k] == P->lrow_slot[r0 + k];
      P->col_flags[c] |= same ? 1 : 0;
    }
    const int q0 = P->cp[c], q1 = P->cp[c + 1];
    bool star = o1 - o0 >= 1 && o1 - o0 <= kStarMaxUpdates &&
                P->op_ptr[q1] - P->op_ptr[q0] == o1 - o0;
    auto progression = [&](const int *v, int len) {
      for (int k = 2; k < len; k++)
        if (v[k] - v[k - 1] != v[1] - v[0])
          return false;
      return true;
    };
    if (star)
      star = progression(&P->op_l[o0], o1 - o0) &&
             progression(&P->op_u[o0], o1 - o0) &&
             (!(P->col_flags[c] & 1) ||
              progression(&P->lrow_col[P->lrow_ptr[c]], o1 - o0));
    P->col_flags[c] |= star ? 2 : 0;
  }

  // Components of the symmetrized pattern (union-find)
  {
    std::vector<int> parent(n);
    std::iota(parent.begin(), parent.end(), 0);
    auto find = [&](int x) {
      while (parent[x] != x)
        x = parent[x] = parent[parent[x]];
      return x;
    };
    for (int i = 0; i < n; i++)
      for (int p = P->rp[i]; p < P->rp[i + 1]; p++) {
        int a = find(i), b = find(P->ci[p]);
        if (a != b)
          parent[std::max(a, b)] = std::min(a, b);
      }
    P->comp_of.assign(n, -1);
    std::vector<int> id(n, -1);
    int nc = 0;
    for (int i = 0; i < n; i++) {
      int r = find(i);
      if (id[r] < 0)
        id[r] = nc++;
      P->comp_of[i] = id[r];
    }
    P->ncomp = nc;
  }
  build_kernel_levels(P->lu, P->comp_of, P->ncomp, P->k_lu);
  build_kernel_levels(P->l, P->comp_of, P->ncomp, P->k_l);
  build_kernel_levels(P->u, P->comp_of, P->ncomp, P->k_u);
  build_kernel_levels(P->lt, P->comp_of, P->ncomp, P->k_lt);
  compile_fast_programs(P);

  uint64_t h = 1469598103934665603ull;
  h = fnv1a(h, &n, sizeof n);
  h = fnv1a(h, P->rp.data(), sizeof(int) * P->rp.size());
  h = fnv1a(h, P->ci.data(), sizeof(int) * P->ci.size());
  P->hash = h;
  *out = P;
  return PVGLU_OK;
}

extern "C" void pvglu_plan_destroy(pvglu_plan *plan) {
  if (!plan)
    return;
  device_plan_destroy(plan);
  delete plan;
}

extern "C" int pvglu_plan_info(const pvglu_plan *P, int what) {
  if (!P)
    return fail(PVGLU_E_ARG, "plan_info: null plan");
  switch (what) {
  case PVGLU_INFO_N: return P->n;
  case PVGLU_INFO_NNZ: return P->nnz;
  case PVGLU_INFO_NLEV_LU: return P->lu.nlev();
  case PVGLU_INFO_NCHAIN_LU: return P->lu.nchains();
  case PVGLU_INFO_NLEV_L: return P->l.nlev();
  case PVGLU_INFO_NCHAIN_L: return P->l.nchains();
  case PVGLU_INFO_NLEV_U: return P->u.nlev();
  case PVGLU_INFO_NCHAIN_U: return P->u.nchains();
  case PVGLU_INFO_NLEV_LT: return P->lt.nlev();
  case PVGLU_INFO_NCHAIN_LT: return P->lt.nchains();
  case PVGLU_INFO_NOPS: return (int)P->op_l.size();
  case PVGLU_INFO_FUSED_LSOLVE: return P->fuse_lsolve ? 1 : 0;
  case PVGLU_INFO_NCOMP: return P->ncomp;
  case PVGLU_INFO_FAST_LU_CHAINS: return P->fast_lu_chains;
  case PVGLU_INFO_FAST_U_CHAINS: return P->fast_u_chains;
  case PVGLU_INFO_FAST_LU_SEGMENTS: return P->fast_lu_segs;
  case PVGLU_INFO_FAST_U_SEGMENTS: return P->fast_u_segs;
  default: return fail(PVGLU_E_ARG, "plan_info: unknown field %d", what);
  }
}

static const std::vector<int> *plan_array(const pvglu_plan *P, int which) {
  switch (which) {
  case PVGLU_ARR_CSR_ROWPTR: return &P->rp;
  case PVGLU_ARR_CSR_COLIDX: return &P->ci;
  case PVGLU_ARR_CSC_COLPTR: return &P->cp;
  case PVGLU_ARR_CSC_ROWIDX: return &P->ri;
  case PVGLU_ARR_CSR2CSC: return &P->csr2csc;
  case PVGLU_ARR_CSC2CSR: return &P->csc2csr;
  case PVGLU_ARR_CSR_DIAG: return &P->csr_diag;
  case PVGLU_ARR_CSC_DIAG: return &P->csc_diag;
  case PVGLU_ARR_LU_START_LEVELS: return &P->lu.start_levels;
  case PVGLU_ARR_LU_START_CHAINS: return &P->lu.start_chains;
  case PVGLU_ARR_LU_CHAINS: return &P->lu.chains;
  case PVGLU_ARR_L_START_LEVELS: return &P->l.start_levels;
  case PVGLU_ARR_L_START_CHAINS: return &P->l.start_chains;
  case PVGLU_ARR_L_CHAINS: return &P->l.chains;
  case PVGLU_ARR_U_START_LEVELS: return &P->u.start_levels;
  case PVGLU_ARR_U_START_CHAINS: return &P->u.start_chains;
  case PVGLU_ARR_U_CHAINS: return &P->u.chains;
  case PVGLU_ARR_MAX_PER_LEVEL: return &P->max_per_level;
  case PVGLU_ARR_COL_FLAGS: return &P->col_flags;
  case PVGLU_ARR_LU_SEGMENTS: return &P->fast_lu_seg;
  default: return nullptr;
  }
}

extern "C" int pvglu_plan_array_len(const pvglu_plan *P, int which) {
  if (!P)
    return fail(PVGLU_E_ARG, "plan_array_len: null plan");
  auto *a = plan_array(P, which);
  if (!a)
    return fail(PVGLU_E_ARG, "plan_array_len: unknown array %d", which);
  return (int)a->size();
}

extern "C" int pvglu_plan_export(const pvglu_plan *P, int which, int *out) {
  if (!P || !out)
    return fail(PVGLU_E_ARG, "plan_export: bad arguments");
  auto *a = plan_array(P, which);
  if (!a)
    return fail(PVGLU_E_ARG, "plan_export: unknown array %d", which);
  if (!a->empty())
    memcpy(out, a->data(), sizeof(int) * a->size());
  return PVGLU_OK;
}
