// pvglu_amd.cpp -- fill-reducing ordering and zero-diagonal row swaps.
//
// ---------------------------------------------------------------------------
// Third-party notice.  pvglu_amd_order (everything in namespace amd below) is
// a MODIFIED VERSION of AMD: a C++ translation of AMD 1.1's amd_aat.c,
// amd_1.c, amd_2.c and amd_postorder.c that keeps AMD's algorithm, step
// order, tie-breaks and many of its variable names (knt1/knt2, pme1/pme2,
// elenme, nvpiv, degme, wnvi, dext, lemax, wflg/wbig).  Modifications: one
// class over index vectors instead of one integer workspace, workspace growth
// instead of garbage collection, the dense-row and postorder steps folded into
// the driver, no Info/Control arrays.
//
//   AMD Version 1.1 (Jan. 21, 2004), Copyright (c) 2004 by Timothy A. Davis,
//   Patrick R. Amestoy, and Iain S. Duff.
//   Availability: http://www.cise.ufl.edu/research/sparse/amd
//   Used by permission.
//
// AMD 1.1 is distributed under the AMD License of its README (the reference
// vendors the sources, PV/AmdLib/amd.h:5-9, which point to that README, but
// not the README itself).  That license permits use, copying, modification
// and distribution of modified code provided that the Copyright, the License
// and the Availability note are retained and a notice that the code was
// modified is included; AMD is provided AS IS, WITH ABSOLUTELY NO WARRANTY
// EXPRESSED OR IMPLIED, and any use is at the user's own risk.  This block
// is that notice.
// ---------------------------------------------------------------------------
//
// pvglu_amd_order restates the approximate minimum degree ordering of
// Amestoy, Davis and Duff (ACM TOMS 30(3) 2004, Algorithm 837; SIAM J. Matrix
// Anal. Appl. 17(4) 1996) as AMD 1.1 runs it with default controls (dense
// threshold max(16, 10 sqrt(n)), aggressive absorption), the version the
// reference vendors (PV/AmdLib, amd.h:6) and calls from
// algorithm_amd_cpp (PV/004a_algorithm_amd_function.cpp:10-50).  The steps,
// their order and every tie-break follow that algorithm, because the
// permutation is a bit-exact parity target:
//   * pattern of A + A^T built with the same interleaved column scan, so each
//     variable's adjacency list has the same order;
//   * degree lists are LIFO; the pivot is the head of the lowest nonempty list;
//   * the new element's variable list is gathered from the pivot's elements
//     in list order, then its own variables;
//   * approximate external degrees via w(e) = |Le \ Lme|, aggressive
//     element absorption, mass elimination of variables left only in Lme;
//   * indistinguishable variables found by hashing (sum of list entries
//     mod n) and pairwise comparison in bucket order;
//   * assembly-tree postorder with the largest-front child last.
// The workspace grows instead of being garbage-collected: compaction only
// moves lists without reordering them, so the ordering is the same.
//
// pvglu_reorder adds the row-only swaps of remove_zeros_diagonal
// (PV/004b_remove_zeros_diagonal.cpp:12-100, move_only_rows = true), as
// reorder_matrix does (PV/004d_reorder_matrix.cpp:27-47).
#include "pvglu_internal.h"

#include <algorithm>
#include <cmath>
#include <set>
#include <vector>

using namespace pvglu;

namespace {

constexpr int kEmpty = -1;
inline int flip(int i) { return -i - 2; }

// Length of each column of A + A^T (diagonal excluded), as the nonzero count
// the ordering starts from.
void aat_lengths(int n, const int *Ap, const int *Ai, std::vector<int> &len,
                 std::vector<int> &tp) {
  len.assign(n, 0);
  tp.assign(n, 0);
  for (int k = 0; k < n; k++) {
    int p = Ap[k];
    const int p2 = Ap[k + 1];
    while (p < p2) {
      const int j = Ai[p];
      if (j < k) {
        len[j]++;
        len[k]++;
        p++;
      } else {
        if (j == k)
          p++;
        break;
      }
      // walk column j's lower part up to row k, resuming where it stopped
      int pj = tp[j];
      const int pj2 = Ap[j + 1];
      while (pj < pj2) {
        const int i = Ai[pj];
        if (i < k) {
          len[i]++;
          len[j]++;
          pj++;
        } else {
          if (i == k)
            pj++;
          break;
        }
      }
      tp[j] = pj;
    }
    tp[k] = p;
  }
  for (int j = 0; j < n; j++)
    for (int pj = tp[j]; pj < Ap[j + 1]; pj++) {
      len[Ai[pj]]++;
      len[j]++;
    }
}

// Adjacency lists of A + A^T in the order the same scan discovers them.
void aat_lists(int n, const int *Ap, const int *Ai, const std::vector<int> &len,
               std::vector<int> &pe, std::vector<int> &iw) {
  std::vector<int> sp(n), tp(n, 0);
  int pfree = 0;
  pe.assign(n, 0);
  for (int j = 0; j < n; j++) {
    pe[j] = sp[j] = pfree;
    pfree += len[j];
  }
  for (int k = 0; k < n; k++) {
    int p = Ap[k];
    const int p2 = Ap[k + 1];
    while (p < p2) {
      const int j = Ai[p];
      if (j < k) {
        iw[sp[j]++] = k;
        iw[sp[k]++] = j;
        p++;
      } else {
        if (j == k)
          p++;
        break;
      }
      int pj = tp[j];
      const int pj2 = Ap[j + 1];
      while (pj < pj2) {
        const int i = Ai[pj];
        if (i < k) {
          iw[sp[i]++] = j;
          iw[sp[j]++] = i;
          pj++;
        } else {
          if (i == k)
            pj++;
          break;
        }
      }
      tp[j] = pj;
    }
    tp[k] = p;
  }
  for (int j = 0; j < n; j++)
    for (int pj = tp[j]; pj < Ap[j + 1]; pj++) {
      const int i = Ai[pj];
      iw[sp[i]++] = j;
      iw[sp[j]++] = i;
    }
}

class MinDegree {
public:
  MinDegree(int n, std::vector<int> &pe, std::vector<int> &iw,
            std::vector<int> &len, int pfree)
      : n_(n), pe_(pe), iw_(iw), len_(len), pfree_(pfree), nv_(n, 1),
        next_(n, kEmpty), last_(n, kEmpty), head_(n, kEmpty), elen_(n, 0),
        degree_(len), w_(n, 1) {}

  void run(int *perm) {
    int dense = (int)(10.0 * std::sqrt((double)n_));
    dense = std::min(n_, std::max(16, dense));
    wbig_ = 2147483647 - n_;
    // initial degree lists; isolated variables are eliminated at once and
    // dense ones are set aside to be ordered last
    for (int i = 0; i < n_; i++) {
      const int deg = degree_[i];
      if (deg == 0) {
        elen_[i] = flip(1);
        nel_++;
        pe_[i] = kEmpty;
        w_[i] = 0;
      } else if (deg > dense) {
        ndense_++;
        nv_[i] = 0;
        elen_[i] = kEmpty;
        nel_++;
        pe_[i] = kEmpty;
      } else {
        push_degree(i, deg);
      }
    }
    while (nel_ < n_)
      eliminate_next();
    finish(perm);
  }

private:
  void push_degree(int i, int deg) {
    const int in = head_[deg];
    if (in != kEmpty)
      last_[in] = i;
    next_[i] = in;
    head_[deg] = i;
  }
  void unlink_degree(int i) {
    const int il = last_[i], in = next_[i];
    if (in != kEmpty)
      last_[in] = il;
    if (il != kEmpty)
      next_[il] = in;
    else
      head_[degree_[i]] = in;
  }
  int &at(int p) {
    if (p >= (int)iw_.size())
      iw_.resize(std::max<size_t>(iw_.size() * 2, p + 1));
    return iw_[p];
  }
  void reset_marks_if_needed() {
    if (wflg_ >= wbig_) {
      for (int x = 0; x < n_; x++)
        if (w_[x] != 0)
          w_[x] = 1;
      wflg_ = 2;
    }
  }

  void eliminate_next() {
    // pivot: head of the lowest nonempty degree list
    int deg = mindeg_;
    int me = kEmpty;
    for (; deg < n_; deg++) {
      me = head_[deg];
      if (me != kEmpty)
        break;
    }
    mindeg_ = deg;
    {
      const int in = next_[me];
      if (in != kEmpty)
        last_[in] = kEmpty;
      head_[deg] = in;
    }
    const int elenme = elen_[me];
    int nvpiv = nv_[me];
    nel_ += nvpiv;
    nv_[me] = -nvpiv;
    int degme = 0;
    int pme1, pme2;

    // Construct the new element Lme.  Variables are flagged by negating nv.
    if (elenme == 0) {
      // no elements adjacent: Lme is built in place over me's variables
      pme1 = pe_[me];
      pme2 = pme1 - 1;
      for (int p = pme1; p <= pme1 + len_[me] - 1; p++) {
        const int i = iw_[p];
        const int nvi = nv_[i];
        if (nvi > 0) {
          degme += nvi;
          nv_[i] = -nvi;
          iw_[++pme2] = i;
          unlink_degree(i);
        }
      }
    } else {
      // union of the adjacent elements' lists and me's variables, appended
      int p = pe_[me];
      pme1 = pfree_;
      const int slenme = len_[me] - elenme;
      for (int knt1 = 1; knt1 <= elenme + 1; knt1++) {
        int e, pj, ln;
        if (knt1 > elenme) {
          e = me;
          pj = p;
          ln = slenme;
        } else {
          e = iw_[p++];
          pj = pe_[e];
          ln = len_[e];
        }
        for (int knt2 = 1; knt2 <= ln; knt2++) {
          const int i = iw_[pj++];
          const int nvi = nv_[i];
          if (nvi > 0) {
            degme += nvi;
            nv_[i] = -nvi;
            at(pfree_++) = i;
            unlink_degree(i);
          }
        }
        if (e != me) { // element e is absorbed into me
          pe_[e] = flip(me);
          w_[e] = 0;
        }
      }
      pme2 = pfree_ - 1;
    }
    degree_[me] = degme;
    pe_[me] = pme1;
    len_[me] = pme2 - pme1 + 1;
    elen_[me] = flip(nvpiv + degme);

    // w(e) = |Le \ Lme| for every element e adjacent to a variable of Lme
    reset_marks_if_needed();
    for (int pme = pme1; pme <= pme2; pme++) {
      const int i = iw_[pme];
      const int eln = elen_[i];
      if (eln > 0) {
        const int nvi = -nv_[i];
        const int wnvi = wflg_ - nvi;
        for (int p = pe_[i]; p <= pe_[i] + eln - 1; p++) {
          const int e = iw_[p];
          int we = w_[e];
          if (we >= wflg_)
            we -= nvi;
          else if (we != 0)
            we = degree_[e] + wnvi;
          w_[e] = we;
        }
      }
    }

    // degree update and element absorption for each variable of Lme
    for (int pme = pme1; pme <= pme2; pme++) {
      const int i = iw_[pme];
      const int p1 = pe_[i];
      const int p2 = p1 + elen_[i] - 1;
      int pn = p1;
      unsigned int hash = 0;
      int dg = 0;
      for (int p = p1; p <= p2; p++) { // aggressive absorption
        const int e = iw_[p];
        const int we = w_[e];
        if (we != 0) {
          const int dext = we - wflg_;
          if (dext > 0) {
            dg += dext;
            iw_[pn++] = e;
            hash += (unsigned int)e;
          } else {
            pe_[e] = flip(me);
            w_[e] = 0;
          }
        }
      }
      elen_[i] = pn - p1 + 1;
      const int p3 = pn;
      const int p4 = p1 + len_[i];
      for (int p = p2 + 1; p < p4; p++) {
        const int j = iw_[p];
        const int nvj = nv_[j];
        if (nvj > 0) {
          dg += nvj;
          iw_[pn++] = j;
          hash += (unsigned int)j;
        }
      }
      if (elen_[i] == 1 && p3 == pn) {
        // mass elimination: i is adjacent to me only
        pe_[i] = flip(me);
        const int nvi = -nv_[i];
        degme -= nvi;
        nvpiv += nvi;
        nel_ += nvi;
        nv_[i] = 0;
        elen_[i] = kEmpty;
      } else {
        degree_[i] = std::min(degree_[i], dg);
        // me becomes the first element of i's list
        iw_[pn] = iw_[p3];
        iw_[p3] = iw_[p1];
        iw_[p1] = me;
        len_[i] = pn - p1 + 1;
        hash = hash % (unsigned int)n_;
        const int j = head_[hash];
        if (j <= kEmpty) { // bucket empty or already a hash bucket
          next_[i] = flip(j);
          head_[hash] = flip(i);
        } else { // head_[hash] is a degree list: chain through last_
          next_[i] = last_[j];
          last_[j] = i;
        }
        last_[i] = (int)hash;
      }
    }
    degree_[me] = degme;
    lemax_ = std::max(lemax_, degme);
    wflg_ += lemax_;
    reset_marks_if_needed();

    // supervariable detection within each hash bucket
    for (int pme = pme1; pme <= pme2; pme++) {
      int i = iw_[pme];
      if (nv_[i] >= 0)
        continue;
      const unsigned int hash = (unsigned int)last_[i];
      const int j = head_[hash];
      if (j == kEmpty)
        i = kEmpty;
      else if (j < kEmpty) {
        i = flip(j);
        head_[hash] = kEmpty;
      } else {
        i = last_[j];
        last_[j] = kEmpty;
      }
      while (i != kEmpty && next_[i] != kEmpty) {
        const int ln = len_[i], eln = elen_[i];
        for (int p = pe_[i] + 1; p <= pe_[i] + ln - 1; p++)
          w_[iw_[p]] = wflg_;
        int jlast = i;
        int jj = next_[i];
        while (jj != kEmpty) {
          bool ok = len_[jj] == ln && elen_[jj] == eln;
          for (int p = pe_[jj] + 1; ok && p <= pe_[jj] + ln - 1; p++)
            if (w_[iw_[p]] != wflg_)
              ok = false;
          if (ok) { // jj is absorbed into supervariable i
            pe_[jj] = flip(i);
            nv_[i] += nv_[jj];
            nv_[jj] = 0;
            elen_[jj] = kEmpty;
            jj = next_[jj];
            next_[jlast] = jj;
          } else {
            jlast = jj;
            jj = next_[jj];
          }
        }
        wflg_++;
        i = next_[i];
      }
    }

    // finalize the new element: restore nv, put survivors back in the
    // degree lists with their external degree plus |Lme|
    int p = pme1;
    const int nleft = n_ - nel_;
    for (int pme = pme1; pme <= pme2; pme++) {
      const int i = iw_[pme];
      const int nvi = -nv_[i];
      if (nvi > 0) {
        nv_[i] = nvi;
        int dg = degree_[i] + degme - nvi;
        dg = std::min(dg, nleft - nvi);
        push_degree(i, dg);
        last_[i] = kEmpty;
        mindeg_ = std::min(mindeg_, dg);
        degree_[i] = dg;
        iw_[p++] = i;
      }
    }
    nv_[me] = nvpiv;
    len_[me] = p - pme1;
    if (len_[me] == 0) {
      pe_[me] = kEmpty;
      w_[me] = 0;
    }
    if (elenme != 0)
      pfree_ = p;
  }

  void finish(int *perm) {
    for (int i = 0; i < n_; i++) {
      pe_[i] = flip(pe_[i]);
      elen_[i] = flip(elen_[i]);
    }
    // every absorbed variable points at its principal element
    for (int i = 0; i < n_; i++) {
      if (nv_[i] != 0)
        continue;
      int j = pe_[i];
      if (j == kEmpty)
        continue;
      while (nv_[j] == 0)
        j = pe_[j];
      const int e = j;
      j = i;
      while (nv_[j] == 0) {
        const int jn = pe_[j];
        pe_[j] = e;
        j = jn;
      }
    }
    // postorder of the assembly tree; fsize = elen (front sizes)
    std::vector<int> order(n_, kEmpty);
    postorder(order);
    std::vector<int> byk(n_, kEmpty), pos(n_, kEmpty);
    for (int e = 0; e < n_; e++)
      if (order[e] != kEmpty)
        byk[order[e]] = e;
    int nel = 0;
    for (int k = 0; k < n_; k++) {
      const int e = byk[k];
      if (e == kEmpty)
        break;
      pos[e] = nel;
      nel += nv_[e];
    }
    for (int i = 0; i < n_; i++) {
      if (nv_[i] != 0)
        continue;
      const int e = pe_[i];
      if (e != kEmpty)
        pos[i] = pos[e]++;
      else
        pos[i] = nel++; // dense rows go last
    }
    for (int i = 0; i < n_; i++)
      perm[pos[i]] = i;
  }

  void postorder(std::vector<int> &order) {
    std::vector<int> child(n_, kEmpty), sibling(n_, kEmpty), stack(n_);
    for (int j = n_ - 1; j >= 0; j--)
      if (nv_[j] > 0 && pe_[j] != kEmpty) {
        sibling[j] = child[pe_[j]];
        child[pe_[j]] = j;
      }
    // move the child with the largest front (last one on ties) to the end
    for (int i = 0; i < n_; i++) {
      if (nv_[i] <= 0 || child[i] == kEmpty)
        continue;
      int fprev = kEmpty, maxfr = kEmpty, bigprev = kEmpty, big = kEmpty;
      for (int f = child[i]; f != kEmpty; f = sibling[f]) {
        if (elen_[f] >= maxfr) {
          maxfr = elen_[f];
          bigprev = fprev;
          big = f;
        }
        fprev = f;
      }
      const int fnext = sibling[big];
      if (fnext != kEmpty) {
        if (bigprev == kEmpty)
          child[i] = fnext;
        else
          sibling[bigprev] = fnext;
        sibling[big] = kEmpty;
        sibling[fprev] = big;
      }
    }
    int k = 0;
    for (int root = 0; root < n_; root++) {
      if (pe_[root] != kEmpty || nv_[root] <= 0)
        continue;
      // iterative depth-first postorder, children in sibling order
      int head = 0;
      stack[0] = root;
      while (head >= 0) {
        const int i = stack[head];
        if (child[i] != kEmpty) {
          for (int f = child[i]; f != kEmpty; f = sibling[f])
            head++;
          int h = head;
          for (int f = child[i]; f != kEmpty; f = sibling[f])
            stack[h--] = f;
          child[i] = kEmpty;
        } else {
          head--;
          order[i] = k++;
        }
      }
    }
  }

  int n_;
  std::vector<int> &pe_, &iw_, &len_;
  int pfree_;
  std::vector<int> nv_, next_, last_, head_, elen_, degree_, w_;
  int nel_ = 0, mindeg_ = 0, ndense_ = 0, lemax_ = 0, wflg_ = 2, wbig_ = 0;
};

} // namespace

extern "C" int pvglu_amd_order(int n, const int *Ap, const int *Ai, int *P) {
  clear_error();
  if (n < 0 || !Ap || !Ai || !P)
    return fail(PVGLU_E_ARG, "amd_order: bad arguments");
  if (n == 0)
    return PVGLU_OK;
  // input check of AMD 1.1: pointers monotone, indices sorted, in range,
  // no duplicates
  if (Ap[0] != 0)
    return fail(PVGLU_E_ARG, "amd_order: Ap[0] != 0");
  for (int j = 0; j < n; j++) {
    if (Ap[j] > Ap[j + 1])
      return fail(PVGLU_E_ARG, "amd_order: pointers not monotone");
    int last = -1;
    for (int p = Ap[j]; p < Ap[j + 1]; p++) {
      if (Ai[p] <= last || Ai[p] >= n)
        return fail(PVGLU_E_ARG, "amd_order: indices not sorted/in range");
      last = Ai[p];
    }
  }
  std::vector<int> len, tp, pe;
  aat_lengths(n, Ap, Ai, len, tp);
  long nzaat = 0;
  for (int v : len)
    nzaat += v;
  std::vector<int> iw((size_t)(nzaat + nzaat / 5 + n) + 1, 0);
  aat_lists(n, Ap, Ai, len, pe, iw);
  MinDegree md(n, pe, iw, len, (int)nzaat);
  md.run(P);
  return PVGLU_OK;
}

// Row-only swaps that remove structurally zero diagonals
// (PV/004b_remove_zeros_diagonal.cpp:12-100 with move_only_rows,
// PV/004c_exchange_rows_columns.cpp:74-89).  rows[k] is the current row k;
// cols[j] the set of rows with an entry in column j.
static int remove_zeros_diagonal(int n, std::vector<std::vector<int>> &rows,
                                 std::vector<int> &perm) {
  std::vector<std::set<int>> cols(n);
  for (int i = 0; i < n; i++)
    for (int j : rows[i])
      cols[j].insert(i);
  for (int r = 0; r < n; r++) {
    if (std::binary_search(rows[r].begin(), rows[r].end(), r))
      continue;
    // would (r,r) fill from some (r,c), (c,r) with c < r?
    bool would_fill = false;
    for (int c : rows[r]) {
      if (c >= r)
        break;
      if (cols[r].count(c)) {
        would_fill = true;
        break;
      }
    }
    if (would_fill)
      continue;
    auto it = cols[r].upper_bound(r);
    if (it == cols[r].end())
      return fail(PVGLU_E_ARG, "reorder: no row to swap into row %d", r);
    const int s = *it;
    for (int c : rows[r]) {
      cols[c].erase(r);
    }
    for (int c : rows[s]) {
      cols[c].erase(s);
    }
    std::swap(rows[r], rows[s]);
    for (int c : rows[r])
      cols[c].insert(r);
    for (int c : rows[s])
      cols[c].insert(s);
    std::swap(perm[r], perm[s]);
  }
  return PVGLU_OK;
}

extern "C" int pvglu_reorder(int n, const int *rowptr, const int *colidx,
                             int *P, int *Prow) {
  clear_error();
  if (n <= 0 || !rowptr || !colidx || !P || !Prow)
    return fail(PVGLU_E_ARG, "reorder: bad arguments");
  // AMD on the row-major boolean pattern (004a:13-32), then the symmetric
  // permutation new(Pinv[i], Pinv[j]) = old(i,j) (004a:44)
  int st = pvglu_amd_order(n, rowptr, colidx, P);
  if (st)
    return st;
  std::vector<int> pinv(n);
  for (int k = 0; k < n; k++)
    pinv[P[k]] = k;
  std::vector<std::vector<int>> rows(n);
  for (int i = 0; i < n; i++) {
    auto &dst = rows[pinv[i]];
    for (int p = rowptr[i]; p < rowptr[i + 1]; p++)
      dst.push_back(pinv[colidx[p]]);
    std::sort(dst.begin(), dst.end());
  }
  std::vector<int> perm(n);
  for (int i = 0; i < n; i++)
    perm[i] = i;
  st = remove_zeros_diagonal(n, rows, perm);
  if (st)
    return st;
  std::copy(perm.begin(), perm.end(), Prow);
  return PVGLU_OK;
}
