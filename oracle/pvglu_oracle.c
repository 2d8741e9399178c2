/*
 * pvglu_oracle.c -- CPU oracle for the PV Newton-step GLU path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker, never by the product library.
 * Built with -ffp-contract=off so the floating-point operation sequence is
 * the one written here (one rounding per + - * /).
 *
 * Each function restates the reference algorithm cited in pvglu_oracle.h;
 * the pins that check each function are listed there.
 */
#include "pvglu_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ utils */

static void *xmalloc(size_t bytes) {
  void *p = malloc(bytes ? bytes : 1);
  if (!p)
    abort();
  return p;
}

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

typedef struct {
  int *v;
  int len, cap;
} ivec;

static void ivec_push(ivec *a, int x) {
  if (a->len == a->cap) {
    a->cap = a->cap ? 2 * a->cap : 8;
    a->v = (int *)realloc(a->v, sizeof(int) * (size_t)a->cap);
    if (!a->v)
      abort();
  }
  a->v[a->len++] = x;
}

/* binary min-heap of ints */
static void heap_push(ivec *h, int x) {
  ivec_push(h, x);
  int i = h->len - 1;
  while (i > 0) {
    int p = (i - 1) / 2;
    if (h->v[p] <= h->v[i])
      break;
    int t = h->v[p];
    h->v[p] = h->v[i];
    h->v[i] = t;
    i = p;
  }
}

static int heap_pop(ivec *h) {
  int top = h->v[0];
  h->v[0] = h->v[--h->len];
  int i = 0;
  for (;;) {
    int l = 2 * i + 1, r = l + 1, m = i;
    if (l < h->len && h->v[l] < h->v[m])
      m = l;
    if (r < h->len && h->v[r] < h->v[m])
      m = r;
    if (m == i)
      break;
    int t = h->v[m];
    h->v[m] = h->v[i];
    h->v[i] = t;
    i = m;
  }
  return top;
}

/* ------------------------------------------------------ 005a symbolic fill */

int orc_symbolic_fill(int n, const int *rp, const int *ci, int **out_rp,
                      int **out_ci) {
  /* Row-merge symbolic factorization: row i of the filled pattern is
   * A(i,:) united with the U part (columns > k) of filled row k for every k
   * in the L part of filled row i, visited in increasing k (new L entries
   * created while merging are > k, so a min-heap order is complete). */
  ivec *rows = (ivec *)calloc((size_t)(n > 0 ? n : 1), sizeof(ivec));
  int *mark = (int *)xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++)
    mark[i] = -1;
  ivec heap = {0, 0, 0};
  for (int i = 0; i < n; i++) {
    ivec *row = &rows[i];
    heap.len = 0;
    for (int p = rp[i]; p < rp[i + 1]; p++) {
      int c = ci[p];
      if (mark[c] == i)
        continue;
      mark[c] = i;
      ivec_push(row, c);
      if (c < i)
        heap_push(&heap, c);
    }
    while (heap.len) {
      int k = heap_pop(&heap);
      const ivec *rk = &rows[k];
      for (int q = 0; q < rk->len; q++) {
        int c = rk->v[q];
        if (c <= k || mark[c] == i)
          continue;
        mark[c] = i;
        ivec_push(row, c);
        if (c < i)
          heap_push(&heap, c);
      }
    }
    qsort(row->v, (size_t)row->len, sizeof(int), cmp_int);
  }
  int nnz = 0;
  for (int i = 0; i < n; i++)
    nnz += rows[i].len;
  int *orp = (int *)xmalloc(sizeof(int) * (size_t)(n + 1));
  int *oci = (int *)xmalloc(sizeof(int) * (size_t)nnz);
  orp[0] = 0;
  for (int i = 0; i < n; i++) {
    memcpy(oci + orp[i], rows[i].v, sizeof(int) * (size_t)rows[i].len);
    orp[i + 1] = orp[i] + rows[i].len;
    free(rows[i].v);
  }
  free(rows);
  free(mark);
  free(heap.v);
  *out_rp = orp;
  *out_ci = oci;
  return nnz;
}

/* ------------------------------------------------------------- 005b scan */

void orc_inclusive_scan(int *a, int n) {
  for (int i = 1; i < n; i++)
    a[i] += a[i - 1];
}

/* -------------------------------------------------------- 006 dependencies */

int orc_deps(int kind, int n, const int *rp, const int *ci, int **dptr,
             int **didx) {
  int *cnt = (int *)calloc((size_t)n + 1, sizeof(int));
  for (int i = 0; i < n; i++)
    for (int p = rp[i]; p < rp[i + 1]; p++) {
      int j = ci[p];
      if (kind == 0 && i < j)
        cnt[j + 1]++; /* column j gains dependency i */
      else if (kind == 1 && j < i)
        cnt[i + 1]++;
      else if (kind == 2 && j > i)
        cnt[i + 1]++;
    }
  for (int i = 0; i < n; i++)
    cnt[i + 1] += cnt[i];
  int total = cnt[n];
  int *idx = (int *)xmalloc(sizeof(int) * (size_t)total);
  int *fill = (int *)xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++)
    fill[i] = cnt[i];
  /* rows visited in ascending order => every list comes out ascending */
  for (int i = 0; i < n; i++)
    for (int p = rp[i]; p < rp[i + 1]; p++) {
      int j = ci[p];
      if (kind == 0 && i < j)
        idx[fill[j]++] = i;
      else if (kind == 1 && j < i)
        idx[fill[i]++] = j;
      else if (kind == 2 && j > i)
        idx[fill[i]++] = j;
    }
  free(fill);
  *dptr = cnt;
  *didx = idx;
  return total;
}

/* ------------------------------------------------- 007/008 levels + chains */

int orc_levels(int n, const int *dptr, const int *didx, int **start_levels,
               int **start_chains, int **chains, int *nchains) {
  /* Basic levels (007:4-60): repeated ascending scans that admit a column
   * once all its dependencies sit in strictly earlier rounds.  That is the
   * longest-path level; columns of one level keep ascending order. */
  int *level = (int *)xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *indeg = (int *)xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *rcnt = (int *)calloc((size_t)n + 1, sizeof(int));
  for (int i = 0; i < n; i++) {
    level[i] = 0;
    indeg[i] = dptr[i + 1] - dptr[i];
    for (int p = dptr[i]; p < dptr[i + 1]; p++)
      rcnt[didx[p] + 1]++;
  }
  for (int i = 0; i < n; i++)
    rcnt[i + 1] += rcnt[i];
  int *radj = (int *)xmalloc(sizeof(int) * (size_t)(rcnt[n] ? rcnt[n] : 1));
  int *rfill = (int *)xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++)
    rfill[i] = rcnt[i];
  for (int i = 0; i < n; i++)
    for (int p = dptr[i]; p < dptr[i + 1]; p++)
      radj[rfill[didx[p]]++] = i;
  int *queue = (int *)xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int qh = 0, qt = 0;
  for (int i = 0; i < n; i++)
    if (indeg[i] == 0)
      queue[qt++] = i;
  while (qh < qt) {
    int i = queue[qh++];
    for (int p = rcnt[i]; p < rcnt[i + 1]; p++) {
      int d = radj[p];
      if (level[d] < level[i] + 1)
        level[d] = level[i] + 1;
      if (--indeg[d] == 0)
        queue[qt++] = d;
    }
  }
  free(indeg);
  free(rcnt);
  free(radj);
  free(rfill);
  free(queue);
  if (qt != n) {
    free(level);
    return -1;
  }
  int nbasic = 0;
  for (int i = 0; i < n; i++)
    if (level[i] + 1 > nbasic)
      nbasic = level[i] + 1;
  int *bstart = (int *)calloc((size_t)nbasic + 1, sizeof(int));
  for (int i = 0; i < n; i++)
    bstart[level[i] + 1]++;
  for (int l = 0; l < nbasic; l++)
    bstart[l + 1] += bstart[l];
  int *bcols = (int *)xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *bfill = (int *)xmalloc(sizeof(int) * (size_t)(nbasic + 1));
  for (int l = 0; l < nbasic; l++)
    bfill[l] = bstart[l];
  for (int i = 0; i < n; i++)
    bcols[bfill[level[i]]++] = i;
  free(bfill);
  free(level);

  /* Fusion (007:89-144,146-210).  Chains are kept as linked lists of
   * columns: chain_head/chain_tail per chain, next_in_chain per column. */
  int *pos = (int *)xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *next_in_chain = (int *)xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *chain_head = (int *)xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *chain_tail = (int *)xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *level_first_chain = (int *)xmalloc(sizeof(int) * (size_t)(nbasic + 1));
  int *hook = (int *)xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *hook_used = (int *)xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++) {
    pos[i] = -1;
    next_in_chain[i] = -1;
    hook_used[i] = -1;
  }
  int nlev = 0, nch = 0, cur_first = 0;
  for (int bl = 0; bl < nbasic; bl++) {
    int b0 = bstart[bl], b1 = bstart[bl + 1];
    int can_fuse = bl > 0;
    for (int t = b0; can_fuse && t < b1; t++) {
      int col = bcols[t];
      /* chain_column_belongs_to (007:62-87): all positioned dependencies
       * must name one chain */
      int found = -1, ok = 1;
      for (int p = dptr[col]; p < dptr[col + 1]; p++) {
        int pp = pos[didx[p]];
        if (pp == -1 || pp == found)
          continue;
        if (found != -1) {
          ok = 0;
          break;
        }
        found = pp;
      }
      if (!ok || found == -1) {
        can_fuse = 0;
        break;
      }
      /* no two columns of this basic level may hook the same chain */
      if (hook_used[found] == bl) {
        can_fuse = 0;
        break;
      }
      hook_used[found] = bl;
      hook[t - b0] = found;
    }
    if (can_fuse) {
      for (int t = b0; t < b1; t++) {
        int col = bcols[t];
        int ch = cur_first + hook[t - b0];
        next_in_chain[chain_tail[ch]] = col;
        chain_tail[ch] = col;
        pos[col] = hook[t - b0];
      }
    } else {
      for (int i = 0; i < n; i++)
        pos[i] = -1;
      cur_first = nch;
      level_first_chain[nlev++] = nch;
      for (int t = b0; t < b1; t++) {
        int col = bcols[t];
        chain_head[nch] = chain_tail[nch] = col;
        pos[col] = t - b0;
        nch++;
      }
    }
  }
  level_first_chain[nlev] = nch;
  int *sl = (int *)xmalloc(sizeof(int) * (size_t)(nlev + 1));
  int *sc = (int *)xmalloc(sizeof(int) * (size_t)(nch + 1));
  int *cc = (int *)xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int w = 0;
  for (int l = 0; l < nlev; l++)
    sl[l] = level_first_chain[l];
  sl[nlev] = nch;
  for (int c = 0; c < nch; c++) {
    sc[c] = w;
    for (int col = chain_head[c]; col != -1; col = next_in_chain[col])
      cc[w++] = col;
  }
  sc[nch] = w;
  free(pos);
  free(next_in_chain);
  free(chain_head);
  free(chain_tail);
  free(level_first_chain);
  free(hook);
  free(hook_used);
  free(bstart);
  free(bcols);
  *start_levels = sl;
  *start_chains = sc;
  *chains = cc;
  *nchains = nch;
  return nlev;
}

void orc_max_per_level(int nlev, const int *start_levels,
                       const int *start_chains, const int *chains,
                       const int *rp, int *out) {
  for (int l = 0; l < nlev; l++) {
    int m = 0;
    for (int c = start_levels[l]; c < start_levels[l + 1]; c++)
      for (int t = start_chains[c]; t < start_chains[c + 1]; t++) {
        int col = chains[t];
        int len = rp[col + 1] - rp[col];
        if (len > m)
          m = len;
      }
    out[l] = m;
  }
}

/* -------------------------------------------------------------- 010 CSR/CSC */

void orc_csr_csc(int n, const int *rp, const int *ci, int *cp, int *ri,
                 int *csr2csc, int *csc2csr, int *csr_diag, int *csc_diag) {
  int nnz = rp[n];
  for (int j = 0; j <= n; j++)
    cp[j] = 0;
  for (int p = 0; p < nnz; p++)
    cp[ci[p] + 1]++;
  for (int j = 0; j < n; j++)
    cp[j + 1] += cp[j];
  int *fill = (int *)xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int j = 0; j < n; j++)
    fill[j] = cp[j];
  for (int i = 0; i < n; i++)
    for (int p = rp[i]; p < rp[i + 1]; p++) {
      int q = fill[ci[p]]++;
      ri[q] = i; /* rows visited ascending: columns come out sorted */
      csr2csc[p] = q;
      csc2csr[q] = p;
    }
  free(fill);
  for (int i = 0; i < n; i++) {
    int p = rp[i];
    while (p < rp[i + 1] && ci[p] <= i)
      p++;
    csr_diag[i] = p; /* first strictly right of the diagonal (010:88-106) */
  }
  for (int j = 0; j < n; j++) {
    int q = cp[j];
    while (q < cp[j + 1] && ri[q] <= j)
      q++;
    csc_diag[j] = q; /* first strictly below the diagonal (010:108-127) */
  }
}

/* ------------------------------------------------------- GLU CSC sequential */

int orc_glu_csc_sequential(int n, const int *rp, const int *ci,
                           const int *csr2csc, const int *cp, const int *ri,
                           double *vals, const int *csr_diag,
                           const int *csc_diag, int *bad_col) {
  for (int j = 0; j < n; j++) {
    if (cp[j] == csc_diag[j]) {
      if (bad_col)
        *bad_col = j;
      return -2;
    }
    double pivot = vals[csc_diag[j] - 1];
    if (fabs(pivot) < 1e-8) {
      if (bad_col)
        *bad_col = j;
      return -1;
    }
    for (int q = csc_diag[j]; q < cp[j + 1]; q++)
      vals[q] /= pivot;
    for (int p = csr_diag[j]; p < rp[j + 1]; p++) {
      int c = ci[p];
      double u = vals[csr2csc[p]];
      for (int q = csc_diag[j]; q < cp[j + 1]; q++) {
        int i = ri[q];
        double l = vals[q];
        for (int t = cp[c]; t < cp[c + 1]; t++) {
          if (ri[t] != i)
            continue;
          vals[t] -= l * u;
          break;
        }
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------ triangular solves */

void orc_substitute_lu_csc(int n, const int *cp, const int *ri,
                           const double *vals, const int *csc_diag,
                           double *b) {
  for (int j = 0; j < n; j++)
    for (int q = csc_diag[j]; q < cp[j + 1]; q++)
      b[ri[q]] -= vals[q] * b[j];
  for (int j = n - 1; j >= 0; j--) {
    int d = csc_diag[j] - 1;
    b[j] /= vals[d];
    for (int q = cp[j]; q < d; q++)
      b[ri[q]] -= vals[q] * b[j];
  }
}

void orc_solve_transpose_csc(int n, const int *cp, const int *ri,
                             const double *vals, const int *csc_diag,
                             double *g) {
  /* U^T z = g: z_j = (g_j - sum_{i<j} U(i,j) z_i) / U(j,j), j ascending */
  for (int j = 0; j < n; j++) {
    int d = csc_diag[j] - 1;
    double s = g[j];
    for (int q = cp[j]; q < d; q++)
      s -= vals[q] * g[ri[q]];
    g[j] = s / vals[d];
  }
  /* L^T y = z: y_j = z_j - sum_{i>j} L(i,j) y_i, j descending */
  for (int j = n - 1; j >= 0; j--) {
    double s = g[j];
    for (int q = csc_diag[j]; q < cp[j + 1]; q++)
      s -= vals[q] * g[ri[q]];
    g[j] = s;
  }
}

/* --------------------------------------------------------- 011 Newton step */

int orc_newton_update(int ne, double *v_old, const double *v_new,
                      long double *diff, double *max_abs_delta) {
  long double acc = 0;
  double mx = 0;
  for (int e = 0; e < ne; e++) {
    double old = v_old[e], nw = v_new[e], m = 1.0, cand = 0;
    for (;;) {
      if (m < 1e-13)
        return -1; /* 011:58-62 */
      cand = old * (1.0 - m) + nw * m;
      /* is_value_change_justified (011:6-17) */
      int ok = !isnan(old) && !isnan(cand) && !(fabs(old - cand) > 1.0);
      if (ok)
        break;
      m /= 2.0;
    }
    double d = fabs(cand - old);
    acc += (long double)d;
    if (d > mx)
      mx = d;
    v_old[e] = cand;
  }
  if (diff)
    *diff = acc;
  if (max_abs_delta)
    *max_abs_delta = mx;
  return 0;
}

/* ------------------------------------------------------------ permutations */

void orc_apply_permutation(int n, double *vec, const int *perm) {
  double *out = (double *)xmalloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++)
    out[perm[i]] = vec[i];
  memcpy(vec, out, sizeof(double) * (size_t)n);
  free(out);
}

void orc_invert_permutation(int n, const int *perm, int *inv) {
  for (int i = 0; i < n; i++)
    inv[perm[i]] = i;
}

typedef struct {
  int c;
  double v;
} cv_pair;

static int cmp_cv(const void *a, const void *b) {
  int x = ((const cv_pair *)a)->c, y = ((const cv_pair *)b)->c;
  return (x > y) - (x < y);
}

static void build_rows(int n, const int *rp, const int *src_row_of_new,
                       const int *ci, const double *v, const int *colmap,
                       int *orp, int *oci, double *ov) {
  orp[0] = 0;
  for (int k = 0; k < n; k++) {
    int i = src_row_of_new[k];
    orp[k + 1] = orp[k] + (rp[i + 1] - rp[i]);
  }
  for (int k = 0; k < n; k++) {
    int i = src_row_of_new[k];
    int len = rp[i + 1] - rp[i];
    cv_pair *tmp = (cv_pair *)xmalloc(sizeof(cv_pair) * (size_t)(len ? len : 1));
    for (int t = 0; t < len; t++) {
      int j = ci[rp[i] + t];
      tmp[t].c = colmap ? colmap[j] : j;
      tmp[t].v = v ? v[rp[i] + t] : 0.0;
    }
    qsort(tmp, (size_t)len, sizeof(cv_pair), cmp_cv);
    for (int t = 0; t < len; t++) {
      oci[orp[k] + t] = tmp[t].c;
      if (ov)
        ov[orp[k] + t] = tmp[t].v;
    }
    free(tmp);
  }
}

void orc_permute_sym(int n, const int *rp, const int *ci, const double *v,
                     const int *P, int *orp, int *oci, double *ov) {
  int *Pinv = (int *)xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  orc_invert_permutation(n, P, Pinv);
  /* new row k holds old row P[k]; columns mapped through Pinv */
  build_rows(n, rp, P, ci, v, Pinv, orp, oci, ov);
  free(Pinv);
}

void orc_permute_rows(int n, const int *rp, const int *ci, const double *v,
                      const int *p, int *orp, int *oci, double *ov) {
  /* new[p[i]] = old[i] => new row k holds old row pinv[k] */
  int *pinv = (int *)xmalloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  orc_invert_permutation(n, p, pinv);
  build_rows(n, rp, pinv, ci, v, NULL, orp, oci, ov);
  free(pinv);
}

/* ------------------------------------------------ 004b remove zero diagonal */

static int row_has(const ivec *row, int c) {
  for (int t = 0; t < row->len; t++)
    if (row->v[t] == c)
      return 1;
  return 0;
}

int orc_remove_zeros_diagonal(int n, int *rp, int *ci, double *v, int *perm) {
  /* rows[] hold column lists (with values in rvals); cols[] hold row lists.
   * A row swap moves whole rows; the column lists are relabelled. */
  ivec *rows = (ivec *)calloc((size_t)(n > 0 ? n : 1), sizeof(ivec));
  double **rvals = (double **)calloc((size_t)(n > 0 ? n : 1), sizeof(double *));
  ivec *cols = (ivec *)calloc((size_t)(n > 0 ? n : 1), sizeof(ivec));
  for (int i = 0; i < n; i++) {
    int len = rp[i + 1] - rp[i];
    rvals[i] = (double *)xmalloc(sizeof(double) * (size_t)(len ? len : 1));
    for (int p = rp[i]; p < rp[i + 1]; p++) {
      ivec_push(&rows[i], ci[p]);
      rvals[i][p - rp[i]] = v ? v[p] : 0.0;
      ivec_push(&cols[ci[p]], i);
    }
  }
  int status = 0;
  for (int r = 0; r < n; r++) {
    if (row_has(&rows[r], r))
      continue;
    /* would_fill (004b:31-53): some c<r with (r,c) and (c,r) present */
    int would_fill = 0;
    for (int t = 0; t < rows[r].len && !would_fill; t++) {
      int c = rows[r].v[t];
      if (c >= r)
        break;
      if (row_has(&cols[r], c))
        would_fill = 1;
    }
    if (would_fill)
      continue;
    int partner = -1; /* first row below r with an entry in column r */
    for (int t = 0; t < cols[r].len; t++)
      if (cols[r].v[t] > r) {
        partner = cols[r].v[t];
        break;
      }
    if (partner == -1) {
      status = -1;
      break;
    }
    /* relabel r <-> partner inside every column list touched */
    for (int s = 0; s < 2; s++) {
      int rr = s ? partner : r;
      for (int t = 0; t < rows[rr].len; t++) {
        ivec *col = &cols[rows[rr].v[t]];
        for (int u = 0; u < col->len; u++) {
          if (col->v[u] == r)
            col->v[u] = -2; /* temporary mark */
          else if (col->v[u] == partner)
            col->v[u] = -3;
        }
      }
    }
    for (int s = 0; s < 2; s++) {
      int rr = s ? partner : r;
      for (int t = 0; t < rows[rr].len; t++) {
        ivec *col = &cols[rows[rr].v[t]];
        int changed = 0;
        for (int u = 0; u < col->len; u++) {
          if (col->v[u] == -2) {
            col->v[u] = partner;
            changed = 1;
          } else if (col->v[u] == -3) {
            col->v[u] = r;
            changed = 1;
          }
        }
        if (changed)
          qsort(col->v, (size_t)col->len, sizeof(int), cmp_int);
      }
    }
    ivec tr = rows[r];
    rows[r] = rows[partner];
    rows[partner] = tr;
    double *tv = rvals[r];
    rvals[r] = rvals[partner];
    rvals[partner] = tv;
    int tp = perm[r];
    perm[r] = perm[partner];
    perm[partner] = tp;
  }
  if (status == 0) {
    int w = 0;
    rp[0] = 0;
    for (int i = 0; i < n; i++) {
      for (int t = 0; t < rows[i].len; t++) {
        ci[w] = rows[i].v[t];
        if (v)
          v[w] = rvals[i][t];
        w++;
      }
      rp[i + 1] = w;
    }
  }
  for (int i = 0; i < n; i++) {
    free(rows[i].v);
    free(cols[i].v);
    free(rvals[i]);
  }
  free(rows);
  free(cols);
  free(rvals);
  return status;
}

/* -------------------------------------------------- single-diode module */

void orc_module_current(double V, double Iph, double I0, double a, double Rs,
                        double Rsh, double *I, double *dIdV, double *dIdIph) {
  /* g(I) = Iph - I0 (exp((V + I Rs)/a) - 1) - (V + I Rs)/Rsh - I is strictly
   * decreasing in I; bracket the root, then Newton steps kept inside the
   * bracket (bisection when a step leaves it). */
#define ORC_G(x, out)                                                          \
  do {                                                                         \
    double u_ = (V + (x)*Rs) / a;                                              \
    double e_ = u_ > 700.0 ? INFINITY : exp(u_);                               \
    (out) = Iph - I0 * (e_ - 1.0) - (V + (x)*Rs) / Rsh - (x);                  \
  } while (0)
  double lo = -1.0, hi = 1.0, glo, ghi;
  ORC_G(lo, glo);
  while (glo <= 0) {
    lo *= 2;
    ORC_G(lo, glo);
  }
  ORC_G(hi, ghi);
  while (ghi >= 0) {
    hi *= 2;
    ORC_G(hi, ghi);
  }
  double x = 0.5 * (lo + hi);
  for (int it = 0; it < 400; it++) {
    double gx;
    ORC_G(x, gx);
    if (gx == 0)
      break;
    if (gx > 0)
      lo = x;
    else
      hi = x;
    double u = (V + x * Rs) / a;
    double e = exp(u);
    double dg = -I0 * Rs / a * e - Rs / Rsh - 1.0;
    double xn = x - gx / dg;
    if (!(xn > lo && xn < hi))
      xn = 0.5 * (lo + hi);
    if (xn == x || hi - lo <= 4e-16 * fabs(x))
      break;
    x = xn;
  }
#undef ORC_G
  double e = exp((V + x * Rs) / a);
  double dgdI = -I0 * Rs / a * e - Rs / Rsh - 1.0;
  double dgdV = -I0 / a * e - 1.0 / Rsh;
  *I = x;
  if (dIdV)
    *dIdV = -dgdV / dgdI;
  if (dIdIph)
    *dIdIph = -1.0 / dgdI;
}

void orc_module_current_many(int m, const double *V, const double *Iph,
                             double I0, double a, double Rs, const double *Rsh,
                             double *I, double *dIdV, double *dIdIph) {
  for (int e = 0; e < m; e++)
    orc_module_current(V[e], Iph[e], I0, a, Rs, Rsh[e], &I[e],
                       dIdV ? &dIdV[e] : NULL, dIdIph ? &dIdIph[e] : NULL);
}

void orc_ref_diodes(int m, const int *kind, const double *V, double T,
                    double *f, double *df) {
  const double q = 1.60217662E-19, k = 1.38064852E-23, na = 48.548;
  const double I0 = 4.86E-10 * (400.0 / 175.0);
  const double qkT = q / k / (273.15 + T);
  for (int e = 0; e < m; e++) {
    if (kind[e] == 0) {
      f[e] = I0 * (exp(qkT * V[e] / na) - 1.0);
      df[e] = I0 * qkT / na * exp(V[e] * qkT / na);
    } else {
      f[e] = 0.3E-6 * (exp(qkT * V[e]) - 1);
      df[e] = qkT * 0.3E-6 * (exp(qkT * V[e]));
    }
  }
}
