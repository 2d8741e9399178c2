// pvglu_device.hip -- host half of the numeric C-ABI: device plans, schedule
// selection and launching (the kernels are in pvglu_kernels.h, one
// translation unit per instance width), the 011 Newton update kernel, the
// autotuner and the multi-pattern work lists.
//
// Numeric path of the reference: PV/factorize_electrical_system_CUDA.cu:61-152,
// PV/factorize_matrix_GLU_CUDA.cuh:51-211, PV/triangular_substitution.cu:61-271.
#include <hip/hip_runtime.h>

#include "pvglu_internal.h"
#include "pvglu_launch.h"
#include "pvglu_update.h"

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
  // instance activity flags of the launch being issued (masked launches
  // only; null otherwise)
  const int *act = nullptr;
  // stamping on load of the launch being issued (stamped launches only)
  const int *stamp_seg = nullptr;
  const double *stamp_geq = nullptr, *stamp_gvs = nullptr,
               *stamp_ieq = nullptr, *stamp_xk = nullptr;
  const int *stamp_live = nullptr;
  std::vector<void *> allocs;
  // component groups (first component of each group, 1 / 2 / 4 groups) and
  // the side streams and events of the grouped schedules
  std::vector<int> comp_bounds[3];
  bool side_ok = false;
  hipStream_t side[3] = {};
  hipEvent_t ev_fork = nullptr, ev_skew[3] = {}, ev_join[3] = {};
};

} // namespace pvglu


static Pat pat_of(const DevicePlan *d) {
  return Pat{d->cp,       d->ri,        d->op_ptr,    d->op_l,     d->op_u,
             d->diag_slot, d->lrow_ptr, d->lrow_slot, d->lrow_col, d->urow_ptr,
             d->urow_slot, d->urow_col, d->lu_rec,  d->lu_seg,    d->lu_run,
             d->lu_fast,   d->u_rec,    d->u_seg,  d->u_run,     d->u_fast,
             d->l_rec,     d->l_seg,    d->l_run,  d->l_fast,    d->ut_rec,
             d->ut_seg,    d->ut_run,   d->ut_fast, d->lt_rec,   d->lt_seg,
             d->lt_run,    d->lt_fast,  d->drf,    d->status, d->act,
             d->stamp_seg, d->stamp_geq, d->stamp_gvs, d->stamp_ieq,
             d->stamp_xk,  d->stamp_live};
}
static Lev lev_of(const DevLevels &L) {
  return Lev{L.level_chain_ptr, L.chain_col_ptr, L.cols,
             L.comp_ptr,        L.comp_lv_begin, L.comp_lv_end,
             L.comp_lv_level};
}

__global__ void k_fill_int(int *p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = v;
}

// ----------------------------------------------------------- 011 update
// update_elements (damped_step per element, upd_tile_log2 launch shape) with
// nothing after each element.
__global__ void __launch_bounds__(256)
    k_newton_update(int ne, int B, const int *__restrict__ pos,
                    const int *__restrict__ neg, const double *__restrict__ x,
                    double *__restrict__ v, unsigned long long *delta_sum,
                    unsigned long long *delta_max, const int *active,
                    int *status) {
  update_elements<false>(
      ne, B, pos, neg, x, nullptr, v, delta_sum, delta_max, active, status,
      [](int, int) { return NoPre{}; }, [](int, int, double, double, NoPre) {},
      nullptr, nullptr);
}

// The fixed-point sums of k_newton_update converted to volts in place.
__global__ void k_fixed_to_delta(int B, double *delta_sum) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < B) {
    unsigned long long *u = (unsigned long long *)delta_sum;
    delta_sum[k] = fixed_to_delta(u[k]);
  }
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

void begin_call() {
  clear_error();
  (void)hipGetLastError();
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
  begin_call();
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
      (st = dev_upload(d, &d->drf, P->col_flags)) ||
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
                                  int &threads, bool fit = false) {
  variant = P->sched_variant;
  W = P->sched_width;
  // measured on MI355X: HYBRID wins from batch 16 up (40x24x28 plant:
  // batch 64 0.061 ms vs COMPONENT 0.110; 400x24x28, batch 32: 0.198 vs
  // 0.282 ms), COMPONENT below (batch 1: 0.050 vs 0.091 ms)
  if (!valid_width(W))
    W = B >= 64 ? 64 : (B >= 32 ? 32 : (B >= 16 ? 16 : (B >= 4 ? 4 : 1)));
  // fit: a width chosen for a larger batch narrowed to this one (a batch
  // that shrank, such as the batched Newton's after compaction)
  if (fit)
    while (W > 1 && W > B)
      W >>= 1;
  if (variant == PVGLU_SCHED_AUTO)
    variant = (B < 16 && P->ncomp >= 8) ? PVGLU_SCHED_COMPONENT
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
                             hipStream_t s, bool stream) {
  const int tiles = (B + W - 1) / W;
  const long threads = (long)nch * tiles * W;
  const int blocks = (int)((threads + 255) / 256);
  if (blocks == 0)
    return;
  Launch<W>::level(pat_of(P->dev), lev_of(L), op, first, nch, tiles, A, LU,
                   b, x, B, s, stream);
}

// Streaming cache policy (non-temporal hints on the value streams, widths 32
// and 64): on request, or by default when the batch's algorithmic bytes
// exceed twice the 256 MB Infinity Cache (512 x 40x24x28: 0.286 -> 0.260 ms;
// 32 x 400x24x28 at width 32: 0.197 -> 0.184 ms; on a single instance or the
// 128-instance mixed batch the hints cost time).
static bool use_streaming(const pvglu_plan *P, int W, int B) {
  if (W < 32 || P->cache_policy == PVGLU_CACHE_DEFAULT)
    return false;
  if (P->cache_policy == PVGLU_CACHE_STREAMING)
    return true;
  const double bytes = (double)B * (24.0 * P->nnz + 16.0 * P->n);
  return bytes > 512e6;
}

// Chains per component at or below which a level counts as narrow.
constexpr int kNarrowChains = 4;

// The launches of one numeric call (mode 0 refactor, 1 refactor + solve,
// 2 solve, 3 transposed solve) under LEVEL (hybrid = false) or HYBRID: a
// level launch of level lv of level set `set` (0 LU, 1 L, 2 U, 3 L^T), or a
// narrow-run launch over the level ranges R (HYBRID merges each run of two
// or more consecutive narrow steps; `widest` = most chains per component in
// it).
struct Item {
  bool run;
  int set, lv, op, widest;
  LvRange R;
};

static void mode_ops(const pvglu_plan *P, int mode, int ops[4]) {
  for (int k = 0; k < 4; ++k)
    ops[k] = -1;
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
}

static std::vector<Item> build_items(const pvglu_plan *P, int mode,
                                     bool hybrid) {
  const DevLevels *sets[4] = {&P->dev->lu, &P->dev->l, &P->dev->u,
                              &P->dev->lt};
  int ops[4];
  mode_ops(P, mode, ops);
  struct Step {
    int set, lv, width;
  };
  std::vector<Step> steps;
  for (int k = 0; k < 4; ++k)
    if (ops[k] >= 0)
      for (int lv = 0; lv < sets[k]->nlev; ++lv)
        steps.push_back({k, lv, sets[k]->h_level_comp_chains[lv]});
  std::vector<Item> items;
  const int nsteps = (int)steps.size();
  for (int i = 0; i < nsteps;) {
    int j = i, widest = 1;
    if (hybrid)
      while (j < nsteps && steps[j].width <= kNarrowChains) {
        widest = std::max(widest, steps[j].width);
        ++j;
      }
    Item it{};
    if (j - i < 2) {
      it.run = false;
      it.set = steps[i].set;
      it.lv = steps[i].lv;
      it.op = ops[it.set];
      items.push_back(it);
      ++i;
      continue;
    }
    it.run = true;
    it.widest = widest;
    for (int k = 0; k < 4; ++k) {
      it.R.lo[k] = 0;
      it.R.hi[k] = 0;
      it.R.op[k] = ops[k] < 0 ? 0 : ops[k];
    }
    for (int t = i; t < j; ++t) {
      const Step &sp = steps[t];
      if (it.R.lo[sp.set] == it.R.hi[sp.set])
        it.R.lo[sp.set] = sp.lv;
      it.R.hi[sp.set] = sp.lv + 1;
    }
    items.push_back(it);
    i = j;
  }
  return items;
}

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
  if (G > 1) {
    // under stream capture the groups are not forked: they gain nothing on
    // MI355X (eager or captured, measured) and a 4-stream fork/join during
    // capture has crashed the host runtime
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone)
      G = 1;
  }
  const int gi = group_index(G);
  G = 1 << gi;
  const DevLevels *sets[4] = {&d->lu, &d->l, &d->u, &d->lt};
  const std::vector<Item> items = build_items(P, mode, hybrid);
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
    const bool stream = use_streaming(P, W, B);
    for (const Item &it : items) {
      if (!it.run) {
        const DevLevels &L = *sets[it.set];
        const int *r = &L.h_grp[gi][((size_t)g * L.nlev + it.lv) * 2];
        // a stamped call forms the string segments' A values on load in its
        // LU level launches (stamped_lu_levels lists them)
        const int op = (d->stamp_seg && it.op == OP_REFACTOR_FUSED)
                           ? (int)OP_REFACTOR_STAMPED
                           : it.op;
        launch_one_level<W>(P, L, r[0], r[1] - r[0], op, A, LU, b, x, B, st,
                            stream);
      } else {
        int slots = 1;
        while (slots < it.widest)
          slots *= 2;
        const int threads = std::max(64, slots * W);
        Launch<W>::hybrid(pat, lev_of(d->lu), lev_of(d->l), lev_of(d->u),
                          lev_of(d->lt), it.R, c0, c1 - c0, tiles, threads, A,
                          LU, b, x, B, st, stream);
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
  Launch<W>::component(pat_of(d), lev_of(d->lu), lev_of(d->l), lev_of(d->u),
                       lev_of(d->lt), mode, P->fuse_lsolve, P->ncomp, tiles,
                       threads, A, LU, b, x, B, s);
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

// mode: 0 refactor, 1 refactor+solve, 2 solve, 3 transposed solve.  act:
// optional device flags per instance; tiles without an active instance are
// skipped (their factors and solution keep their earlier values).
static int run_numeric(pvglu_plan *P, int mode, const double *A, double *LU,
                       const double *b, double *x, int B, void *stream,
                       const int *act = nullptr, bool fit = false,
                       const StampOnLoad *so = nullptr) {
  begin_call();
  int st = require_device(P);
  if (st)
    return st;
  struct MaskScope { // the flags and stamping apply to this call only
    DevicePlan *d;
    MaskScope(DevicePlan *d, const int *a, const StampOnLoad *so) : d(d) {
      d->act = a;
      if (so) {
        d->stamp_seg = so->seg;
        d->stamp_geq = so->geq;
        d->stamp_gvs = so->gvs;
        d->stamp_ieq = so->ieq;
        d->stamp_xk = so->xk;
        d->stamp_live = so->live;
      }
    }
    ~MaskScope() {
      d->act = nullptr;
      d->stamp_seg = nullptr;
      d->stamp_geq = d->stamp_gvs = d->stamp_ieq = d->stamp_xk = nullptr;
      d->stamp_live = nullptr;
    }
  } scope(P->dev, act, so);
  if (B <= 0)
    return fail(PVGLU_E_ARG, "batch must be positive");
  if ((mode != 2 && mode != 3 && (!A || !LU)) ||
      ((mode == 2 || mode == 3) && !LU) || (mode >= 1 && (!b || !x)))
    return fail(PVGLU_E_ARG, "null device pointer");
  hipStream_t s = (hipStream_t)stream;
  int variant, W, threads;
  pick_default_schedule(P, B, variant, W, threads, fit);
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

namespace pvglu {
int refactor_solve_masked(pvglu_plan *P, const double *a, double *lu,
                          const double *b, double *x, int batch, void *stream,
                          const int *active) {
  return run_numeric(P, 1, a, lu, b, x, batch, stream, active, true);
}

int refactor_solve_stamped(pvglu_plan *P, const StampOnLoad &so,
                           const double *a, double *lu, const double *b,
                           double *x, int batch, void *stream,
                           const int *active) {
  if (!so.seg || !so.geq || !so.gvs)
    return fail(PVGLU_E_ARG, "stamped refactor: null stamping arrays");
  return run_numeric(P, 1, a, lu, b, x, batch, stream, active, true, &so);
}

int stamped_lu_levels(pvglu_plan *P, int batch, std::vector<int> &levels) {
  levels.clear();
  int st = require_device(P);
  if (st)
    return st;
  int variant, W, threads;
  pick_default_schedule(P, batch, variant, W, threads, true);
  if (variant == PVGLU_SCHED_COMPONENT || !P->fuse_lsolve)
    return PVGLU_OK;
  for (const Item &it : build_items(P, 1, variant == PVGLU_SCHED_HYBRID))
    if (!it.run && it.set == 0)
      levels.push_back(it.lv);
  return PVGLU_OK;
}
} // namespace pvglu

extern "C" int pvglu_solve_transpose(pvglu_plan *P, const double *lu,
                                     const double *g, double *y, int batch,
                                     void *stream) {
  return run_numeric(P, 3, nullptr, (double *)lu, g, y, batch, stream);
}

extern "C" int pvglu_status(pvglu_plan *P, int *code, int *column,
                            int *instance) {
  begin_call();
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
  begin_call();
  int st = require_device(P);
  if (st)
    return st;
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, (hipStream_t)stream,
                     P->dev->status, 4, 0);
  HIP_TRY(hipGetLastError());
  return PVGLU_OK;
}

extern "C" int pvglu_set_schedule(pvglu_plan *P, int variant, int width) {
  begin_call();
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
  begin_call();
  if (!P)
    return fail(PVGLU_E_ARG, "set_stream_groups: null plan");
  if (groups != 1 && groups != 2 && groups != 4)
    return fail(PVGLU_E_ARG, "set_stream_groups: %d not 1, 2 or 4", groups);
  P->sched_groups = groups;
  return PVGLU_OK;
}

extern "C" int pvglu_set_cache_policy(pvglu_plan *P, int policy) {
  begin_call();
  if (!P)
    return fail(PVGLU_E_ARG, "set_cache_policy: null plan");
  if (policy != PVGLU_CACHE_AUTO && policy != PVGLU_CACHE_STREAMING &&
      policy != PVGLU_CACHE_DEFAULT)
    return fail(PVGLU_E_ARG, "set_cache_policy: unknown policy %d", policy);
  P->cache_policy = policy;
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

int *pvglu::device_status(pvglu_plan *P) {
  return P->dev ? P->dev->status : nullptr;
}

extern "C" int pvglu_newton_update(pvglu_plan *P, int ne, int batch,
                                   const int *pos, const int *neg,
                                   const double *x, double *v_elem,
                                   double *delta_sum, double *delta_max,
                                   const int *active, void *stream) {
  begin_call();
  int st = require_device(P);
  if (st)
    return st;
  if (ne <= 0 || batch <= 0 || !pos || !neg || !x || !v_elem || !delta_sum ||
      !delta_max)
    return fail(PVGLU_E_ARG, "newton_update: bad arguments");
  const int lt = upd_tile_log2(batch);
  const int per_block = (256 >> lt) * kUpdPerThread;
  const int chunks = (ne + per_block - 1) / per_block;
  // zero the per-instance reductions (fill kernel: capture-safe)
  hipLaunchKernelGGL(k_fill_int, dim3((2 * batch + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, (int *)delta_sum, 2 * batch, 0);
  hipLaunchKernelGGL(k_fill_int, dim3((2 * batch + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, (int *)delta_max, 2 * batch, 0);
  hipLaunchKernelGGL(k_newton_update,
                     dim3(chunks, (batch + (1 << lt) - 1) >> lt), dim3(256), 0,
                     (hipStream_t)stream, ne, batch, pos, neg, x, v_elem,
                     (unsigned long long *)delta_sum,
                     (unsigned long long *)delta_max, active, P->dev->status);
  hipLaunchKernelGGL(k_fixed_to_delta, dim3((batch + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, batch, delta_sum);
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
  int variant, width, groups, policy;
};
static std::map<std::pair<uint64_t, int>, TuneChoice> g_tune_cache;

extern "C" int pvglu_autotune(pvglu_plan *P, int batch, int reps,
                              double *best_ms) {
  begin_call();
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
      P->cache_policy = it->second.policy;
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
            save_g = P->sched_groups, save_c = P->cache_policy;
  double best = 1e300;
  int bv = PVGLU_SCHED_LEVEL, bw = 1, bg = 1, bc = PVGLU_CACHE_DEFAULT;
  const int widths[] = {1, 2, 4, 8, 16, 32, 64};
  for (int variant = PVGLU_SCHED_LEVEL; variant <= PVGLU_SCHED_HYBRID && !st;
       ++variant) {
    for (int w : widths) {
      if (w > 1 && w / 2 >= batch)
        continue; // wider tiles only idle lanes
      for (int cfg = 0; cfg < 6; ++cfg) {
        // (groups, cache policy): streaming only where it can apply
        const int groups = 1 << (cfg % 3);
        const int policy = cfg < 3 ? PVGLU_CACHE_DEFAULT : PVGLU_CACHE_STREAMING;
        if (groups > 1 && (variant == PVGLU_SCHED_COMPONENT ||
                           !P->dev->side_ok || groups > P->ncomp))
          continue;
        if (policy == PVGLU_CACHE_STREAMING &&
            (w < 32 || variant == PVGLU_SCHED_COMPONENT))
          continue;
        P->sched_variant = variant;
        P->sched_width = w;
        P->sched_groups = groups;
        P->cache_policy = policy;
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
          bc = policy;
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
    P->cache_policy = save_c;
    return st;
  }
  // clear any status the stand-in values may have recorded
  int zero[4] = {0, 0, 0, 0};
  HIP_TRY(hipMemcpy(P->dev->status, zero, sizeof zero, hipMemcpyHostToDevice));
  P->sched_variant = bv;
  P->sched_width = bw;
  P->sched_groups = bg;
  P->cache_policy = bc;
  P->tuned_batch = batch;
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune_cache[{P->hash, batch}] = {bv, bw, bg, bc};
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
  // multi-pattern HYBRID steps, per mode (1 refactor + solve, 3 transposed
  // solve): usable when every group's step sequence has the same shape
  bool steps_ok[4] = {false, false, false, false};
  int step_width = 64;
  std::vector<std::vector<Item>> items[4];
};

extern "C" int pvglu_multi_create(int ngroups, pvglu_plan *const *plans,
                                  const int *batches, pvglu_multi **out) {
  begin_call();
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
  // HYBRID steps across the groups: one wide-level launch per wide level and
  // one narrow-run launch per run, each covering every group, when the step
  // sequences agree (same topology at different sizes) and the batches fill
  // at least 16 lanes; otherwise the component-style work list above
  m->step_width = minB >= 32 ? 64 : 16;
  for (int mode : {1, 3}) {
    bool ok = minB >= 16;
    for (int g = 0; g < ngroups && ok; g++) {
      m->items[mode].push_back(build_items(plans[g], mode, true));
      const auto &a = m->items[mode][0], &b = m->items[mode][g];
      ok = a.size() == b.size();
      for (size_t i = 0; ok && i < a.size(); i++)
        ok = a[i].run == b[i].run &&
             (a[i].run || (a[i].set == b[i].set && a[i].op == b[i].op));
    }
    m->steps_ok[mode] = ok;
  }
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
  Launch<W>::multi(args, mode, blocks, m->threads, s);
  HIP_TRY(hipGetLastError());
  return PVGLU_OK;
}

template <int W>
static int launch_multi_steps_w(pvglu_multi *m, int mode,
                                const double *const *a, double *const *lu,
                                const double *const *b, double *const *x,
                                hipStream_t s) {
  const int ng = m->ngroups;
  const size_t nitems = m->items[mode][0].size();
  for (size_t i = 0; i < nitems; i++) {
    if (!m->items[mode][0][i].run) {
      MultiLevel ml{};
      ml.ngroups = ng;
      int blocks = 0;
      for (int g = 0; g < ng; g++) {
        const pvglu_plan *P = m->plans[g];
        const DevicePlan *d = P->dev;
        const Item &it = m->items[mode][g][i];
        const DevLevels *sets[4] = {&d->lu, &d->l, &d->u, &d->lt};
        const DevLevels &L = *sets[it.set];
        const int first = L.h_level_chain_ptr[it.lv];
        const int nch = L.h_level_chain_ptr[it.lv + 1] - first;
        const int tiles = (m->batches[g] + W - 1) / W;
        ml.block_start[g] = blocks;
        blocks += (int)(((long)nch * tiles * W + 255) / 256);
        ml.first[g] = first;
        ml.nch[g] = nch;
        ml.tiles[g] = tiles;
        ml.B[g] = m->batches[g];
        ml.pat[g] = pat_of(d);
        ml.lev[g] = lev_of(L);
        ml.a[g] = a ? a[g] : nullptr;
        ml.luv[g] = lu[g];
        ml.b[g] = b[g];
        ml.x[g] = x[g];
      }
      ml.block_start[ng] = blocks;
      Launch<W>::level_multi(ml, m->items[mode][0][i].op, blocks, s);
    } else {
      MultiRun mr{};
      mr.ngroups = ng;
      int blocks = 0, threads = 64;
      for (int g = 0; g < ng; g++) {
        const pvglu_plan *P = m->plans[g];
        const DevicePlan *d = P->dev;
        const Item &it = m->items[mode][g][i];
        int slots = 1;
        while (slots < it.widest)
          slots *= 2;
        threads = std::max(threads, slots * W);
        const int tiles = (m->batches[g] + W - 1) / W;
        mr.block_start[g] = blocks;
        blocks += P->ncomp * tiles;
        mr.tiles[g] = tiles;
        mr.B[g] = m->batches[g];
        mr.R[g] = it.R;
        mr.pat[g] = pat_of(d);
        mr.lu[g] = lev_of(d->lu);
        mr.l[g] = lev_of(d->l);
        mr.u[g] = lev_of(d->u);
        mr.lt[g] = lev_of(d->lt);
        mr.a[g] = a ? a[g] : nullptr;
        mr.luv[g] = lu[g];
        mr.b[g] = b[g];
        mr.x[g] = x[g];
      }
      mr.block_start[ng] = blocks;
      Launch<W>::run_multi(mr, blocks, threads, s);
    }
  }
  HIP_TRY(hipGetLastError());
  return PVGLU_OK;
}

static int launch_multi(pvglu_multi *m, int mode, const double *const *a,
                        double *const *lu, const double *const *b,
                        double *const *x, void *stream) {
  begin_call();
  if (!m || !lu || !b || !x || (mode != 3 && !a))
    return fail(PVGLU_E_ARG, "multi: null pointer arrays");
  for (int g = 0; g < m->ngroups; g++)
    if (!lu[g] || !b[g] || !x[g] || (mode != 3 && !a[g]))
      return fail(PVGLU_E_ARG, "multi: null device pointer in group %d", g);
  hipStream_t s = (hipStream_t)stream;
  if (m->steps_ok[mode])
    return m->step_width == 64
               ? launch_multi_steps_w<64>(m, mode, a, lu, b, x, s)
               : launch_multi_steps_w<16>(m, mode, a, lu, b, x, s);
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
