// pvglu_fill.hip -- fill-in discovery (005a) and prefix sum (005b) on gfx950.
//
// Same fixed point as obtain_fillin_matrix_CUDA
// (PV/005a_obtain_fillin_matrix_CUDA.cu:18-236): the no-pivot symbolic LU
// closure  (i,k) in L and (k,c) in U, k < min(i,c)  =>  (i,c).  Design:
//   * one wavefront per row owns a 32768-column bitmap window in LDS; the U
//     parts of the rows named by the row's L entries are OR-ed into it with
//     LDS atomics, the row's own entries are masked, and popcounts give the
//     new-entry count (count pass) -- or the union is enumerated in order with
//     a wavefront prefix sum and written as the merged sorted row (write pass);
//   * counts -> exclusive scan -> exact new row offsets: no scratch growth,
//     no retry loop;
//   * rows whose L part does not touch a row that changed skip later passes
//     (the reference's active row/column masks, 005a:177-187).
// The prefix sum replaces call_cumsum_CUDA (PV/005b_cumsum_CUDA.cu:22-31)
// with a single-workgroup tiled scan: coalesced 16 B per lane, wavefront
// shuffles inside a tile, a running carry across tiles.
#include <hip/hip_runtime.h>

#include "pvglu_internal.h"

#include <mutex>
#include <vector>

using namespace pvglu;

#define HIP_TRY(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail(PVGLU_E_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                         \
  } while (0)

static constexpr int kWinBits = 32768;           // columns per LDS window
static constexpr int kWinWords = kWinBits / 32;  // 1024 words = 4 KB
static constexpr int kWordsPerLane = kWinWords / 64;

__device__ __forceinline__ int wave_sum(int v) {
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int wave_excl_scan(int v, int lane) {
  int incl = v;
  for (int off = 1; off < 64; off <<= 1) {
    int t = __shfl_up(incl, off, 64);
    if (lane >= off)
      incl += t;
  }
  return incl - v;
}

// mode 0: count new entries of row i -> cnt[i]
// mode 1: write the merged row i at new_rp[i] (rows with cnt == 0 copied)
// One wavefront (64 threads) per block, one row per block.
__global__ void __launch_bounds__(64)
    k_fill_rows(int mode, int n, const int *__restrict__ rp,
                const int *__restrict__ ci, const unsigned char *__restrict__ active,
                int *__restrict__ cnt, const int *__restrict__ new_rp,
                int *__restrict__ new_ci) {
  __shared__ unsigned int bits[kWinWords];
  const int i = blockIdx.x;
  const int lane = threadIdx.x;
  const int r0 = rp[i], r1 = rp[i + 1];
  const bool act = active[i] != 0;
  if (mode == 1 && (!act || cnt[i] == 0)) {
    const int o = new_rp[i];
    for (int p = r0 + lane; p < r1; p += 64)
      new_ci[o + (p - r0)] = ci[p];
    return;
  }
  if (mode == 0 && !act) {
    if (lane == 0)
      cnt[i] = 0;
    return;
  }
  // smallest L entry k: new columns are > k
  int kmin = n;
  for (int p = r0; p < r1; ++p) {
    int k = ci[p];
    if (k >= i)
      break;
    kmin = k;
    break;
  }
  if (kmin == n) { // no L part: nothing can fill
    if (mode == 0) {
      if (lane == 0)
        cnt[i] = 0;
    } else {
      const int o = new_rp[i];
      for (int p = r0 + lane; p < r1; p += 64)
        new_ci[o + (p - r0)] = ci[p];
    }
    return;
  }
  int lo = kmin + 1;
  if (mode == 1 && r1 > r0 && ci[r0] < lo)
    lo = ci[r0];
  lo = (lo / kWinBits) * kWinBits;
  int total_new = 0;
  int out = mode == 1 ? new_rp[i] : 0;
  for (int ws = lo; ws < n; ws += kWinBits) {
    const int we = min(n, ws + kWinBits);
    for (int w = lane; w < kWinWords; w += 64)
      bits[w] = 0u;
    __syncthreads();
    // OR in the U parts of rows k for every L entry (i,k)
    for (int p = r0; p < r1; ++p) {
      const int k = ci[p];
      if (k >= i)
        break;
      const int s0 = rp[k], s1 = rp[k + 1];
      for (int t = s0 + lane; t < s1; t += 64) {
        const int c = ci[t];
        if (c > k && c >= ws && c < we)
          atomicOr(&bits[(c - ws) >> 5], 1u << ((c - ws) & 31));
      }
    }
    __syncthreads();
    // own entries: masked (count) or counted in the union (write)
    int own = 0;
    for (int p = r0 + lane; p < r1; p += 64) {
      const int c = ci[p];
      if (c >= ws && c < we) {
        const unsigned m = 1u << ((c - ws) & 31);
        if (mode == 0)
          atomicAnd(&bits[(c - ws) >> 5], ~m);
        else {
          unsigned old = atomicOr(&bits[(c - ws) >> 5], m);
          own += (old & m) ? 0 : 1; // entries not already produced by a merge
        }
      }
    }
    (void)own;
    __syncthreads();
    int pc = 0;
    for (int w = 0; w < kWordsPerLane; ++w)
      pc += __popc(bits[lane * kWordsPerLane + w]);
    if (mode == 0) {
      total_new += wave_sum(pc);
    } else {
      const int base = wave_excl_scan(pc, lane);
      int o = out + base;
      for (int w = 0; w < kWordsPerLane; ++w) {
        unsigned word = bits[lane * kWordsPerLane + w];
        while (word) {
          const int b = __ffs(word) - 1;
          word &= word - 1;
          new_ci[o++] = ws + (lane * kWordsPerLane + w) * 32 + b;
        }
      }
      out += wave_sum(pc);
    }
    __syncthreads();
  }
  if (mode == 0 && lane == 0)
    cnt[i] = total_new;
}

// new_len[i] = old_len[i] + cnt[i], written at rp_next[i+1]; rp_next[0] = 0
__global__ void k_new_lengths(int n, const int *rp, const int *cnt,
                              int *rp_next) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0)
    rp_next[0] = 0;
  if (i < n)
    rp_next[i + 1] = rp[i + 1] - rp[i] + cnt[i];
}

// Rows to revisit: a row that changed, or whose L part names a changed row.
__global__ void k_next_active(int n, const int *rp, const int *ci,
                              const int *cnt, unsigned char *active) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  unsigned char a = cnt[i] > 0;
  for (int p = rp[i]; p < rp[i + 1] && !a; ++p) {
    const int k = ci[p];
    if (k >= i)
      break;
    a = cnt[k] > 0;
  }
  active[i] = a;
}

__global__ void k_fill_u8(unsigned char *p, int n, unsigned char v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = v;
}

// Single-workgroup tiled inclusive scan (1024 threads, 4 ints per thread).
__global__ void __launch_bounds__(1024) k_scan_inclusive(int *a, int n) {
  __shared__ int wsum[16];
  __shared__ int carry_s;
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  if (t == 0)
    carry_s = 0;
  __syncthreads();
  for (int base = 0; base < n; base += 4096) {
    int v[4];
    const int i0 = base + 4 * t;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      v[k] = (i0 + k < n) ? a[i0 + k] : 0;
#pragma unroll
    for (int k = 1; k < 4; ++k)
      v[k] += v[k - 1];
    const int tsum = v[3];
    int incl = tsum;
    for (int off = 1; off < 64; off <<= 1) {
      int y = __shfl_up(incl, off, 64);
      if (lane >= off)
        incl += y;
    }
    if (lane == 63)
      wsum[wid] = incl;
    __syncthreads();
    int woff = 0;
    for (int w = 0; w < wid; ++w)
      woff += wsum[w];
    const int carry = carry_s;
    const int off = carry + woff + incl - tsum;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (i0 + k < n)
        a[i0 + k] = v[k] + off;
    __syncthreads();
    if (t == 1023) {
      int tot = 0;
      for (int w = 0; w < 16; ++w)
        tot += wsum[w];
      carry_s = carry + tot;
    }
    __syncthreads();
  }
}

extern "C" int pvglu_inclusive_scan_device(int *d, int n, void *stream) {
  begin_call();
  if (!d || n < 0)
    return fail(PVGLU_E_ARG, "inclusive_scan: bad arguments");
  if (n == 0)
    return PVGLU_OK;
  hipLaunchKernelGGL(k_scan_inclusive, dim3(1), dim3(1024), 0,
                     (hipStream_t)stream, d, n);
  HIP_TRY(hipGetLastError());
  return PVGLU_OK;
}

extern "C" int pvglu_obtain_fillin(int n, const int *rowptr, const int *colidx,
                                   int **filled_rowptr, int **filled_colidx,
                                   int *filled_nnz, void *stream) {
  begin_call();
  if (n <= 0 || !rowptr || !colidx || !filled_rowptr || !filled_colidx)
    return fail(PVGLU_E_ARG, "obtain_fillin: bad arguments");
  for (int i = 0; i < n; i++) {
    if (rowptr[i + 1] < rowptr[i])
      return fail(PVGLU_E_ARG, "obtain_fillin: rowptr not monotone");
    for (int p = rowptr[i]; p < rowptr[i + 1]; p++)
      if (colidx[p] < 0 || colidx[p] >= n ||
          (p > rowptr[i] && colidx[p - 1] >= colidx[p]))
        return fail(PVGLU_E_ARG, "obtain_fillin: row %d not sorted/in range",
                    i);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(PVGLU_E_HIP, "obtain_fillin: no HIP device");
  hipStream_t s = (hipStream_t)stream;
  const int nnz0 = rowptr[n];
  int *d_rp = nullptr, *d_ci = nullptr, *d_cnt = nullptr, *d_rp2 = nullptr,
      *d_ci2 = nullptr;
  unsigned char *d_act = nullptr;
  int st = PVGLU_OK;
  auto cleanup = [&]() {
    (void)hipFree(d_rp);
    (void)hipFree(d_ci);
    (void)hipFree(d_cnt);
    (void)hipFree(d_rp2);
    (void)hipFree(d_ci2);
    (void)hipFree(d_act);
  };
#define TRY_C(expr)                                                            \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      st = fail(PVGLU_E_HIP, "%s: %s", #expr, hipGetErrorString(e_));          \
      cleanup();                                                               \
      return st;                                                               \
    }                                                                          \
  } while (0)
  TRY_C(hipMalloc(&d_rp, sizeof(int) * (n + 1)));
  TRY_C(hipMalloc(&d_ci, sizeof(int) * std::max(nnz0, 1)));
  TRY_C(hipMalloc(&d_cnt, sizeof(int) * n));
  TRY_C(hipMalloc(&d_rp2, sizeof(int) * (n + 1)));
  TRY_C(hipMalloc(&d_act, n));
  TRY_C(hipMemcpyAsync(d_rp, rowptr, sizeof(int) * (n + 1),
                       hipMemcpyHostToDevice, s));
  if (nnz0)
    TRY_C(hipMemcpyAsync(d_ci, colidx, sizeof(int) * nnz0,
                         hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_fill_u8, dim3((n + 255) / 256), dim3(256), 0, s, d_act,
                     n, (unsigned char)1);
  int nnz = nnz0;
  for (int pass = 0; pass < n + 1; ++pass) {
    hipLaunchKernelGGL(k_fill_rows, dim3(n), dim3(64), 0, s, 0, n, d_rp, d_ci,
                       d_act, d_cnt, nullptr, nullptr);
    hipLaunchKernelGGL(k_new_lengths, dim3((n + 256) / 256), dim3(256), 0, s,
                       n, d_rp, d_cnt, d_rp2);
    hipLaunchKernelGGL(k_scan_inclusive, dim3(1), dim3(1024), 0, s, d_rp2,
                       n + 1);
    TRY_C(hipGetLastError());
    int nnz_new = 0;
    TRY_C(hipMemcpyAsync(&nnz_new, d_rp2 + n, sizeof(int),
                         hipMemcpyDeviceToHost, s));
    TRY_C(hipStreamSynchronize(s));
    if (nnz_new == nnz)
      break; // closure reached
    if (nnz_new < nnz || nnz_new > (long)n * n) {
      cleanup();
      return fail(PVGLU_E_HIP, "obtain_fillin: inconsistent pass (%d -> %d)",
                  nnz, nnz_new);
    }
    TRY_C(hipMalloc(&d_ci2, sizeof(int) * nnz_new));
    hipLaunchKernelGGL(k_fill_rows, dim3(n), dim3(64), 0, s, 1, n, d_rp, d_ci,
                       d_act, d_cnt, d_rp2, d_ci2);
    hipLaunchKernelGGL(k_next_active, dim3((n + 255) / 256), dim3(256), 0, s,
                       n, d_rp2, d_ci2, d_cnt, d_act);
    TRY_C(hipGetLastError());
    TRY_C(hipStreamSynchronize(s));
    std::swap(d_rp, d_rp2);
    TRY_C(hipFree(d_ci));
    d_ci = d_ci2;
    d_ci2 = nullptr;
    nnz = nnz_new;
  }
  int *h_rp = (int *)malloc(sizeof(int) * (n + 1));
  int *h_ci = (int *)malloc(sizeof(int) * std::max(nnz, 1));
  if (!h_rp || !h_ci) {
    free(h_rp);
    free(h_ci);
    cleanup();
    return fail(PVGLU_E_ARG, "obtain_fillin: host allocation failed");
  }
  TRY_C(hipMemcpyAsync(h_rp, d_rp, sizeof(int) * (n + 1),
                       hipMemcpyDeviceToHost, s));
  TRY_C(hipMemcpyAsync(h_ci, d_ci, sizeof(int) * nnz, hipMemcpyDeviceToHost,
                       s));
  TRY_C(hipStreamSynchronize(s));
#undef TRY_C
  cleanup();
  *filled_rowptr = h_rp;
  *filled_colidx = h_ci;
  if (filled_nnz)
    *filled_nnz = nnz;
  return PVGLU_OK;
}

// ---------------------------------------------------------------------------
// pvglu_obtain_fillin_noalloc: the same passes with every device buffer in
// one file-scope array, for callers that cannot allocate device memory.
// Capacity: n <= kNaMaxN rows and kNaMaxNnz nonzeros of any intermediate or
// the final filled pattern; the array (about 21 MB) is held on every device
// the library runs on until the process ends.  Layout (int words): two row
// pointer arrays, the per-row counts, the active-row mask (bytes) and two
// column index arrays, swapped between passes.  Calls that overlap on one
// device share the array and race.
static constexpr int kNaMaxN = 1 << 19;
static constexpr int kNaMaxNnz = 1 << 21;
static constexpr size_t kNaRp = (size_t)kNaMaxN + 64; // n + 1, padded
static constexpr size_t kNaWords =
    2 * kNaRp + kNaMaxN + kNaMaxN / 4 + 2 * (size_t)kNaMaxNnz;
static __device__ __align__(256) int scratch[kNaWords];

// Device address of `scratch` per device, resolved once per device (with
// that device current) on the first call there.
static std::mutex g_na_mu;
static int *g_na_ptr[64];

static int na_scratch(int **out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64)
    return fail(PVGLU_E_HIP, "obtain_fillin_noalloc: device %d unsupported", dev);
  std::lock_guard<std::mutex> g(g_na_mu);
  if (!g_na_ptr[dev]) {
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_na_ptr[dev] = (int *)p;
  }
  *out = g_na_ptr[dev];
  return PVGLU_OK;
}

extern "C" int pvglu_obtain_fillin_noalloc(int n, const int *rowptr,
                                           const int *colidx,
                                           int *filled_rowptr,
                                           int *filled_colidx, int capacity,
                                           int *filled_nnz, void *stream) {
  begin_call();
  if (n <= 0 || !rowptr || !colidx || !filled_rowptr || !filled_colidx ||
      !filled_nnz || capacity < 0)
    return fail(PVGLU_E_ARG, "obtain_fillin_noalloc: bad arguments");
  if (n > kNaMaxN)
    return fail(PVGLU_E_ARG,
                "obtain_fillin_noalloc: n = %d exceeds the maximum of %d rows",
                n, kNaMaxN);
  if (rowptr[0] != 0 || rowptr[n] < 0 || rowptr[n] > kNaMaxNnz)
    return fail(PVGLU_E_ARG,
                "obtain_fillin_noalloc: %d nonzeros exceed the maximum of %d",
                rowptr[n], kNaMaxNnz);
  for (int i = 0; i < n; i++) {
    if (rowptr[i + 1] < rowptr[i])
      return fail(PVGLU_E_ARG, "obtain_fillin_noalloc: rowptr not monotone");
    for (int p = rowptr[i]; p < rowptr[i + 1]; p++)
      if (colidx[p] < 0 || colidx[p] >= n ||
          (p > rowptr[i] && colidx[p - 1] >= colidx[p]))
        return fail(PVGLU_E_ARG,
                    "obtain_fillin_noalloc: row %d not sorted/in range", i);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(PVGLU_E_HIP, "obtain_fillin_noalloc: no HIP device");
  int *base = nullptr;
  int st = na_scratch(&base);
  if (st)
    return st;
  hipStream_t s = (hipStream_t)stream;
  int *d_rp = base, *d_rp2 = base + kNaRp, *d_cnt = base + 2 * kNaRp;
  unsigned char *d_act = (unsigned char *)(d_cnt + kNaMaxN);
  int *d_ci = d_cnt + kNaMaxN + kNaMaxN / 4, *d_ci2 = d_ci + kNaMaxNnz;
  const int nnz0 = rowptr[n];
  HIP_TRY(hipMemcpyAsync(d_rp, rowptr, sizeof(int) * (n + 1),
                         hipMemcpyHostToDevice, s));
  if (nnz0)
    HIP_TRY(hipMemcpyAsync(d_ci, colidx, sizeof(int) * nnz0,
                           hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_fill_u8, dim3((n + 255) / 256), dim3(256), 0, s, d_act,
                     n, (unsigned char)1);
  int nnz = nnz0;
  for (int pass = 0; pass < n + 1; ++pass) {
    hipLaunchKernelGGL(k_fill_rows, dim3(n), dim3(64), 0, s, 0, n, d_rp, d_ci,
                       d_act, d_cnt, nullptr, nullptr);
    hipLaunchKernelGGL(k_new_lengths, dim3((n + 256) / 256), dim3(256), 0, s,
                       n, d_rp, d_cnt, d_rp2);
    hipLaunchKernelGGL(k_scan_inclusive, dim3(1), dim3(1024), 0, s, d_rp2,
                       n + 1);
    HIP_TRY(hipGetLastError());
    int nnz_new = 0;
    HIP_TRY(hipMemcpyAsync(&nnz_new, d_rp2 + n, sizeof(int),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (nnz_new == nnz)
      break; // closure reached
    if (nnz_new < nnz || nnz_new > (long)n * n)
      return fail(PVGLU_E_HIP,
                  "obtain_fillin_noalloc: inconsistent pass (%d -> %d)", nnz,
                  nnz_new);
    if (nnz_new > kNaMaxNnz)
      return fail(PVGLU_E_ARG,
                  "obtain_fillin_noalloc: the filled pattern reaches %d "
                  "nonzeros, beyond the maximum of %d",
                  nnz_new, kNaMaxNnz);
    hipLaunchKernelGGL(k_fill_rows, dim3(n), dim3(64), 0, s, 1, n, d_rp, d_ci,
                       d_act, d_cnt, d_rp2, d_ci2);
    hipLaunchKernelGGL(k_next_active, dim3((n + 255) / 256), dim3(256), 0, s,
                       n, d_rp2, d_ci2, d_cnt, d_act);
    HIP_TRY(hipGetLastError());
    std::swap(d_rp, d_rp2);
    std::swap(d_ci, d_ci2);
    nnz = nnz_new;
  }
  *filled_nnz = nnz;
  if (nnz > capacity) {
    HIP_TRY(hipStreamSynchronize(s));
    return fail(PVGLU_E_ARG,
                "obtain_fillin_noalloc: the filled pattern has %d nonzeros, "
                "the output holds %d",
                nnz, capacity);
  }
  HIP_TRY(hipMemcpyAsync(filled_rowptr, d_rp, sizeof(int) * (n + 1),
                         hipMemcpyDeviceToHost, s));
  if (nnz)
    HIP_TRY(hipMemcpyAsync(filled_colidx, d_ci, sizeof(int) * nnz,
                           hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return PVGLU_OK;
}
