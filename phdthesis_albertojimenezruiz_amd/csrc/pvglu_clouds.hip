// pvglu_clouds.hip -- irradiance of the config plants on the device.
//
// The analogue of the reference's cloud kernel (PV/001_clouds.cu:96-155) for
// the config cloud model: each (module, instance) sums the shadow of the
// instance's 8 drifting Gaussian blobs at the instance's time and writes
// G = 1000 * max(0, 1 - shade) W/m^2 into [module][instance], the layout the
// batched stamping and Newton read.  The blob parameters come from the same
// seeded host draws as pvglu_pv_irradiance, so the two generators differ only
// in the last bits of exp().
#include <hip/hip_runtime.h>

#include "pvglu_module.h"
#include "pvglu_pv.h"

#include <vector>

using namespace pvglu;

namespace pvglu {
struct CloudDev {
  int device = -1;
  double *pos = nullptr;   // x[nmod], y[nmod]
  double *prm = nullptr;   // [instance][kCloudParams + 1]: blobs, time, flag
  int cap = 0;             // instances prm can hold
  std::vector<double> host;
};

void cloud_dev_destroy(CloudDev *cd) {
  if (!cd)
    return;
  (void)hipFree(cd->pos);
  (void)hipFree(cd->prm);
  delete cd;
}
} // namespace pvglu

constexpr int kStride = kCloudParams + 2; // params, time, clear-sky flag

__global__ void __launch_bounds__(256)
    k_clouds(int nmod, int B, const double *__restrict__ pos,
             const double *__restrict__ prm, double *__restrict__ g) {
  PVGLU_ROW_INDEX(nmod, B);
  const int m = row_;
  const double *p = prm + (size_t)inst * kStride;
  if (p[kCloudParams + 1] != 0.0) { // uniform 1000 W/m^2
    g[gid] = 1000.0;
    return;
  }
  const double t = p[kCloudParams];
  const double x = pos[m], y = pos[nmod + m];
  double shade = 0.0;
  for (int k = 0; k < 8; k++) {
    const double dx = x - (p[k] + p[16 + k] * t);
    const double dy = y - (p[8 + k] + p[24 + k] * t);
    const double s = p[32 + k];
    shade += p[40 + k] * exp(-(dx * dx + dy * dy) / (2.0 * s * s));
  }
  g[gid] = 1000.0 * fmax(0.0, 1.0 - shade);
}

#define HIP_TRY(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail(PVGLU_E_HIP, "%s: %s", #expr, hipGetErrorString(e_));        \
  } while (0)

extern "C" int pvglu_pv_irradiance_device(pvglu_pv *pv, int batch,
                                          const int64_t *scenarios,
                                          const double *times, double *g_dev,
                                          void *stream) {
  begin_call();
  if (!pv || batch <= 0 || !scenarios || !times || !g_dev)
    return fail(PVGLU_E_ARG, "pv_irradiance_device: bad arguments");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(PVGLU_E_HIP, "pv_irradiance_device: no HIP device");
  hipStream_t s = (hipStream_t)stream;
  if (!pv->cd) {
    auto *cd = new CloudDev();
    pv->cd = cd;
    HIP_TRY(hipGetDevice(&cd->device));
    HIP_TRY(hipMalloc(&cd->pos, sizeof(double) * 2 * pv->nmod));
    std::vector<double> xy(pv->pos_x);
    xy.insert(xy.end(), pv->pos_y.begin(), pv->pos_y.end());
    HIP_TRY(hipMemcpy(cd->pos, xy.data(), sizeof(double) * xy.size(),
                      hipMemcpyHostToDevice));
  }
  CloudDev *cd = pv->cd;
  if (batch > cd->cap) {
    HIP_TRY(hipStreamSynchronize(s)); // earlier launches may read prm
    (void)hipFree(cd->prm);
    cd->prm = nullptr;
    cd->cap = 0;
    HIP_TRY(hipMalloc(&cd->prm, sizeof(double) * kStride * (size_t)batch));
    cd->cap = batch;
  }
  cd->host.assign((size_t)kStride * batch, 0.0);
  for (int k = 0; k < batch; k++) {
    double *p = &cd->host[(size_t)k * kStride];
    if (scenarios[k] < 0)
      p[kCloudParams + 1] = 1.0;
    else
      cloud_params(pv, scenarios[k], p);
    p[kCloudParams] = times[k];
  }
  HIP_TRY(hipMemcpyAsync(cd->prm, cd->host.data(),
                         sizeof(double) * cd->host.size(),
                         hipMemcpyHostToDevice, s));
  const RowGrid rg = row_grid(pv->nmod, batch);
  hipLaunchKernelGGL(k_clouds, rg.grid, rg.block, 0, s, pv->nmod, batch,
                     cd->pos, cd->prm, g_dev);
  HIP_TRY(hipGetLastError());
  // the staging vector is reused by the next call: finish the copy first
  HIP_TRY(hipStreamSynchronize(s));
  return PVGLU_OK;
}
