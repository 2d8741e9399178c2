// pvglu_pv.h -- the config PV plant object shared by the host drop-in driver
// and the batched device Newton.
#pragma once

#include "pvglu_circuit.h"
#include "pvglu_internal.h"

#include <cstdint>
#include <vector>

namespace pvglu {
struct NewtonDev; // device state of the batched Newton (pvglu_newton.hip)
struct DropinDev; // device buffers of the drop-in driver
void newton_dev_destroy(NewtonDev *nd);
void dropin_dev_destroy(DropinDev *dd);

// std::mt19937_64-based uniform in [0,1) with a fixed 53-bit construction so
// draws do not depend on the standard library's distribution code.
double uniform01(uint64_t word);

struct CloudDev; // device buffers of the cloud generator (pvglu_clouds.hip)
void cloud_dev_destroy(CloudDev *cd);
constexpr int kCloudParams = 48;
} // namespace pvglu

struct pvglu_pv {
  int I = 0, S = 0, M = 0;
  uint64_t seed = 0;
  int n = 0, nmod = 0, ncab = 0;
  pvglu::Circuit circuit;
  // 0-based node indices (-1 = ground)
  std::vector<int> mod_pos, mod_neg, cab_a, cab_b, bus;
  std::vector<double> mod_rsh, mod_iph, cab_g, pos_x, pos_y;

  // symbolic state, built once (prepare_matrices, PV/000_PVPP_simulator.cpp:10-93)
  bool prepared = false;
  std::vector<int> P, Pinv, Prow;
  pvglu_plan *plan = nullptr;
  // linear part of the MNA system (cables, Norton buses), unpermuted
  pvglu::SparseMap lin_A;
  std::vector<double> lin_b;

  int form = PVGLU_NEWTON_CORRECTION; // Newton step form of the drivers
  pvglu::NewtonDev *nd = nullptr;
  pvglu::DropinDev *dd = nullptr;
  pvglu::CloudDev *cd = nullptr;
};

namespace pvglu {
int pv_prepare(pvglu_pv *pv); // symbolic setup (host reorder + GPU fill + plan)
void cloud_params(const pvglu_pv *pv, int64_t scenario, double *prm);
}
