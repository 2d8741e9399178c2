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
  int skip_frozen = 1; // batched Newton: skip tiles of frozen instances
  int fused_stamping = 1; // batched Newton: string A values stamped on load
  pvglu::NewtonDev *nd = nullptr;
  pvglu::DropinDev *dd = nullptr;
  pvglu::CloudDev *cd = nullptr;
};

// The reference's own plant (PV/003_create_electric_system.cpp:184-382):
// blocks of panels_x x panels_y panels (internal node, series and shunt
// resistance, diode, photocurrent, bypass diode per panel), each block's DC
// side tied to ground and to its inverter (DC voltage source + AC injection),
// the inverters' AC nodes in a ring through the point of common coupling
// (a 25 kV voltage source).
struct pvglu_refplant {
  int blocks = 0, px = 0, py = 0, interconnection = 0, bypass = 1;
  double volts_per_panel = 0, irradiance = 0, temperature = 0;
  int n = 0, npanels = 0;
  pvglu::Circuit circuit;
  int pcc_node = 0, pcc_current = -1; // PCC AC node and its current unknown
  std::vector<double> panel_g;        // irradiance per panel (W/m^2)
  bool prepared = false;
  std::vector<int> P, Pinv, Prow;
  pvglu_plan *plan = nullptr;
  int form = PVGLU_NEWTON_CORRECTION;
  pvglu::DropinDev *dd = nullptr;
};

namespace pvglu {
// prepare_matrices (PV/000_PVPP_simulator.cpp:10-93) for any circuit: the
// pattern of the system at virtual conductance 1.0, reorder_matrix (AMD +
// row swaps, Prow checked to be an involution), GPU fill-in, plan, upload.
int prepare_circuit(const Circuit &c, std::vector<int> &P,
                    std::vector<int> &Pinv, std::vector<int> &Prow,
                    pvglu_plan **plan);
// The reference's call sequence (PV/000_PVPP_simulator.cpp:144-235) for a
// prepared circuit, in either Newton form; inverters refreshed after each
// iteration (PV/011_calculate_new_iteration.cpp:88-109).  x_out: unknowns in
// circuit order.  stats[0..3]: iterations, host ms, device ms, total ms.
int dropin_newton(Circuit &c, pvglu_plan *plan, const std::vector<int> &P,
                  const std::vector<int> &Pinv, const std::vector<int> &Prow,
                  DropinDev *&dd, int form, double tol, int ladder,
                  int fixed_iters, double *x_out, double *stats);
int pv_prepare(pvglu_pv *pv); // symbolic setup (host reorder + GPU fill + plan)
void cloud_params(const pvglu_pv *pv, int64_t scenario, double *prm);
}
