// pvglu_pv.cpp -- the config PV plants on the reference's circuit API.
//
// Topology (per inverter): S strings of M single-diode modules in series;
// module k sits between node k-1 (node 0 = ground) and node k; node M is the
// string's combiner, tied by a cable (log-uniform 5..50 S) to the inverter's
// DC bus; the bus is a Norton source of 1e4 S to ground in parallel with
// 1e4 * 600 A.  Unknowns: n = I (S M + 1) node voltages.  Elements are
// registered like the reference's plant builder does
// (PV/003_create_electric_system.cpp:89-126) through add_linear_element /
// add_nonlinear_voltage_source (PV/002_base_electric_system.cpp:78-108,164-173).
// A module is one nonlinear element in explicit terminal form (Rs folded by
// the Lambert W solution), so each module adds exactly one node.
#include "pvglu_module.h"
#include "pvglu_pv.h"

#include <cmath>
#include <cstring>
#include <random>

using namespace pvglu;

namespace pvglu {

double uniform01(uint64_t w) { return (double)(w >> 11) * 0x1.0p-53; }

static double log_uniform(std::mt19937_64 &rng, double lo, double hi) {
  return std::exp(std::log(lo) + uniform01(rng()) * (std::log(hi) - std::log(lo)));
}

int prepare_circuit(const Circuit &c, std::vector<int> &P,
                    std::vector<int> &Pinv, std::vector<int> &Prow,
                    pvglu_plan **plan) {
  // prepare_matrices (PV/000_PVPP_simulator.cpp:10-93): pattern of the
  // system assembled with virtual conductance 1.0, reordered, fill-in found,
  // symbolic plan built.
  SparseMap A;
  std::vector<double> b;
  c.create_matrix_and_vector(1.0, A, b);
  std::vector<int> rp, ci;
  A.csr(rp, ci);
  const int n = c.unknowns();
  P.assign(n, 0);
  Prow.assign(n, 0);
  int st = pvglu_reorder(n, rp.data(), ci.data(), P.data(), Prow.data());
  if (st)
    return st;
  Pinv.assign(n, 0);
  for (int i = 0; i < n; i++)
    Pinv[P[i]] = i;
  // The reordered matrix of reorder_matrix: AMD symmetric permutation, then
  // the recorded row swaps (row k holds row Prow[k]).  The Newton loop later
  // applies exchange_rows(Prow) (new[Prow[i]] = old[i]); both agree only when
  // Prow is an involution (disjoint swaps).
  for (int i = 0; i < n; i++)
    if (Prow[Prow[i]] != i)
      return fail(PVGLU_E_ARG, "prepare: row permutation is not an involution");
  SparseMap R = A.interchange_row_columns(Pinv.data());
  SparseMap R2(n, n);
  for (int k = 0; k < n; k++)
    R2.rows[k] = R.rows[Prow[k]];
  R2.csr(rp, ci);
  int *frp = nullptr, *fci = nullptr, fnnz = 0;
  st = pvglu_obtain_fillin(n, rp.data(), ci.data(), &frp, &fci, &fnnz, nullptr);
  if (st)
    return st;
  st = pvglu_plan_create(n, frp, fci, plan);
  free(frp);
  free(fci);
  if (st)
    return st;
  return pvglu_plan_upload(*plan);
}

int pv_prepare(pvglu_pv *pv) {
  if (pv->prepared)
    return PVGLU_OK;
  int st = prepare_circuit(pv->circuit, pv->P, pv->Pinv, pv->Prow, &pv->plan);
  if (st)
    return st;
  pv->prepared = true;
  return PVGLU_OK;
}

} // namespace pvglu

extern "C" int pvglu_pv_create(int inverters, int strings, int modules,
                               uint64_t seed, pvglu_pv **out) {
  clear_error();
  if (!out || inverters <= 0 || strings <= 0 || modules <= 0)
    return fail(PVGLU_E_ARG, "pv_create: bad arguments");
  auto *pv = new pvglu_pv();
  pv->I = inverters;
  pv->S = strings;
  pv->M = modules;
  pv->seed = seed;
  std::mt19937_64 rng(seed);
  Circuit &C = pv->circuit;
  const int nm_total = inverters * strings * modules;
  pv->mod_iph.assign(nm_total, kIphPerIrradiance * 1000.0);
  for (int inv = 0; inv < inverters; inv++) {
    std::vector<int> combiners;
    for (int s = 0; s < strings; s++) {
      int prev = 0;
      for (int k = 1; k <= modules; k++) {
        const int node = C.create_node();
        const int e = (int)pv->mod_rsh.size();
        const double rsh = log_uniform(rng, 200.0, 600.0);
        pv->mod_rsh.push_back(rsh);
        pv->mod_pos.push_back(node - 1);
        pv->mod_neg.push_back(prev - 1);
        // element current from + (node) to - (prev) is -I(V)
        C.add_nonlinear_voltage_source(
            [pv, e](double V) {
              return -module_current(V, pv->mod_iph[e], pv->mod_rsh[e]).I;
            },
            [pv, e](double V) {
              return -module_current(V, pv->mod_iph[e], pv->mod_rsh[e]).dIdV;
            },
            0.0, node, prev);
        prev = node;
      }
      combiners.push_back(prev);
      pv->cab_g.push_back(log_uniform(rng, 5.0, 50.0));
    }
    const int bus = C.create_node();
    pv->bus.push_back(bus - 1);
    for (int s = 0; s < strings; s++) {
      const double g = pv->cab_g[inv * strings + s];
      C.add_linear_element(ADMITTANCE, g, combiners[s], bus);
      pv->cab_a.push_back(combiners[s] - 1);
      pv->cab_b.push_back(bus - 1);
    }
    C.add_linear_element(ADMITTANCE, kNortonG, bus, 0);
    C.add_linear_element(CURRENT_SOURCE, kNortonG * kNortonV, bus, 0);
  }
  pv->n = C.unknowns();
  pv->nmod = nm_total;
  pv->ncab = (int)pv->cab_g.size();
  // modules on a regular grid over 1 km^2
  const int g = (int)std::ceil(std::sqrt((double)nm_total));
  const double h = 1000.0 / g;
  pv->pos_x.resize(nm_total);
  pv->pos_y.resize(nm_total);
  for (int m = 0; m < nm_total; m++) {
    pv->pos_x[m] = (m % g + 0.5) * h;
    pv->pos_y[m] = (m / g + 0.5) * h;
  }
  // Linear part of the system (for device stamping)
  pv->lin_A = SparseMap(pv->n, pv->n);
  pv->lin_b.assign(pv->n, 0.0);
  for (int c = 0; c < pv->ncab; c++) {
    int a = pv->cab_a[c], b = pv->cab_b[c];
    double gc = pv->cab_g[c];
    pv->lin_A.write_add(a, a, gc);
    pv->lin_A.write_add(b, b, gc);
    pv->lin_A.write_add(a, b, -gc);
    pv->lin_A.write_add(b, a, -gc);
  }
  for (int bidx : pv->bus) {
    pv->lin_A.write_add(bidx, bidx, kNortonG);
    pv->lin_b[bidx] += kNortonG * kNortonV;
  }
  *out = pv;
  return PVGLU_OK;
}

extern "C" void pvglu_pv_destroy(pvglu_pv *pv) {
  if (!pv)
    return;
  newton_dev_destroy(pv->nd);
  dropin_dev_destroy(pv->dd);
  cloud_dev_destroy(pv->cd);
  pvglu_plan_destroy(pv->plan);
  delete pv;
}

extern "C" int pvglu_pv_info(const pvglu_pv *pv, int what) {
  if (!pv)
    return fail(PVGLU_E_ARG, "pv_info: null");
  switch (what) {
  case PVGLU_PV_N: return pv->n;
  case PVGLU_PV_NMODULES: return pv->nmod;
  case PVGLU_PV_NCABLES: return pv->ncab;
  case PVGLU_PV_NNZ_A: {
    SparseMap A;
    std::vector<double> b;
    pv->circuit.create_matrix_and_vector(1.0, A, b);
    return A.nnz();
  }
  default: return fail(PVGLU_E_ARG, "pv_info: unknown field %d", what);
  }
}

extern "C" int pvglu_pv_modules(const pvglu_pv *pv, int *pos, int *neg,
                                double *rsh) {
  if (!pv)
    return fail(PVGLU_E_ARG, "pv_modules: null");
  if (pos)
    memcpy(pos, pv->mod_pos.data(), sizeof(int) * pv->nmod);
  if (neg)
    memcpy(neg, pv->mod_neg.data(), sizeof(int) * pv->nmod);
  if (rsh)
    memcpy(rsh, pv->mod_rsh.data(), sizeof(double) * pv->nmod);
  return PVGLU_OK;
}

extern "C" int pvglu_pv_cables(const pvglu_pv *pv, int *a, int *b, double *g) {
  if (!pv)
    return fail(PVGLU_E_ARG, "pv_cables: null");
  if (a)
    memcpy(a, pv->cab_a.data(), sizeof(int) * pv->ncab);
  if (b)
    memcpy(b, pv->cab_b.data(), sizeof(int) * pv->ncab);
  if (g)
    memcpy(g, pv->cab_g.data(), sizeof(double) * pv->ncab);
  return PVGLU_OK;
}

extern "C" int pvglu_pv_buses(const pvglu_pv *pv, int *bus) {
  if (!pv || !bus)
    return fail(PVGLU_E_ARG, "pv_buses: bad arguments");
  memcpy(bus, pv->bus.data(), sizeof(int) * pv->I);
  return PVGLU_OK;
}

extern "C" int pvglu_pv_positions(const pvglu_pv *pv, double *x, double *y) {
  if (!pv || !x || !y)
    return fail(PVGLU_E_ARG, "pv_positions: bad arguments");
  memcpy(x, pv->pos_x.data(), sizeof(double) * pv->nmod);
  memcpy(y, pv->pos_y.data(), sizeof(double) * pv->nmod);
  return PVGLU_OK;
}

namespace pvglu {
// Seeded cloud field of one scenario: 8 Gaussian shadow blobs over a
// 1000 W/m^2 clear sky, sigma 50..300 m, depth 0.2..0.8, drift 5..15 m/s at a
// uniform heading.  prm = [cx(8) cy(8) vx(8) vy(8) sigma(8) depth(8)].
void cloud_params(const pvglu_pv *pv, int64_t scenario, double *prm) {
  std::mt19937_64 rng(pv->seed + 1 + (uint64_t)scenario);
  for (int k = 0; k < 8; k++) {
    prm[k] = 1000.0 * uniform01(rng());
    prm[8 + k] = 1000.0 * uniform01(rng());
    prm[32 + k] = 50.0 + 250.0 * uniform01(rng());
    prm[40 + k] = 0.2 + 0.6 * uniform01(rng());
    const double sp = 5.0 + 10.0 * uniform01(rng());
    const double an = 2.0 * M_PI * uniform01(rng());
    prm[16 + k] = sp * std::cos(an);
    prm[24 + k] = sp * std::sin(an);
  }
}
} // namespace pvglu

extern "C" int pvglu_pv_irradiance(const pvglu_pv *pv, int64_t scenario,
                                   double t, double *g_out) {
  if (!pv || !g_out)
    return fail(PVGLU_E_ARG, "pv_irradiance: bad arguments");
  if (scenario < 0) {
    for (int m = 0; m < pv->nmod; m++)
      g_out[m] = 1000.0;
    return PVGLU_OK;
  }
  double prm[kCloudParams];
  cloud_params(pv, scenario, prm);
  const double *cx = prm, *cy = prm + 8, *vx = prm + 16, *vy = prm + 24,
               *sg = prm + 32, *dp = prm + 40;
  for (int m = 0; m < pv->nmod; m++) {
    double shade = 0.0;
    for (int k = 0; k < 8; k++) {
      const double dx = pv->pos_x[m] - (cx[k] + vx[k] * t);
      const double dy = pv->pos_y[m] - (cy[k] + vy[k] * t);
      shade += dp[k] * std::exp(-(dx * dx + dy * dy) / (2.0 * sg[k] * sg[k]));
    }
    g_out[m] = 1000.0 * std::max(0.0, 1.0 - shade);
  }
  return PVGLU_OK;
}

extern "C" int pvglu_pv_pattern(const pvglu_pv *pv, int *rowptr,
                                int *colidx) {
  if (!pv || !rowptr)
    return fail(PVGLU_E_ARG, "pv_pattern: bad arguments");
  SparseMap A;
  std::vector<double> b;
  pv->circuit.create_matrix_and_vector(1.0, A, b);
  std::vector<int> rp, ci;
  A.csr(rp, ci);
  memcpy(rowptr, rp.data(), sizeof(int) * rp.size());
  if (colidx)
    memcpy(colidx, ci.data(), sizeof(int) * ci.size());
  return PVGLU_OK;
}

extern "C" int pvglu_pv_module_current(double V, double Iph, double rsh,
                                       double *I, double *dIdV,
                                       double *dIdIph) {
  ModuleEval e = module_current(V, Iph, rsh);
  if (I)
    *I = e.I;
  if (dIdV)
    *dIdV = e.dIdV;
  if (dIdIph)
    *dIdIph = e.dIdIph;
  return PVGLU_OK;
}
