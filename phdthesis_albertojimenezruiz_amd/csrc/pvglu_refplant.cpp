// pvglu_refplant.cpp -- the reference's own PV plant on the circuit API.
//
// create_full_electric_system (PV/003_create_electric_system.cpp:313-382)
// with create_block_of_panels_simple (:184-311) and add_solar_panel
// (:87-126): per block, panels_y strings of panels_x panels between two
// block nodes (optionally cross-tied between strings), wires (1e-5 Ohm) to
// the block's two external nodes, the negative one tied to ground by a wire,
// an inverter (DC voltage source of panels_x * volts_per_panel, AC injection)
// onto two AC nodes; the AC nodes of consecutive blocks joined by wires, and
// the first and last block's AC nodes joined to the point of common coupling,
// a 25 kV AC voltage source to ground.  A panel: an internal node, Rs to the
// panel's + node, Rsh and the diode (a nonlinear element) to its - node, the
// photocurrent into the internal node, and a bypass diode from - to +.
// The panel constants are the reference's (:7-25); the diode models
// func_diode_panel / func_diode_bypass (:27-85) are written out in the same
// operation order.  The unknowns are numbered as the reference creates them
// (panel internal nodes and string nodes interleaved, voltage-source
// currents where the sources are added), so its voltage sources put zero
// diagonals in the system and reorder_matrix's row swaps are needed, and the
// AC ring, the parallel strings and the cross-ties give fill-in.
#include "pvglu_pv.h"

#include <cmath>
#include <cstring>
#include <vector>

using namespace pvglu;

namespace {

// PV/003_create_electric_system.cpp:7-25 (same constants, same arithmetic)
constexpr double kQ = 1.60217662E-19;
constexpr double kK = 1.38064852E-23;
constexpr double kNa = 48.548;
constexpr double kT0 = 25;
constexpr double kG0 = 1000;
const double kRs = 0.262 * (175.0 / 400.0);
constexpr double kRsh = 90.85;
const double kI0 = 4.86E-10 * (400.0 / 175.0);
constexpr double kK0Percent = -0.45;
const double kIpv = 8.11 * 400.0 / 175.0 * 1.05;
const double kK0 = kK0Percent / 100.0;
constexpr double kWire = 1e-5;      // resistance_wire (PV/main.h:76)
constexpr double kVac = 25000;      // external_voltage_AC (PV/main.h:7)

double qkT_of(double temperature) { return kQ / kK / (273.15 + temperature); }

double diode_panel(double V, double T) {
  const double qkT = qkT_of(T);
  return kI0 * (std::exp(qkT * V / kNa) - 1.0);
}
double diode_panel_d(double V, double T) {
  const double qkT = qkT_of(T);
  return kI0 * qkT / kNa * std::exp(V * qkT / kNa);
}
double diode_bypass(double V, double T) {
  const double qkT = qkT_of(T);
  return 0.3E-6 * (std::exp(qkT * V) - 1);
}
double diode_bypass_d(double V, double T) {
  const double qkT = qkT_of(T);
  return qkT * 0.3E-6 * (std::exp(qkT * V));
}

// add_solar_panel (PV/003_create_electric_system.cpp:87-126)
void add_panel(Circuit &c, int node_in, int node_out, double irradiance,
               double T, bool bypass) {
  const int internal = c.create_node();
  c.add_linear_element(ADMITTANCE, 1.0 / kRs, internal, node_in);
  c.add_linear_element(ADMITTANCE, 1.0 / kRsh, internal, node_out);
  c.add_nonlinear_voltage_source([T](double V) { return diode_panel(V, T); },
                                 [T](double V) { return diode_panel_d(V, T); },
                                 0.0, internal, node_out);
  const double irr = kIpv * irradiance / kG0;
  const double loss = kK0 * (T - kT0);
  c.add_linear_element(CURRENT_SOURCE, irr * (1.0 + loss), internal, node_out);
  if (bypass)
    c.add_nonlinear_voltage_source(
        [T](double V) { return diode_bypass(V, T); },
        [T](double V) { return diode_bypass_d(V, T); }, 0.0, node_out, node_in);
}

// create_block_of_panels_simple (:184-311), without broken panels and the
// aggregated model: returns the block's (+, -) external nodes
std::pair<int, int> add_block(pvglu_refplant *rp, int block) {
  Circuit &c = rp->circuit;
  const int in_before = c.create_node();
  const int out_before = c.create_node();
  std::vector<std::vector<std::pair<int, int>>> nodes(rp->py);
  for (int i = 0; i < rp->py; i++) {
    std::pair<int, int> np{0, 0};
    for (int j = 0; j < rp->px; j++) {
      const int first = j == rp->px - 1 ? in_before : c.create_node();
      const int second = j == 0 ? out_before : np.first;
      np = {first, second};
      nodes[i].push_back(np);
      const double g =
          rp->panel_g[((size_t)block * rp->py + i) * rp->px + j];
      add_panel(c, np.first, np.second, g, rp->temperature, rp->bypass != 0);
    }
  }
  for (int i = 0; i < rp->py - 1; i++)
    for (int j = 0; j < rp->px - 1; j++) {
      bool tie = false;
      switch (rp->interconnection) {
      case PVGLU_INTERCONNECTION_FULL:
        tie = true;
        break;
      case PVGLU_INTERCONNECTION_HONEYCOMB:
        tie = ((i + j) % 2) == 1;
        break;
      case PVGLU_INTERCONNECTION_BRIDGE:
        tie = (i % 3) == 0;
        if ((j % 2) == 0)
          tie = !tie;
        break;
      default:
        break;
      }
      if (tie)
        c.add_linear_element(ADMITTANCE, 1.0 / kWire, nodes[i][j].first,
                             nodes[i + 1][j].first);
    }
  const int ext_in = c.create_node(), ext_out = c.create_node();
  c.add_linear_element(ADMITTANCE, 1.0 / kWire, ext_in, in_before);
  c.add_linear_element(ADMITTANCE, 1.0 / kWire, ext_out, out_before);
  return {ext_in, ext_out};
}

} // namespace

extern "C" int pvglu_refplant_create(int blocks, int panels_x, int panels_y,
                                     double volts_per_panel,
                                     int interconnection, double irradiance,
                                     const double *panel_irradiance,
                                     double temperature, int bypass,
                                     pvglu_refplant **out) {
  clear_error();
  if (!out || blocks <= 0 || panels_x <= 0 || panels_y <= 0 ||
      interconnection < PVGLU_INTERCONNECTION_NOTHING ||
      interconnection > PVGLU_INTERCONNECTION_BRIDGE)
    return fail(PVGLU_E_ARG, "refplant_create: bad arguments");
  auto *rp = new pvglu_refplant();
  rp->blocks = blocks;
  rp->px = panels_x;
  rp->py = panels_y;
  rp->volts_per_panel = volts_per_panel;
  rp->interconnection = interconnection;
  rp->irradiance = irradiance;
  rp->temperature = temperature;
  rp->bypass = bypass;
  rp->npanels = blocks * panels_x * panels_y;
  rp->panel_g.assign((size_t)rp->npanels, irradiance);
  if (panel_irradiance)
    memcpy(rp->panel_g.data(), panel_irradiance, sizeof(double) * rp->npanels);
  Circuit &c = rp->circuit;
  const double v_pcc = panels_x * volts_per_panel;
  std::vector<std::pair<int, int>> main_nodes;
  for (int b = 0; b < blocks; b++) {
    const auto ext = add_block(rp, b);
    // main_nodes.push_back(int_2(create_node_ac(), create_node_ac())) (:331-
    // 332) leaves the order of the two calls to the compiler; taken left to
    // right here
    const int ac1 = c.create_node_ac();
    const int ac2 = c.create_node_ac();
    main_nodes.push_back({ac1, ac2});
    // path so that the DC circuit is fixed to ground
    c.add_linear_element(ADMITTANCE, 1.0 / kWire, ext.second, 0);
    c.create_inverter_dc_ac(v_pcc, {kVac, 0.0}, ext.first, ext.second, ac1,
                            ac2);
    if (b != 0) {
      c.add_linear_element_ac(ADMITTANCE, {1.0 / kWire, 0.0}, main_nodes[b].first,
                              main_nodes[b - 1].first);
      c.add_linear_element_ac(ADMITTANCE, {1.0 / kWire, 0.0},
                              main_nodes[b].second, main_nodes[b - 1].second);
    }
  }
  rp->pcc_node = c.create_node_ac();
  rp->pcc_current =
      c.add_linear_element_ac(VOLTAGE_SOURCE, {kVac, 0.0}, rp->pcc_node, 0);
  c.add_linear_element_ac(ADMITTANCE, {1.0 / kWire, 0.0}, rp->pcc_node,
                          main_nodes[0].first);
  c.add_linear_element_ac(ADMITTANCE, {1.0 / kWire, 0.0}, 0,
                          main_nodes[0].second);
  c.add_linear_element_ac(ADMITTANCE, {1.0 / kWire, 0.0}, rp->pcc_node,
                          main_nodes.back().first);
  c.add_linear_element_ac(ADMITTANCE, {1.0 / kWire, 0.0}, 0,
                          main_nodes.back().second);
  rp->n = c.unknowns();
  *out = rp;
  return PVGLU_OK;
}

extern "C" void pvglu_refplant_destroy(pvglu_refplant *rp) {
  if (!rp)
    return;
  dropin_dev_destroy(rp->dd);
  pvglu_plan_destroy(rp->plan);
  delete rp;
}

extern "C" int pvglu_refplant_info(const pvglu_refplant *rp, int what) {
  if (!rp)
    return fail(PVGLU_E_ARG, "refplant_info: null");
  const Circuit &c = rp->circuit;
  switch (what) {
  case PVGLU_REF_N: return rp->n;
  case PVGLU_REF_NNZ_A: {
    SparseMap A;
    std::vector<double> b;
    c.create_matrix_and_vector(1.0, A, b);
    return A.nnz();
  }
  case PVGLU_REF_NONLINEAR: return c.nonlinear_count();
  case PVGLU_REF_INVERTERS: return c.inverter_count();
  case PVGLU_REF_VSOURCES: return (int)c.voltage_sources().size();
  case PVGLU_REF_PANELS: return rp->npanels;
  case PVGLU_REF_PCC_NODE: return rp->pcc_node;
  case PVGLU_REF_PCC_CURRENT: return rp->pcc_current;
  default: return fail(PVGLU_E_ARG, "refplant_info: unknown field %d", what);
  }
}

extern "C" int pvglu_refplant_pattern(const pvglu_refplant *rp, int *rowptr,
                                      int *colidx) {
  if (!rp || !rowptr)
    return fail(PVGLU_E_ARG, "refplant_pattern: bad arguments");
  SparseMap A;
  std::vector<double> b;
  rp->circuit.create_matrix_and_vector(1.0, A, b);
  std::vector<int> r, ci;
  A.csr(r, ci);
  memcpy(rowptr, r.data(), sizeof(int) * r.size());
  if (colidx)
    memcpy(colidx, ci.data(), sizeof(int) * ci.size());
  return PVGLU_OK;
}

static int refplant_prepare(pvglu_refplant *rp) {
  if (rp->prepared)
    return PVGLU_OK;
  int st = prepare_circuit(rp->circuit, rp->P, rp->Pinv, rp->Prow, &rp->plan);
  if (st)
    return st;
  rp->prepared = true;
  return PVGLU_OK;
}

extern "C" pvglu_plan *pvglu_refplant_plan(pvglu_refplant *rp) {
  clear_error();
  if (!rp) {
    fail(PVGLU_E_ARG, "refplant_plan: null");
    return nullptr;
  }
  if (refplant_prepare(rp))
    return nullptr;
  return rp->plan;
}

extern "C" int pvglu_refplant_ordering(pvglu_refplant *rp, int *P, int *Prow) {
  clear_error();
  if (!rp || !P || !Prow)
    return fail(PVGLU_E_ARG, "refplant_ordering: bad arguments");
  int st = refplant_prepare(rp);
  if (st)
    return st;
  memcpy(P, rp->P.data(), sizeof(int) * rp->n);
  memcpy(Prow, rp->Prow.data(), sizeof(int) * rp->n);
  return PVGLU_OK;
}

// Elements, for an independent restatement of the assembly, in the order
// create_matrix_and_vector stamps them; returns their count.  Per element
// `elem` holds 5 ints (kind, n1, n2, aux, index) and `val` 2 doubles:
// admittance (g), AC admittance (re, im), current source (value), voltage
// source (E; aux = its current's unknown), nonlinear element (aux = 0 panel
// diode, 1 bypass diode; index = its position; val[0] = temperature),
// inverter (aux = its DC source's current unknown; val[0] = V_dc).
extern "C" int pvglu_refplant_elements(const pvglu_refplant *rp, int *elem,
                                       double *val) {
  if (!rp)
    return fail(PVGLU_E_ARG, "refplant_elements: null");
  const Circuit &c = rp->circuit;
  int k = 0;
  auto put = [&](int kind, int n1, int n2, int aux, int idx, double v0,
                 double v1) {
    if (elem) {
      elem[5 * k] = kind;
      elem[5 * k + 1] = n1;
      elem[5 * k + 2] = n2;
      elem[5 * k + 3] = aux;
      elem[5 * k + 4] = idx;
    }
    if (val) {
      val[2 * k] = v0;
      val[2 * k + 1] = v1;
    }
    k++;
  };
  for (auto &e : c.admittances())
    put(PVGLU_ELEM_ADMITTANCE, e.n1, e.n2, 0, 0, e.v, 0);
  for (auto &e : c.ac_admittances())
    put(PVGLU_ELEM_AC_ADMITTANCE, e.n1, e.n2, 0, 0, e.v.real(), e.v.imag());
  for (auto &e : c.current_sources())
    put(PVGLU_ELEM_CURRENT_SOURCE, e.n1, e.n2, 0, 0, e.v, 0);
  for (auto &e : c.voltage_sources())
    put(PVGLU_ELEM_VOLTAGE_SOURCE, e.n1, e.n2, e.x, 0, e.v, 0);
  // nonlinear elements in creation order: each panel's diode, then its
  // bypass diode when present
  for (int e = 0; e < c.nonlinear_count(); e++) {
    const auto t = c.nonlinear_nodes(e);
    const int kind = rp->bypass ? (e % 2) : 0;
    put(PVGLU_ELEM_NONLINEAR, t.first, t.second, kind, e, rp->temperature, 0);
  }
  int i = 0;
  for (auto &e : c.inverters())
    put(PVGLU_ELEM_INVERTER, e.n1_ac, e.n2_ac, e.vs_x, i++, e.v_dc, 0);
  return k;
}

extern "C" int pvglu_refplant_node_unknowns(const pvglu_refplant *rp,
                                            int *node_x) {
  if (!rp)
    return fail(PVGLU_E_ARG, "refplant_node_unknowns: null");
  // nodes are 1..N in creation order; N = unknowns - voltage sources
  const int nnodes = rp->n - (int)rp->circuit.voltage_sources().size();
  if (node_x)
    for (int node = 1; node <= nnodes; node++)
      node_x[node - 1] = rp->circuit.x_of(node);
  return nnodes;
}

extern "C" int pvglu_refplant_set_newton_form(pvglu_refplant *rp, int form) {
  clear_error();
  if (!rp || (form != PVGLU_NEWTON_CORRECTION && form != PVGLU_NEWTON_DIRECT))
    return fail(PVGLU_E_ARG, "refplant_set_newton_form: bad arguments");
  rp->form = form;
  return PVGLU_OK;
}

// Host assembly of the system at the current iterate (nonlinear voltages and
// inverter states), permuted and in the plan's CSC slot order: the values
// factorize_electrical_system receives (PV/000_PVPP_simulator.cpp:185-210).
extern "C" int pvglu_refplant_system(pvglu_refplant *rp, double g_virtual,
                                     double *a_csc, double *b_perm) {
  clear_error();
  if (!rp || !a_csc || !b_perm)
    return fail(PVGLU_E_ARG, "refplant_system: bad arguments");
  int st = refplant_prepare(rp);
  if (st)
    return st;
  const int n = rp->n;
  SparseMap A;
  std::vector<double> b, tmp(n);
  rp->circuit.create_matrix_and_vector(g_virtual, A, b);
  A = A.interchange_row_columns(rp->Pinv.data());
  A = A.exchange_rows(rp->Prow.data());
  for (int i = 0; i < n; i++)
    tmp[rp->Pinv[i]] = b[i];
  for (int i = 0; i < n; i++)
    b_perm[rp->Prow[i]] = tmp[i];
  const pvglu_plan *P = rp->plan;
  for (int c = 0; c < n; c++)
    for (int q = P->cp[c]; q < P->cp[c + 1]; q++)
      a_csc[q] = A.D(P->ri[q], c);
  return PVGLU_OK;
}

// Iterate state: nonlinear element voltages and inverter (V2 re, V2 im, I1);
// get copies them out, set in (either pointer may be NULL).
extern "C" int pvglu_refplant_get_state(const pvglu_refplant *rp, double *v,
                                        double *inv) {
  if (!rp)
    return fail(PVGLU_E_ARG, "refplant_get_state: null");
  const Circuit &c = rp->circuit;
  if (v)
    memcpy(v, c.v_iterate.data(), sizeof(double) * c.v_iterate.size());
  if (inv)
    for (size_t k = 0; k < c.inv_iterate.size(); k++) {
      inv[3 * k] = c.inv_iterate[k].v2r;
      inv[3 * k + 1] = c.inv_iterate[k].v2i;
      inv[3 * k + 2] = c.inv_iterate[k].i1;
    }
  return PVGLU_OK;
}

extern "C" int pvglu_refplant_set_state(pvglu_refplant *rp, const double *v,
                                        const double *inv) {
  if (!rp)
    return fail(PVGLU_E_ARG, "refplant_set_state: null");
  Circuit &c = rp->circuit;
  if (v)
    memcpy(c.v_iterate.data(), v, sizeof(double) * c.v_iterate.size());
  if (inv)
    for (size_t k = 0; k < c.inv_iterate.size(); k++)
      c.inv_iterate[k] = {inv[3 * k], inv[3 * k + 1], inv[3 * k + 2]};
  return PVGLU_OK;
}

// The reference's Newton loop on this plant (PV/000_PVPP_simulator.cpp:144-
// 235): ladder, host assembly, permutations, GPU refactor and solve, the 011
// update with the inverter refresh.  x_out (n): unknowns in circuit order.
// stats[0..5]: iterations, host ms, device ms, total ms, PCC power (real,
// imaginary part of V conj(I) of the PCC source, PV/000_...:237-253).
extern "C" int pvglu_refplant_newton_dropin(pvglu_refplant *rp, double tol,
                                            int ladder, double *x_out,
                                            double *stats) {
  clear_error();
  if (!rp)
    return fail(PVGLU_E_ARG, "refplant_newton_dropin: null");
  int st = refplant_prepare(rp);
  if (st)
    return st;
  std::vector<double> x(rp->n);
  double s4[4] = {0, 0, 0, 0};
  if ((st = dropin_newton(rp->circuit, rp->plan, rp->P, rp->Pinv, rp->Prow,
                          rp->dd, rp->form, tol, ladder, 0, x.data(), s4)))
    return st;
  const Circuit &c = rp->circuit;
  const std::complex<double> V = c.voltage_drop_ac(x, rp->pcc_node, 0);
  const std::complex<double> I(x[rp->pcc_current], x[rp->pcc_current + 1]);
  const std::complex<double> S = V * std::conj(I);
  if (x_out)
    memcpy(x_out, x.data(), sizeof(double) * rp->n);
  if (stats) {
    memcpy(stats, s4, sizeof s4);
    stats[4] = S.real();
    stats[5] = S.imag();
  }
  return PVGLU_OK;
}
