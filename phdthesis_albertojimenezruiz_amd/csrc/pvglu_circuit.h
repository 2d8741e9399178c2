// pvglu_circuit.h -- host circuit model with the reference's API shape.
//
// SparseMap mirrors Sparse_Matrix<double> (PV/sparse_matrix.h:6-150): rows of
// ordered maps, symmetric permutation new(Pinv[i],Pinv[j]) = old(i,j)
// (interchange_row_columns, :117-134) and row permutation new[p[i]] = old[i]
// (exchange_rows, :136-147).  Circuit mirrors basic_electric_circuit
// (PV/electrical_system.h:27-115, PV/002_base_electric_system.cpp): nodes
// (0 = ground), AC nodes as a (real, imaginary) pair of consecutive nodes,
// linear admittances, current sources and voltage sources (each voltage
// source adds its current as an unknown), AC admittances as 2x2 real blocks,
// DC/AC inverters (a DC voltage source whose power is injected on the AC side
// as a current fixed from the previous iteration), and nonlinear two-terminal
// elements with companion stamping G = f'(V), I_eq = f(V) - G V plus the
// homotopy's virtual conductance (PV/002b_sistema_electrico_base_matriz.cpp).
// Unknowns are numbered in creation order, nodes and voltage-source
// currents interleaved as they are created (obt_f, PV/002_...:185-204).
#pragma once

#include <cmath>
#include <complex>
#include <functional>
#include <map>
#include <vector>

namespace pvglu {

struct SparseMap {
  int m = 0, n = 0;
  std::vector<std::map<int, double>> rows;
  SparseMap() = default;
  SparseMap(int m_, int n_) : m(m_), n(n_), rows(m_) {}
  void write(int i, int j, double v) { rows[i][j] = v; }
  double D(int i, int j) const {
    auto it = rows[i].find(j);
    return it == rows[i].end() ? 0.0 : it->second;
  }
  void write_add(int i, int j, double v) { rows[i][j] += v; }
  bool exists(int i, int j) const { return rows[i].count(j) != 0; }
  int nnz() const {
    int s = 0;
    for (auto &r : rows)
      s += (int)r.size();
    return s;
  }
  SparseMap transpose() const {
    SparseMap t(n, m);
    for (int i = 0; i < m; i++)
      for (auto &e : rows[i])
        t.rows[e.first].emplace(i, e.second);
    return t;
  }
  SparseMap interchange_row_columns(const int *new_order) const {
    SparseMap r(m, n);
    for (int i = 0; i < m; i++) {
      auto &dst = r.rows[new_order[i]];
      for (auto &e : rows[i])
        dst[new_order[e.first]] = e.second;
    }
    return r;
  }
  SparseMap exchange_rows(const int *new_order) const {
    SparseMap r(m, n);
    for (int i = 0; i < m; i++)
      r.rows[new_order[i]] = rows[i];
    return r;
  }
  void csr(std::vector<int> &rp, std::vector<int> &ci) const {
    rp.assign(1, 0);
    ci.clear();
    for (auto &r : rows) {
      for (auto &e : r)
        ci.push_back(e.first);
      rp.push_back((int)ci.size());
    }
  }
};

enum ElementType { ADMITTANCE, CURRENT_SOURCE, VOLTAGE_SOURCE };

class Circuit {
public:
  using Fn = std::function<double(double)>;
  using cplx = std::complex<double>;

  // 1-based node; its unknown is x_of(node)
  int create_node() {
    node_x_.push_back(nx_++);
    return (int)node_x_.size();
  }
  // AC node: the real part's node; the imaginary part is the next node
  int create_node_ac() {
    const int re = create_node();
    create_node();
    return re;
  }
  static int imag_of(int node) { return node == 0 ? 0 : node + 1; }
  int x_of(int node) const { return node_x_[node - 1]; }
  int unknowns() const { return nx_; }

  // Returns the unknown of a voltage source's current, else -1.
  int add_linear_element(ElementType t, double v, int n1, int n2) {
    switch (t) {
    case ADMITTANCE:
      admittances_.push_back({n1, n2, v, -1});
      return -1;
    case CURRENT_SOURCE:
      sources_.push_back({n1, n2, v, -1});
      return -1;
    case VOLTAGE_SOURCE:
      vsources_.push_back({n1, n2, v, nx_});
      return nx_++;
    }
    return -1;
  }
  // add_linear_element_ac (PV/002_base_electric_system.cpp:110-144): an AC
  // admittance is kept complex; sources split into real and imaginary parts.
  int add_linear_element_ac(ElementType t, cplx v, int n1, int n2) {
    switch (t) {
    case ADMITTANCE:
      ac_admittances_.push_back({n1, n2, v});
      return -1;
    case CURRENT_SOURCE:
      add_linear_element(CURRENT_SOURCE, v.real(), n1, n2);
      add_linear_element(CURRENT_SOURCE, v.imag(), imag_of(n1), imag_of(n2));
      return -1;
    case VOLTAGE_SOURCE: {
      const int r = add_linear_element(VOLTAGE_SOURCE, v.real(), n1, n2);
      add_linear_element(VOLTAGE_SOURCE, v.imag(), imag_of(n1), imag_of(n2));
      return r;
    }
    }
    return -1;
  }
  // element current f(V) flows n1 -> n2 through the element, V = V(n1)-V(n2)
  int add_nonlinear_voltage_source(Fn f, Fn df, double v0, int n1, int n2) {
    nonlinear_.push_back({n1, n2, std::move(f), std::move(df)});
    v_iterate.push_back(v0);
    return (int)nonlinear_.size() - 1;
  }
  // create_inverter_dc_ac (PV/002_base_electric_system.cpp:146-162): a DC
  // voltage source V_dc between the DC nodes, whose power P0 = V_dc I1 leaves
  // on the AC side between the AC nodes as the current I2 = P0 V2 / |V2|^2,
  // with I1 (the source's current) and V2 (the AC voltage) taken from the
  // previous iteration (PV/002b_sistema_electrico_base_matriz.cpp:205-237).
  int create_inverter_dc_ac(double v_dc, cplx v_initial_ac, int n1_dc,
                            int n2_dc, int n1_ac, int n2_ac) {
    const int vs = add_linear_element(VOLTAGE_SOURCE, v_dc, n1_dc, n2_dc);
    inverters_.push_back({n1_ac, n2_ac, vs, v_dc});
    inv_iterate.push_back({v_initial_ac.real(), v_initial_ac.imag(), 0.0});
    return (int)inverters_.size() - 1;
  }

  double voltage_drop_dc(const std::vector<double> &x, int n1, int n2) const {
    double r = 0;
    if (n1)
      r = x[x_of(n1)];
    if (n2)
      r -= x[x_of(n2)];
    return r;
  }
  cplx voltage_drop_ac(const std::vector<double> &x, int n1, int n2) const {
    return {voltage_drop_dc(x, n1, n2),
            voltage_drop_dc(x, imag_of(n1), imag_of(n2))};
  }
  void create_matrix_and_vector(double virtual_g, SparseMap &A,
                                std::vector<double> &b) const;
  // b - A x of create_matrix_and_vector's system at unknowns x, evaluated
  // element by element: the sources minus the currents leaving each node
  // through admittances g (V1 - V2), AC admittances Y (V1 - V2), voltage
  // sources (their current unknowns) and nonlinear elements
  // f(V) + f'(V) (Vx - V) [+ virtual_g Vx] (V the element's linearization
  // voltage, Vx its voltage at x), and E - (V1 - V2) in a voltage source's
  // own row.  Each voltage difference is formed before it is scaled.  The
  // right-hand side of the correction form A d = b - A x (no reference
  // counterpart).
  void residual(double virtual_g, const std::vector<double> &x,
                std::vector<double> &r) const;
  // The 011 inverter refresh (PV/011_calculate_new_iteration.cpp:88-109):
  // each inverter's (V2 real, V2 imaginary, I1) from the solution x; returns
  // the sum of their absolute changes (added to the iteration's diff) and
  // the largest one in max_change.
  double refresh_inverters(const std::vector<double> &x,
                           double &max_change);

  int nonlinear_count() const { return (int)nonlinear_.size(); }
  int inverter_count() const { return (int)inverters_.size(); }
  std::pair<int, int> nonlinear_nodes(int e) const {
    return {nonlinear_[e].n1, nonlinear_[e].n2};
  }
  // element lists, for exporting a circuit (node numbers 1-based, 0 ground)
  struct Lin {
    int n1, n2;
    double v;
    int x; // voltage sources: the current's unknown
  };
  struct AcLin {
    int n1, n2;
    cplx v;
  };
  struct Inverter {
    int n1_ac, n2_ac, vs_x;
    double v_dc;
  };
  const std::vector<Lin> &admittances() const { return admittances_; }
  const std::vector<Lin> &current_sources() const { return sources_; }
  const std::vector<Lin> &voltage_sources() const { return vsources_; }
  const std::vector<AcLin> &ac_admittances() const { return ac_admittances_; }
  const std::vector<Inverter> &inverters() const { return inverters_; }
  double nonlinear_f(int e, double V) const { return nonlinear_[e].f(V); }
  double nonlinear_df(int e, double V) const { return nonlinear_[e].df(V); }

  std::vector<double> v_iterate; // vector_voltages_nonlinear_currentsources_iterate
  struct InvState {
    double v2r, v2i, i1;
  };
  std::vector<InvState> inv_iterate; // list_values_inverter_dc_ac_iterate

private:
  struct NonLin {
    int n1, n2;
    Fn f, df;
  };
  void stamp_g(SparseMap &A, int n1, int n2, double g) const {
    if (n1)
      A.write_add(x_of(n1), x_of(n1), g);
    if (n2)
      A.write_add(x_of(n2), x_of(n2), g);
    if (n1 && n2) {
      A.write_add(x_of(n1), x_of(n2), -g);
      A.write_add(x_of(n2), x_of(n1), -g);
    }
  }
  void stamp_i(std::vector<double> &b, int n1, int n2, double v) const {
    if (n1)
      b[x_of(n1)] += v;
    if (n2)
      b[x_of(n2)] -= v;
  }
  // one 2x2 block of an AC admittance: rows (p, p+1), columns (q, q+1)
  // receive [[Re v, -Im v], [Im v, Re v]] (add_submatrix_2x2_complex_impedance)
  void stamp_block(SparseMap &A, int p, int q, cplx v) const {
    if (!p || !q)
      return;
    const int i = x_of(p), j = x_of(q);
    A.write_add(i, j, v.real());
    A.write_add(i, j + 1, -v.imag());
    A.write_add(i + 1, j, v.imag());
    A.write_add(i + 1, j + 1, v.real());
  }
  int nx_ = 0;
  std::vector<int> node_x_;
  std::vector<Lin> admittances_, sources_, vsources_;
  std::vector<AcLin> ac_admittances_;
  std::vector<NonLin> nonlinear_;
  std::vector<Inverter> inverters_;
};

// create_matrix_and_vector (PV/002b_sistema_electrico_base_matriz.cpp:158-240):
// admittances, AC admittances, current sources, voltage sources, nonlinear
// elements (companion model, then the virtual conductance), inverter
// injections -- in that order, since entries shared by several elements are
// summed in it.
inline void Circuit::create_matrix_and_vector(double virtual_g, SparseMap &A,
                                              std::vector<double> &b) const {
  A = SparseMap(nx_, nx_);
  b.assign(nx_, 0.0);
  for (auto &e : admittances_)
    stamp_g(A, e.n1, e.n2, e.v);
  for (auto &e : ac_admittances_) {
    stamp_block(A, e.n1, e.n1, e.v);
    stamp_block(A, e.n1, e.n2, -e.v);
    stamp_block(A, e.n2, e.n1, -e.v);
    stamp_block(A, e.n2, e.n2, e.v);
  }
  for (auto &e : sources_)
    stamp_i(b, e.n1, e.n2, e.v);
  for (auto &e : vsources_) {
    if (e.n1) {
      A.write_add(x_of(e.n1), e.x, -1.0);
      A.write_add(e.x, x_of(e.n1), 1.0);
    }
    if (e.n2) {
      A.write_add(x_of(e.n2), e.x, 1.0);
      A.write_add(e.x, x_of(e.n2), -1.0);
    }
    b[e.x] += e.v;
  }
  for (size_t k = 0; k < nonlinear_.size(); k++) {
    const auto &e = nonlinear_[k];
    const double V = v_iterate[k];
    const double g = e.df(V);
    const double ieq = e.f(V) - g * V;
    stamp_g(A, e.n1, e.n2, g);
    stamp_i(b, e.n1, e.n2, -ieq);
    if (virtual_g > 0)
      stamp_g(A, e.n1, e.n2, virtual_g);
  }
  for (size_t k = 0; k < inverters_.size(); k++) {
    const auto &e = inverters_[k];
    const auto &s = inv_iterate[k];
    const double den = s.v2r * s.v2r + s.v2i * s.v2i;
    const double p0 = e.v_dc * s.i1;
    stamp_i(b, e.n1_ac, e.n2_ac, -(p0 * s.v2r / den));
    stamp_i(b, imag_of(e.n1_ac), imag_of(e.n2_ac), -(p0 * s.v2i / den));
  }
}

inline void Circuit::residual(double virtual_g, const std::vector<double> &x,
                              std::vector<double> &r) const {
  r.assign(nx_, 0.0);
  auto leave = [&](int n1, int n2, double i) {
    if (n1)
      r[x_of(n1)] -= i;
    if (n2)
      r[x_of(n2)] += i;
  };
  for (auto &e : admittances_)
    leave(e.n1, e.n2, e.v * voltage_drop_dc(x, e.n1, e.n2));
  for (auto &e : ac_admittances_) {
    const cplx i = e.v * voltage_drop_ac(x, e.n1, e.n2);
    leave(e.n1, e.n2, i.real());
    leave(imag_of(e.n1), imag_of(e.n2), i.imag());
  }
  for (auto &e : sources_)
    stamp_i(r, e.n1, e.n2, e.v);
  for (auto &e : vsources_) {
    leave(e.n1, e.n2, -x[e.x]); // the source's current enters n1
    r[e.x] += e.v - voltage_drop_dc(x, e.n1, e.n2);
  }
  for (size_t k = 0; k < nonlinear_.size(); k++) {
    const auto &e = nonlinear_[k];
    const double V = v_iterate[k];
    const double Vx = voltage_drop_dc(x, e.n1, e.n2);
    double i = e.f(V) + e.df(V) * (Vx - V);
    if (virtual_g > 0)
      i = i + virtual_g * Vx;
    leave(e.n1, e.n2, i);
  }
  for (size_t k = 0; k < inverters_.size(); k++) {
    const auto &e = inverters_[k];
    const auto &s = inv_iterate[k];
    const double den = s.v2r * s.v2r + s.v2i * s.v2i;
    const double p0 = e.v_dc * s.i1;
    stamp_i(r, e.n1_ac, e.n2_ac, -(p0 * s.v2r / den));
    stamp_i(r, imag_of(e.n1_ac), imag_of(e.n2_ac), -(p0 * s.v2i / den));
  }
}

inline double Circuit::refresh_inverters(const std::vector<double> &x,
                                         double &max_change) {
  double sum = 0.0;
  max_change = 0.0;
  for (size_t k = 0; k < inverters_.size(); k++) {
    const auto &e = inverters_[k];
    InvState &s = inv_iterate[k];
    const InvState old = s;
    s.v2r = voltage_drop_dc(x, e.n1_ac, e.n2_ac);
    s.v2i = voltage_drop_dc(x, imag_of(e.n1_ac), imag_of(e.n2_ac));
    s.i1 = x[e.vs_x];
    const double d[3] = {std::fabs(s.v2r - old.v2r), std::fabs(s.v2i - old.v2i),
                         std::fabs(s.i1 - old.i1)};
    sum += d[0] + d[1] + d[2];
    for (double v : d)
      max_change = std::fmax(max_change, v);
  }
  return sum;
}

} // namespace pvglu
