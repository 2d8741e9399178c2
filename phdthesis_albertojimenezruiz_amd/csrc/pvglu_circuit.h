// pvglu_circuit.h -- host circuit model with the reference's API shape.
//
// SparseMap mirrors Sparse_Matrix<double> (PV/sparse_matrix.h:6-150): rows of
// ordered maps, symmetric permutation new(Pinv[i],Pinv[j]) = old(i,j)
// (interchange_row_columns, :117-134) and row permutation new[p[i]] = old[i]
// (exchange_rows, :136-147).  Circuit mirrors basic_electric_circuit
// (PV/electrical_system.h:27-115): nodes (0 = ground), linear admittances and
// current sources, nonlinear two-terminal elements with companion stamping
// G = f'(V), I_eq = f(V) - G V, plus the homotopy's virtual conductance
// (PV/002b_sistema_electrico_base_matriz.cpp:141-203).
#pragma once

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

enum ElementType { ADMITTANCE, CURRENT_SOURCE };

class Circuit {
public:
  using Fn = std::function<double(double)>;
  int create_node() { return ++nodes_; } // 1-based; index = node - 1
  int unknowns() const { return nodes_; }
  void add_linear_element(ElementType t, double v, int n1, int n2) {
    (t == ADMITTANCE ? admittances_ : sources_).push_back({n1, n2, v});
  }
  // element current f(V) flows n1 -> n2 through the element, V = V(n1)-V(n2)
  int add_nonlinear_voltage_source(Fn f, Fn df, double v0, int n1, int n2) {
    nonlinear_.push_back({n1, n2, std::move(f), std::move(df)});
    v_iterate.push_back(v0);
    return (int)nonlinear_.size() - 1;
  }
  double voltage_drop_dc(const std::vector<double> &x, int n1, int n2) const {
    double r = 0;
    if (n1)
      r = x[n1 - 1];
    if (n2)
      r -= x[n2 - 1];
    return r;
  }
  void create_matrix_and_vector(double virtual_g, SparseMap &A,
                                std::vector<double> &b) const;
  // b - A x of create_matrix_and_vector's system at node voltages x
  // (x[k - 1] for node k), evaluated branch by branch: the sources minus the
  // currents leaving each node through admittances v (V1 - V2) and through
  // nonlinear elements f(V) + f'(V) (Vx - V) [+ virtual_g Vx], V the
  // element's linearization voltage and Vx its voltage at x.  The right-hand
  // side of the correction form A d = b - A x (no reference counterpart).
  void residual(double virtual_g, const std::vector<double> &x,
                std::vector<double> &r) const;
  int nonlinear_count() const { return (int)nonlinear_.size(); }
  std::pair<int, int> nonlinear_nodes(int e) const {
    return {nonlinear_[e].n1, nonlinear_[e].n2};
  }

  std::vector<double> v_iterate; // vector_voltages_nonlinear_currentsources_iterate

private:
  struct Lin {
    int n1, n2;
    double v;
  };
  struct NonLin {
    int n1, n2;
    Fn f, df;
  };
  static void stamp_g(SparseMap &A, int n1, int n2, double g) {
    if (n1)
      A.write_add(n1 - 1, n1 - 1, g);
    if (n2)
      A.write_add(n2 - 1, n2 - 1, g);
    if (n1 && n2) {
      A.write_add(n1 - 1, n2 - 1, -g);
      A.write_add(n2 - 1, n1 - 1, -g);
    }
  }
  static void stamp_i(std::vector<double> &b, int n1, int n2, double v) {
    if (n1)
      b[n1 - 1] += v;
    if (n2)
      b[n2 - 1] -= v;
  }
  int nodes_ = 0;
  std::vector<Lin> admittances_, sources_;
  std::vector<NonLin> nonlinear_;
};

inline void Circuit::create_matrix_and_vector(double virtual_g, SparseMap &A,
                                              std::vector<double> &b) const {
  A = SparseMap(nodes_, nodes_);
  b.assign(nodes_, 0.0);
  for (auto &e : admittances_)
    stamp_g(A, e.n1, e.n2, e.v);
  for (auto &e : sources_)
    stamp_i(b, e.n1, e.n2, e.v);
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
}

inline void Circuit::residual(double virtual_g, const std::vector<double> &x,
                              std::vector<double> &r) const {
  r.assign(nodes_, 0.0);
  auto leave = [&](int n1, int n2, double i) {
    if (n1)
      r[n1 - 1] -= i;
    if (n2)
      r[n2 - 1] += i;
  };
  for (auto &e : admittances_)
    leave(e.n1, e.n2, e.v * voltage_drop_dc(x, e.n1, e.n2));
  for (auto &e : sources_)
    stamp_i(r, e.n1, e.n2, e.v);
  for (size_t k = 0; k < nonlinear_.size(); k++) {
    const auto &e = nonlinear_[k];
    const double V = v_iterate[k];
    const double Vx = voltage_drop_dc(x, e.n1, e.n2);
    double i = e.f(V) + e.df(V) * (Vx - V);
    if (virtual_g > 0)
      i = i + virtual_g * Vx;
    leave(e.n1, e.n2, i);
  }
}

} // namespace pvglu
