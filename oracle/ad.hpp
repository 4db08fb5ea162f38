// Algorithmic differentiation of the oracle (test infrastructure).
//
// Rev: reverse mode with a global tape.  Every operation on Rev values
// appends a node with the partial derivatives with respect to its (at most
// three) operands; a reverse sweep from seeded output adjoints gives exact
// vector-Jacobian products of any templated oracle function.  This is the
// reference against which the hand-written adjoints are checked ("the
// transpose of the assembled CPU operator", one row combination at a time).
//
// Fwd: forward mode (one tangent), exact Jacobian-vector products, for the
// dot-product identity <gbar, J v> = <J^T gbar, v>.
#pragma once

#include <cmath>
#include <cstddef>
#include <vector>

namespace sxo {

struct Tape {
  struct Node {
    int p[3];
    double d[3];
  };
  std::vector<Node> nodes;
  std::vector<double> adj;
  int push(int p0, double d0, int p1 = -1, double d1 = 0.0, int p2 = -1, double d2 = 0.0) {
    nodes.push_back({{p0, p1, p2}, {d0, d1, d2}});
    return static_cast<int>(nodes.size()) - 1;
  }
  void clear() {
    nodes.clear();
    adj.clear();
  }
  // adjoints of every node given seeds (node id -> adjoint)
  void reverse() {
    adj.resize(nodes.size(), 0.0);
    for (size_t k = nodes.size(); k-- > 0;) {
      const double a = adj[k];
      if (a == 0.0) continue;
      const Node& n = nodes[k];
      for (int q = 0; q < 3; ++q)
        if (n.p[q] >= 0) adj[n.p[q]] += n.d[q] * a;
    }
  }
  static Tape& get() {
    static Tape t;
    return t;
  }
};

struct Rev {
  double v = 0.0;
  int id = -1;  // -1: a constant
  Rev() = default;
  Rev(double x) : v(x) {}  // NOLINT: constants convert implicitly
  static Rev input(double x) {
    Rev r(x);
    r.id = Tape::get().push(-1, 0.0);
    return r;
  }
};
inline Rev make_rev(double v, int p0, double d0, int p1 = -1, double d1 = 0.0, int p2 = -1,
                    double d2 = 0.0) {
  Rev r(v);
  if (p0 >= 0 || p1 >= 0 || p2 >= 0) r.id = Tape::get().push(p0, d0, p1, d1, p2, d2);
  return r;
}
inline Rev operator+(const Rev& a, const Rev& b) { return make_rev(a.v + b.v, a.id, 1.0, b.id, 1.0); }
inline Rev operator-(const Rev& a, const Rev& b) { return make_rev(a.v - b.v, a.id, 1.0, b.id, -1.0); }
inline Rev operator*(const Rev& a, const Rev& b) { return make_rev(a.v * b.v, a.id, b.v, b.id, a.v); }
inline Rev operator/(const Rev& a, const Rev& b) {
  return make_rev(a.v / b.v, a.id, 1.0 / b.v, b.id, -a.v / (b.v * b.v));
}
inline Rev operator-(const Rev& a) { return make_rev(-a.v, a.id, -1.0); }
inline Rev fma_(const Rev& a, const Rev& b, const Rev& c) {
  return make_rev(std::fma(a.v, b.v, c.v), a.id, b.v, b.id, a.v, c.id, 1.0);
}
inline Rev sqrt_(const Rev& a) {
  const double s = std::sqrt(a.v);
  return make_rev(s, a.id, 0.5 / s);
}
inline double value_of(const Rev& a) { return a.v; }

struct Fwd {
  double v = 0.0, d = 0.0;
  Fwd() = default;
  Fwd(double x) : v(x) {}  // NOLINT
  Fwd(double x, double dx) : v(x), d(dx) {}
};
inline Fwd operator+(const Fwd& a, const Fwd& b) { return {a.v + b.v, a.d + b.d}; }
inline Fwd operator-(const Fwd& a, const Fwd& b) { return {a.v - b.v, a.d - b.d}; }
inline Fwd operator*(const Fwd& a, const Fwd& b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Fwd operator/(const Fwd& a, const Fwd& b) {
  return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)};
}
inline Fwd operator-(const Fwd& a) { return {-a.v, -a.d}; }
inline Fwd fma_(const Fwd& a, const Fwd& b, const Fwd& c) {
  return {std::fma(a.v, b.v, c.v), a.d * b.v + a.v * b.d + c.d};
}
inline Fwd sqrt_(const Fwd& a) {
  const double s = std::sqrt(a.v);
  return {s, 0.5 * a.d / s};
}
inline double value_of(const Fwd& a) { return a.v; }

}  // namespace sxo
