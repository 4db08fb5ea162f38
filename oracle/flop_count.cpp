// Exact floating-point operation counts of the DG right-hand sides, obtained
// by running the oracle's templated restatement of the reference path with a
// counting scalar type.  Test and measurement infrastructure (the counts are
// the algorithmic flops of the roofline); not part of the product.
//
// Counting rules: +, -, *, / and sqrt are one flop each, a fused multiply-add
// is two, negation and copies are free.  One element with every face
// internal (the element is its own neighbour, as in a periodic brick) gives
// the per-element count of a periodic brick of equal extents.
//
// Usage: flop_count gh|sw N0 N1 N2   ->  one JSON object on stdout
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace sxo {
struct Counted;
}
namespace sxo_count {
inline long long g_add = 0, g_mul = 0, g_div = 0, g_fma = 0, g_sqrt = 0;
}

namespace sxo {
struct Counted {
  double v = 0.0;
  Counted() = default;
  Counted(double x) : v(x) {}  // NOLINT: implicit from constants, as double
};
inline Counted operator+(Counted a, Counted b) { ++sxo_count::g_add; return a.v + b.v; }
inline Counted operator-(Counted a, Counted b) { ++sxo_count::g_add; return a.v - b.v; }
inline Counted operator*(Counted a, Counted b) { ++sxo_count::g_mul; return a.v * b.v; }
inline Counted operator/(Counted a, Counted b) { ++sxo_count::g_div; return a.v / b.v; }
inline Counted operator-(Counted a) { return -a.v; }
inline Counted& operator+=(Counted& a, Counted b) { a = a + b; return a; }
inline Counted& operator-=(Counted& a, Counted b) { a = a - b; return a; }
inline Counted& operator*=(Counted& a, Counted b) { a = a * b; return a; }
inline Counted fma_(Counted a, Counted b, Counted c) {
  ++sxo_count::g_fma;
  return std::fma(a.v, b.v, c.v);
}
inline Counted sqrt_(Counted a) { ++sxo_count::g_sqrt; return std::sqrt(a.v); }
inline double value_of(Counted a) { return a.v; }
}  // namespace sxo

#include "sxo.hpp"

using sxo::Counted;

static void reset() {
  sxo_count::g_add = sxo_count::g_mul = sxo_count::g_div = sxo_count::g_fma = sxo_count::g_sqrt = 0;
}

static long long flops() {
  return sxo_count::g_add + sxo_count::g_mul + sxo_count::g_div + 2 * sxo_count::g_fma +
         sxo_count::g_sqrt;
}

int main(int argc, char** argv) {
  if (argc != 5 || (std::strcmp(argv[1], "gh") != 0 && std::strcmp(argv[1], "sw") != 0)) {
    std::fprintf(stderr, "usage: %s gh|sw N0 N1 N2\n", argv[0]);
    return 2;
  }
  const bool gh = std::strcmp(argv[1], "gh") == 0;
  const int ext[3] = {std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4])};
  for (int d = 0; d < 3; ++d)
    if (ext[d] < 2 || ext[d] > 20) return 2;
  const int np = ext[0] * ext[1] * ext[2];
  const int nc = gh ? 50 : 5;
  const int naux = gh ? 3 : 1;
  const int nbr[6] = {0, 0, 0, 0, 0, 0};
  std::vector<Counted> u(static_cast<size_t>(nc) * np), aux(static_cast<size_t>(naux) * np),
      inv(9 * static_cast<size_t>(np)), dt(static_cast<size_t>(nc) * np);
  // a Minkowski-like state (positive lapse, positive-definite spatial metric)
  for (int p = 0; p < np; ++p) {
    const double w = 1e-3 * std::sin(0.37 * p);
    for (int c = 0; c < nc; ++c) u[static_cast<size_t>(c) * np + p] = 0.01 * std::cos(c + 0.1 * p);
    if (gh) {
      const double diag[4] = {-1.0, 1.0, 1.0, 1.0};
      for (int a = 0; a < 4; ++a) u[static_cast<size_t>(sxo::sym4(a, a)) * np + p] = diag[a] + w;
    }
    for (int c = 0; c < naux; ++c) aux[static_cast<size_t>(c) * np + p] = 1.0;
    for (int c = 0; c < 9; ++c) inv[static_cast<size_t>(c) * np + p] = (c % 4 == 0) ? 2.0 : 0.0;
  }
  reset();
  if (gh)
    sxo::gh_rhs<Counted>(1, ext, nbr, u.data(), aux.data(), inv.data(), dt.data());
  else
    sxo::sw_rhs<Counted>(1, ext, nbr, u.data(), aux.data(), inv.data(), dt.data());
  std::printf(
      "{\"system\": \"%s\", \"extents\": [%d, %d, %d], \"grid_points\": %d, "
      "\"add_sub\": %lld, \"mul\": %lld, \"div\": %lld, \"fma\": %lld, \"sqrt\": %lld, "
      "\"flops_per_element\": %lld, \"flops_per_point\": %.4f}\n",
      argv[1], ext[0], ext[1], ext[2], np, sxo_count::g_add, sxo_count::g_mul, sxo_count::g_div,
      sxo_count::g_fma, sxo_count::g_sqrt, flops(), static_cast<double>(flops()) / np);
  return 0;
}
