// C API of the oracle's algorithmic derivatives (oracle/ad.hpp): exact
// vector-Jacobian products (reverse mode, tape) and Jacobian-vector products
// (forward mode) of the GH DG right-hand side and of its point-wise terms.
// TEST INFRASTRUCTURE ONLY.
#include <vector>

#include "ad.hpp"
#include "sxo.hpp"

namespace {

using sxo::Fwd;
using sxo::Rev;
using sxo::Tape;

std::vector<Rev> constants(const double* x, size_t n) { return std::vector<Rev>(x, x + n); }

// ubar = J^T gbar for y = f(x) (x: n_in values, y: n_out values).
template <class F>
void tape_vjp(size_t n_in, const double* x, size_t n_out, const double* gbar, double* xbar,
              F f) {
  Tape& t = Tape::get();
  t.clear();
  std::vector<Rev> X(n_in), Y(n_out);
  for (size_t i = 0; i < n_in; ++i) X[i] = Rev::input(x[i]);
  f(X.data(), Y.data());
  t.adj.assign(t.nodes.size(), 0.0);
  for (size_t o = 0; o < n_out; ++o)
    if (Y[o].id >= 0) t.adj[Y[o].id] += gbar[o];
  t.reverse();
  for (size_t i = 0; i < n_in; ++i) xbar[i] = t.adj[X[i].id];
  t.clear();
}

template <class F>
void forward_jvp(size_t n_in, const double* x, const double* v, size_t n_out, double* jv, F f) {
  std::vector<Fwd> X(n_in), Y(n_out);
  for (size_t i = 0; i < n_in; ++i) X[i] = Fwd(x[i], v[i]);
  f(X.data(), Y.data());
  for (size_t o = 0; o < n_out; ++o) jv[o] = Y[o].d;
}

size_t points(int num_elements, const int* ext) {
  return static_cast<size_t>(num_elements) * ext[0] * ext[1] * ext[2];
}

size_t ragged_points(int num_elements, const int* extents) {
  size_t n = 0;
  for (int e = 0; e < num_elements; ++e)
    n += static_cast<size_t>(extents[3 * e]) * extents[3 * e + 1] * extents[3 * e + 2];
  return n;
}

}  // namespace

extern "C" {

// GH DG RHS (uniform brick): ubar = (d dt / d u)^T gbar.
int sxo_gh_rhs_vjp(int num_elements, const int* ext, const int* nbr, const double* u,
                   const double* gammas, const double* inv_j, const double* gbar, double* ubar) {
  const size_t np = points(num_elements, ext);
  auto ga = constants(gammas, 3 * np);
  auto ij = constants(inv_j, 9 * np);
  tape_vjp(50 * np, u, 50 * np, gbar, ubar, [&](const Rev* x, Rev* y) {
    sxo::gh_rhs(num_elements, ext, nbr, x, ga.data(), ij.data(), y);
  });
  return 0;
}

int sxo_gh_rhs_jvp(int num_elements, const int* ext, const int* nbr, const double* u,
                   const double* gammas, const double* inv_j, const double* v, double* jv) {
  const size_t np = points(num_elements, ext);
  std::vector<Fwd> ga(gammas, gammas + 3 * np), ij(inv_j, inv_j + 9 * np);
  forward_jvp(50 * np, u, v, 50 * np, jv, [&](const Fwd* x, Fwd* y) {
    sxo::gh_rhs(num_elements, ext, nbr, x, ga.data(), ij.data(), y);
  });
  return 0;
}

int sxo_gh_rhs_mixed_vjp(int num_elements, const int* extents, const int* nbr, const double* u,
                         const double* gammas, const double* inv_j, const double* gbar,
                         double* ubar) {
  const size_t np = ragged_points(num_elements, extents);
  auto ga = constants(gammas, 3 * np);
  auto ij = constants(inv_j, 9 * np);
  tape_vjp(50 * np, u, 50 * np, gbar, ubar, [&](const Rev* x, Rev* y) {
    sxo::gh_rhs_mixed(num_elements, extents, nbr, x, ga.data(), ij.data(), y);
  });
  return 0;
}

int sxo_gh_rhs_mixed_jvp(int num_elements, const int* extents, const int* nbr, const double* u,
                         const double* gammas, const double* inv_j, const double* v, double* jv) {
  const size_t np = ragged_points(num_elements, extents);
  std::vector<Fwd> ga(gammas, gammas + 3 * np), ij(inv_j, inv_j + 9 * np);
  forward_jvp(50 * np, u, v, 50 * np, jv, [&](const Fwd* x, Fwd* y) {
    sxo::gh_rhs_mixed(num_elements, extents, nbr, x, ga.data(), ij.data(), y);
  });
  return 0;
}

// Point-wise GH time derivative: (ubar [50][n], dubar [150][n]) =
// (d dt / d (u, du))^T gbar, point by point.
int sxo_gh_time_derivative_vjp(int npts, const double* u, const double* du, const double* gammas,
                               const double* gbar, double* ubar, double* dubar) {
  for (int p = 0; p < npts; ++p) {
    double x[200], gb[50], xb[200];
    for (int c = 0; c < 50; ++c) x[c] = u[c * npts + p];
    for (int c = 0; c < 150; ++c) x[50 + c] = du[c * npts + p];
    for (int c = 0; c < 50; ++c) gb[c] = gbar[c * npts + p];
    const Rev g0(gammas[p]), g1(gammas[npts + p]), g2(gammas[2 * npts + p]);
    tape_vjp(200, x, 50, gb, xb, [&](const Rev* in, Rev* out) {
      sxo::GhFaceTemps<Rev> t;
      sxo::gh_time_derivative_point(out, &t, in, in + 50, g0, g1, g2);
    });
    for (int c = 0; c < 50; ++c) ubar[c * npts + p] = xb[c];
    for (int c = 0; c < 150; ++c) dubar[c * npts + p] = xb[50 + c];
  }
  return 0;
}

// GH evolution through `steps` Rk3HesthavenSsp steps (uniform brick):
// ubar0 = (d u_final / d u0)^T gbar_final, by the tape over the whole
// evolution (the transpose of the assembled CPU operator, applied to one
// vector).
int sxo_gh_evolve_vjp(int num_elements, const int* ext, const int* nbr, const double* u0,
                      const double* gammas, const double* inv_j, double dt, int steps,
                      const double* gbar_final, double* ubar0) {
  const size_t n = 50 * points(num_elements, ext);
  auto ga = constants(gammas, 3 * points(num_elements, ext));
  auto ij = constants(inv_j, 9 * points(num_elements, ext));
  tape_vjp(n, u0, n, gbar_final, ubar0, [&](const Rev* x, Rev* y) {
    std::vector<Rev> u(x, x + n), u1(n), u2(n), l(n);
    for (int s = 0; s < steps; ++s) {
      sxo::gh_rhs(num_elements, ext, nbr, u.data(), ga.data(), ij.data(), l.data());
      sxo::rk3_substep(0, n, dt, u.data(), u.data(), l.data(), u1.data());
      sxo::gh_rhs(num_elements, ext, nbr, u1.data(), ga.data(), ij.data(), l.data());
      sxo::rk3_substep(1, n, dt, u.data(), u1.data(), l.data(), u2.data());
      sxo::gh_rhs(num_elements, ext, nbr, u2.data(), ga.data(), ij.data(), l.data());
      sxo::rk3_substep(2, n, dt, u.data(), u2.data(), l.data(), u1.data());
      u.swap(u1);
    }
    for (size_t i = 0; i < n; ++i) y[i] = u[i];
  });
  return 0;
}

// The same on a mixed-p set of elements (ragged layout).
int sxo_gh_evolve_mixed_vjp(int num_elements, const int* extents, const int* nbr,
                            const double* u0, const double* gammas, const double* inv_j,
                            double dt, int steps, const double* gbar_final, double* ubar0) {
  const size_t np = ragged_points(num_elements, extents);
  const size_t n = 50 * np;
  auto ga = constants(gammas, 3 * np);
  auto ij = constants(inv_j, 9 * np);
  tape_vjp(n, u0, n, gbar_final, ubar0, [&](const Rev* x, Rev* y) {
    std::vector<Rev> u(x, x + n), u1(n), u2(n), l(n);
    for (int s = 0; s < steps; ++s) {
      sxo::gh_rhs_mixed(num_elements, extents, nbr, u.data(), ga.data(), ij.data(), l.data());
      sxo::rk3_substep(0, n, dt, u.data(), u.data(), l.data(), u1.data());
      sxo::gh_rhs_mixed(num_elements, extents, nbr, u1.data(), ga.data(), ij.data(), l.data());
      sxo::rk3_substep(1, n, dt, u.data(), u1.data(), l.data(), u2.data());
      sxo::gh_rhs_mixed(num_elements, extents, nbr, u2.data(), ga.data(), ij.data(), l.data());
      sxo::rk3_substep(2, n, dt, u.data(), u2.data(), l.data(), u1.data());
      u.swap(u1);
    }
    for (size_t i = 0; i < n; ++i) y[i] = u[i];
  });
  return 0;
}

// ScalarWave DG RHS by the tape (cross-check of the hand-written adjoint).
int sxo_sw_rhs_vjp(int num_elements, const int* ext, const int* nbr, const double* u,
                   const double* gamma2, const double* inv_j, const double* gbar, double* ubar) {
  const size_t np = points(num_elements, ext);
  auto g2 = constants(gamma2, np);
  auto ij = constants(inv_j, 9 * np);
  tape_vjp(5 * np, u, 5 * np, gbar, ubar, [&](const Rev* x, Rev* y) {
    sxo::sw_rhs(num_elements, ext, nbr, x, g2.data(), ij.data(), y);
  });
  return 0;
}

}  // extern "C"
