// Oracle C entry points for the Python tests (ctypes).  TEST INFRASTRUCTURE
// ONLY: see the header of sxo.hpp.  All arrays are host arrays in SpECTRE's
// Variables layout (component-major, xi fastest); return 0 on success and a
// negative value on an invalid argument.
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "gh_reference_impl.hpp"
#include "sxo.hpp"
#include "sxo_refstyle.hpp"
#include "sxo_flux.hpp"
#include "sxo_gauss.hpp"
#include "sxo_mortar.hpp"

namespace {
bool bad_extents(const int* ext) {
  for (int d = 0; d < 3; ++d)
    if (ext[d] < 2 || ext[d] > 20) return true;
  return false;
}

template <class F>
void run_split(int threads, int num_elements, F&& f) {
  if (threads < 1) threads = 1;
  if (threads > num_elements) threads = num_elements;
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) {
    const int b = static_cast<int>(static_cast<long>(num_elements) * t / threads);
    const int e = static_cast<int>(static_cast<long>(num_elements) * (t + 1) / threads);
    pool.emplace_back([&f, b, e] { f(b, e); });
  }
  for (auto& th : pool) th.join();
}
}  // namespace

extern "C" {

int sxo_lgl_points_and_weights(int n, double* x, double* w) {
  if (n < 2 || n > 20) return -1;
  std::vector<double> xv, wv;
  sxo::lgl_points_and_weights(n, &xv, &wv);
  std::memcpy(x, xv.data(), n * sizeof(double));
  std::memcpy(w, wv.data(), n * sizeof(double));
  return 0;
}

int sxo_differentiation_matrix(int n, double* d) {
  if (n < 2 || n > 20) return -1;
  const auto& m = sxo::SpectralCache::get().D[n];
  std::memcpy(d, m.data(), m.size() * sizeof(double));
  return 0;
}

int sxo_lg_points_and_weights(int n, double* x, double* w) {
  if (n < 2 || n > 20) return -1;
  std::vector<double> xv, wv;
  sxo::lg_points_and_weights(n, &xv, &wv);
  std::memcpy(x, xv.data(), n * sizeof(double));
  std::memcpy(w, wv.data(), n * sizeof(double));
  return 0;
}

int sxo_lg_differentiation_matrix(int n, double* d) {
  if (n < 2 || n > 20) return -1;
  const std::vector<double> m = sxo::lg_differentiation_matrix(n);
  std::memcpy(d, m.data(), m.size() * sizeof(double));
  return 0;
}

// out [4][n]: interpolation rows to -1 and +1, lifting terms at -1 and +1.
int sxo_lg_boundary_terms(int n, double* out) {
  if (n < 2 || n > 20) return -1;
  const std::vector<double> m = sxo::lg_boundary_terms(n);
  std::memcpy(out, m.data(), m.size() * sizeof(double));
  return 0;
}

int sxo_sw_rhs_gauss(int num_elements, const int* ext, const int* nbr, const double* u,
                     const double* gamma2, const double* inv_j, const double* det_inv_j,
                     const double* face_inv_j, double* dt) {
  if (bad_extents(ext) || num_elements < 1) return -1;
  for (int i = 0; i < 6 * num_elements; ++i)
    if (nbr[i] < 0 || nbr[i] >= num_elements) return -1;
  sxo::sw_rhs_gauss(num_elements, ext, nbr, u, gamma2, inv_j, det_inv_j, face_inv_j, dt);
  return 0;
}

int sxo_weak_flux_differentiation_matrix(int n, double* w) {
  if (n < 2 || n > 20) return -1;
  const std::vector<double> m = sxo::weak_flux_differentiation_matrix(n);
  std::memcpy(w, m.data(), m.size() * sizeof(double));
  return 0;
}

// One element: fluxes [3 nc][Np], inv_j / djij [9][Np] -> div [nc][Np].
int sxo_divergence(const double* F, int nc, const int* ext, const double* inv_j, double* div) {
  if (bad_extents(ext) || nc < 1) return -1;
  sxo::divergence(F, nc, ext, inv_j, div);
  return 0;
}

int sxo_weak_divergence(const double* F, int nc, const int* ext, const double* djij,
                        double* div) {
  if (bad_extents(ext) || nc < 1) return -1;
  sxo::weak_divergence(F, nc, ext, djij, div);
  return 0;
}

// One element; F updated in place, dt [nc][Np] updated in place.  v / div_v
// may be null (static mesh); weak != 0 selects the weak form.
int sxo_flux_volume_terms(int nc, const int* ext, const double* u, double* F, const double* v,
                          const double* div_v, int weak, const double* inv_j,
                          const double* djij, const double* det_inv_jac, double* dt) {
  if (bad_extents(ext) || nc < 1) return -1;
  sxo::flux_volume_terms(nc, ext, u, F, v, div_v, weak != 0, inv_j, djij, det_inv_jac, dt);
  return 0;
}

int sxo_mesh_velocity_nonconservative(long n, int nc, const double* du, const double* v,
                                      double* dt) {
  if (n < 0 || nc < 1) return -1;
  sxo::mesh_velocity_nonconservative(static_cast<size_t>(n), nc, du, v, dt);
  return 0;
}

int sxo_logical_partial_derivatives(const double* u, int nc, const int* ext,
                                    double* dlogical) {
  if (bad_extents(ext) || nc < 1) return -1;
  const size_t np = static_cast<size_t>(ext[0]) * ext[1] * ext[2];
  sxo::logical_partial_derivatives(u, nc, ext, dlogical, dlogical + nc * np,
                                   dlogical + 2 * nc * np);
  return 0;
}

int sxo_partial_derivatives(const double* u, int nc, const int* ext,
                            const double* inv_j, double* du) {
  if (bad_extents(ext) || nc < 1) return -1;
  sxo::partial_derivatives(u, nc, ext, inv_j, du);
  return 0;
}

// Transpose of the partial derivatives (one element): dubar [3 nc][Np] ->
// ubar [nc][Np].
int sxo_partial_derivatives_adjoint(const double* dubar, int nc, const int* ext,
                                    const double* inv_j, double* ubar) {
  if (bad_extents(ext) || nc < 1) return -1;
  sxo::partial_derivatives_adjoint(dubar, nc, ext, inv_j, ubar);
  return 0;
}

// Batched partial derivatives over num_elements elements of equal extents.
int sxo_partial_derivatives_batch(int num_elements, const double* u, int nc,
                                  const int* ext, const double* inv_j,
                                  double* du) {
  if (bad_extents(ext) || nc < 1 || num_elements < 0) return -1;
  const size_t np = static_cast<size_t>(ext[0]) * ext[1] * ext[2];
  for (int e = 0; e < num_elements; ++e)
    sxo::partial_derivatives(u + e * nc * np, nc, ext, inv_j + e * 9 * np,
                             du + e * 3 * nc * np);
  return 0;
}

// Pointwise ScalarWave TimeDerivative over npts points.
int sxo_sw_time_derivative(int npts, const double* u, const double* du,
                           const double* gamma2, double* dt) {
  for (int p = 0; p < npts; ++p) {
    double up[5], dup[15], dtp[5];
    for (int c = 0; c < 5; ++c) up[c] = u[c * npts + p];
    for (int c = 0; c < 15; ++c) dup[c] = du[c * npts + p];
    sxo::sw_time_derivative_point(dtp, up, dup, gamma2[p]);
    for (int c = 0; c < 5; ++c) dt[c * npts + p] = dtp[c];
  }
  return 0;
}

// Pointwise GH TimeDerivative over npts points.  gammas [3][npts];
// temps (may be NULL) [12][npts] = gamma1, gamma2, lapse, shift^i,
// inverse spatial metric (6).
int sxo_gh_time_derivative(int npts, const double* u, const double* du,
                           const double* gammas, double* dt, double* temps) {
  for (int p = 0; p < npts; ++p) {
    double up[50], dup[150], dtp[50];
    for (int c = 0; c < 50; ++c) up[c] = u[c * npts + p];
    for (int c = 0; c < 150; ++c) dup[c] = du[c * npts + p];
    sxo::GhFaceTemps<double> t;
    sxo::gh_time_derivative_point(dtp, &t, up, dup, gammas[p], gammas[npts + p],
                                  gammas[2 * npts + p]);
    for (int c = 0; c < 50; ++c) dt[c * npts + p] = dtp[c];
    if (temps != nullptr) {
      temps[0 * npts + p] = t.gamma1;
      temps[1 * npts + p] = t.gamma2;
      temps[2 * npts + p] = t.lapse;
      for (int i = 0; i < 3; ++i) temps[(3 + i) * npts + p] = t.shift[i];
      for (int i = 0; i < 6; ++i) temps[(6 + i) * npts + p] = t.inv_spatial[i];
    }
  }
  return 0;
}

// The same points through the restated Test_DuDt reference implementation,
// with its auxiliary inputs computed from u by the GR helper functions, as in
// Test_DuDt.cpp:521-560 (harmonic gauge: H = 0, dH = 0).
int sxo_gh_reference_impl_from_state(int npts, const double* u, const double* du,
                                     const double* gammas, double* dt) {
  for (int p = 0; p < npts; ++p) {
    sxo::GhReferenceInputs in{};
    for (int c = 0; c < 10; ++c) {
      in.psi[c] = u[c * npts + p];
      in.pi[c] = u[(10 + c) * npts + p];
    }
    for (int c = 0; c < 30; ++c) in.phi[c] = u[(20 + c) * npts + p];
    for (int c = 0; c < 30; ++c) {
      in.d_psi[c] = du[c * npts + p];
      in.d_pi[c] = du[(30 + c) * npts + p];
    }
    for (int c = 0; c < 90; ++c) in.d_phi[c] = du[(60 + c) * npts + p];
    in.gamma0 = gammas[p];
    in.gamma1 = gammas[npts + p];
    in.gamma2 = gammas[2 * npts + p];
    double det;
    sxo::gh_lapse_shift(&in.lapse, in.shift, &det, in.inv_spatial, in.psi);
    double m1 = -1.0 / (in.lapse * in.lapse);
    in.inv_psi[sxo::sym4(0, 0)] = m1;
    for (int i = 0; i < 3; ++i) in.inv_psi[sxo::sym4(0, i + 1)] = -in.shift[i] * m1;
    for (int i = 0; i < 3; ++i)
      for (int j = i; j < 3; ++j)
        in.inv_psi[sxo::sym4(i + 1, j + 1)] =
            in.inv_spatial[sxo::sym3(i, j)] + in.shift[i] * in.shift[j] * m1;
    // spacetime derivative of g: index 0 = -lapse Pi + shift^m Phi_m (the
    // harmonic-gauge value of dt g used for the Christoffel symbols)
    double da[40];
    for (int a = 0; a < 4; ++a)
      for (int b = a; b < 4; ++b) {
        double s = -in.lapse * in.pi[sxo::sym4(a, b)];
        for (int m = 0; m < 3; ++m) s += in.shift[m] * in.phi[sxo::iaa(m, a, b)];
        da[sxo::abb(0, a, b)] = s;
        for (int i = 0; i < 3; ++i) da[sxo::abb(i + 1, a, b)] = in.phi[sxo::iaa(i, a, b)];
      }
    for (int k = 0; k < 4; ++k)
      for (int i = 0; i < 4; ++i)
        for (int j = i; j < 4; ++j)
          in.chr1[sxo::abb(k, i, j)] =
              0.5 * (da[sxo::abb(i, j, k)] + da[sxo::abb(j, i, k)] - da[sxo::abb(k, i, j)]);
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b)
        for (int c = b; c < 4; ++c) {
          double s = 0.0;
          for (int d = 0; d < 4; ++d)
            s += in.inv_psi[sxo::sym4(a, d)] * in.chr1[sxo::abb(d, b, c)];
          in.chr2[sxo::abb(a, b, c)] = s;
        }
    for (int a = 0; a < 4; ++a) {
      double s = 0.0;
      for (int b = 0; b < 4; ++b)
        for (int c = 0; c < 4; ++c)
          s += in.inv_psi[sxo::sym4(b, c)] * in.chr1[sxo::abb(a, b, c)];
      in.trace_chr[a] = s;
    }
    in.normal_vector[0] = 1.0 / in.lapse;
    for (int i = 0; i < 3; ++i) in.normal_vector[i + 1] = -in.shift[i] / in.lapse;
    in.normal_one_form[0] = -in.lapse;
    for (int i = 1; i < 4; ++i) in.normal_one_form[i] = 0.0;
    double dtp[50];
    sxo::gh_reference_impl_point(dtp, in);
    for (int c = 0; c < 50; ++c) dt[c * npts + p] = dtp[c];
  }
  return 0;
}

// Test_DuDt.cpp:294-359: random inputs from std::mt19937 gen(1.) with
// uniform_real_distribution<>(-10, 10), 2 points, tensors filled component by
// component in storage order; then the reference implementation.
// dt_out: [50][2].
int sxo_gh_spec_golden(double* dt_out) {
  const int n_pts = 2;
  std::mt19937 gen(1.);
  auto fill = [&](double* dst, int ncomp, std::vector<std::vector<double>>* keep) {
    std::vector<double> buf(ncomp * n_pts);
    for (int c = 0; c < ncomp; ++c)
      for (int s = 0; s < n_pts; ++s)
        buf[c * n_pts + s] = std::uniform_real_distribution<>(-10, 10)(gen);
    keep->push_back(buf);
    (void)dst;
  };
  std::vector<std::vector<double>> t;
  // order of Test_DuDt.cpp:313-352
  const int sizes[] = {10, 10, 30, 30, 30, 90, 4, 16, 1, 1, 1, 1,
                       3,  6,  10, 40, 40, 4, 4, 4};
  for (int s : sizes) fill(nullptr, s, &t);
  for (int p = 0; p < n_pts; ++p) {
    sxo::GhReferenceInputs in{};
    auto take = [&](int idx, double* dst, int ncomp) {
      for (int c = 0; c < ncomp; ++c) dst[c] = t[idx][c * n_pts + p];
    };
    take(0, in.psi, 10);
    take(1, in.pi, 10);
    take(2, in.phi, 30);
    take(3, in.d_psi, 30);
    take(4, in.d_pi, 30);
    take(5, in.d_phi, 90);
    take(6, in.gauge_function, 4);
    take(7, in.d_gauge_function, 16);
    take(8, &in.gamma0, 1);
    take(9, &in.gamma1, 1);
    take(10, &in.gamma2, 1);
    take(11, &in.lapse, 1);
    take(12, in.shift, 3);
    take(13, in.inv_spatial, 6);
    take(14, in.inv_psi, 10);
    take(15, in.chr1, 40);
    take(16, in.chr2, 40);
    take(17, in.trace_chr, 4);
    take(18, in.normal_one_form, 4);
    take(19, in.normal_vector, 4);
    double dtp[50];
    sxo::gh_reference_impl_point(dtp, in);
    for (int c = 0; c < 50; ++c) dt_out[c * n_pts + p] = dtp[c];
  }
  return 0;
}

// ScalarWave UpwindPenalty on face points: package both sides then boundary
// terms.  u_int/u_ext [5][npts], gamma2 [npts], n_int/n_ext [3][npts] unit
// covectors.  bc [5][npts].
int sxo_sw_upwind(int npts, const double* u_int, const double* g2_int,
                  const double* n_int, const double* u_ext,
                  const double* g2_ext, const double* n_ext, double* bc) {
  for (int p = 0; p < npts; ++p) {
    double ui[5], ue[5], ni[3], ne[3], pi_[16], pe[16], b[5];
    for (int c = 0; c < 5; ++c) { ui[c] = u_int[c * npts + p]; ue[c] = u_ext[c * npts + p]; }
    for (int i = 0; i < 3; ++i) { ni[i] = n_int[i * npts + p]; ne[i] = n_ext[i * npts + p]; }
    sxo::sw_package_point(pi_, ui, g2_int[p], ni);
    sxo::sw_package_point(pe, ue, g2_ext[p], ne);
    sxo::sw_boundary_terms_point(b, pi_, pe);
    for (int c = 0; c < 5; ++c) bc[c * npts + p] = b[c];
  }
  return 0;
}

// GH UpwindPenalty on face points.  u_* [50][npts]; temps_* [12][npts] as in
// sxo_gh_time_derivative; n_* [3][npts] unnormalized covectors (normalized
// here with the inverse spatial metric).  bc [50][npts].
int sxo_gh_upwind(int npts, const double* u_int, const double* t_int,
                  const double* n_int, const double* u_ext, const double* t_ext,
                  const double* n_ext, double* bc) {
  for (int p = 0; p < npts; ++p) {
    double isi[6], ise[6], shi[3], she[3], ni[3], ne[3], nvi[3], nve[3], mi, me;
    for (int i = 0; i < 6; ++i) { isi[i] = t_int[(6 + i) * npts + p]; ise[i] = t_ext[(6 + i) * npts + p]; }
    for (int i = 0; i < 3; ++i) {
      shi[i] = t_int[(3 + i) * npts + p]; she[i] = t_ext[(3 + i) * npts + p];
      ni[i] = n_int[i * npts + p]; ne[i] = n_ext[i * npts + p];
    }
    sxo::gh_normalize(ni, nvi, &mi, isi);
    sxo::gh_normalize(ne, nve, &me, ise);
    double li[4], le[4];
    sxo::gh_char_speeds(li, t_int[p], t_int[2 * npts + p], shi, ni);
    sxo::gh_char_speeds(le, t_ext[p], t_ext[2 * npts + p], she, ne);
    for (int pr = 0; pr < 10; ++pr) {
      double phii[3], phie[3], pki[13], pke[13], b[5];
      for (int i = 0; i < 3; ++i) {
        phii[i] = u_int[(20 + i + 3 * pr) * npts + p];
        phie[i] = u_ext[(20 + i + 3 * pr) * npts + p];
      }
      sxo::gh_package_pair(pki, u_int[pr * npts + p], u_int[(10 + pr) * npts + p],
                           phii, t_int[npts + p], li, ni, nvi);
      sxo::gh_package_pair(pke, u_ext[pr * npts + p], u_ext[(10 + pr) * npts + p],
                           phie, t_ext[npts + p], le, ne, nve);
      sxo::gh_boundary_terms_pair(b, pki, li, pke, le);
      const int comp[5] = {pr, 10 + pr, 20 + 3 * pr, 21 + 3 * pr, 22 + 3 * pr};
      for (int c = 0; c < 5; ++c) bc[comp[c] * npts + p] = b[c];
    }
  }
  return 0;
}

// GH UpwindPenalty with the face quantities given directly (unit normal
// covector and vector, lapse, shift, gamma1, gamma2 per side), as the pypp
// oracle of the reference takes them.  side arrays: u [50][n], aux [9][n] =
// gamma1, gamma2, lapse, shift (3), ... normals [6][n] = n_i (3), n^i (3).
int sxo_gh_upwind_direct(int npts, const double* u_int, const double* aux_int,
                         const double* nrm_int, const double* u_ext, const double* aux_ext,
                         const double* nrm_ext, double* bc) {
  for (int p = 0; p < npts; ++p) {
    double ni[3], nvi[3], ne[3], nve[3], shi[3], she[3];
    for (int i = 0; i < 3; ++i) {
      ni[i] = nrm_int[i * npts + p];
      nvi[i] = nrm_int[(3 + i) * npts + p];
      ne[i] = nrm_ext[i * npts + p];
      nve[i] = nrm_ext[(3 + i) * npts + p];
      shi[i] = aux_int[(3 + i) * npts + p];
      she[i] = aux_ext[(3 + i) * npts + p];
    }
    double li[4], le[4];
    sxo::gh_char_speeds(li, aux_int[p], aux_int[2 * npts + p], shi, ni);
    sxo::gh_char_speeds(le, aux_ext[p], aux_ext[2 * npts + p], she, ne);
    for (int pr = 0; pr < 10; ++pr) {
      double phii[3], phie[3], pki[13], pke[13], b[5];
      for (int i = 0; i < 3; ++i) {
        phii[i] = u_int[(20 + i + 3 * pr) * npts + p];
        phie[i] = u_ext[(20 + i + 3 * pr) * npts + p];
      }
      sxo::gh_package_pair(pki, u_int[pr * npts + p], u_int[(10 + pr) * npts + p], phii,
                           aux_int[npts + p], li, ni, nvi);
      sxo::gh_package_pair(pke, u_ext[pr * npts + p], u_ext[(10 + pr) * npts + p], phie,
                           aux_ext[npts + p], le, ne, nve);
      sxo::gh_boundary_terms_pair(b, pki, li, pke, le);
      const int comp[5] = {pr, 10 + pr, 20 + 3 * pr, 21 + 3 * pr, 22 + 3 * pr};
      for (int c = 0; c < 5; ++c) bc[comp[c] * npts + p] = b[c];
    }
  }
  return 0;
}

// Full DG RHS on a uniform periodic brick.  nbr [6 E].
int sxo_sw_rhs(int num_elements, const int* ext, const int* nbr, const double* u,
               const double* gamma2, const double* inv_j, double* dt) {
  if (bad_extents(ext) || num_elements < 1) return -1;
  sxo::sw_rhs(num_elements, ext, nbr, u, gamma2, inv_j, dt);
  return 0;
}

// ScalarWave RHS with external faces (neighbour < 0) closed by the ghost
// state `ghost` [E][5][Np] (Dirichlet analytic boundary condition).
int sxo_sw_rhs_dirichlet(int num_elements, const int* ext, const int* nbr, const double* u,
                         const double* gamma2, const double* inv_j, const double* ghost,
                         double* dt) {
  if (bad_extents(ext) || num_elements < 1 || ghost == nullptr) return -1;
  sxo::sw_rhs(num_elements, ext, nbr, u, gamma2, inv_j, dt, ghost);
  return 0;
}

// Generalized-harmonic RHS with external faces (neighbour < 0) closed by the
// ghost state `ghost` [E][50][Np] (Dirichlet analytic boundary condition).
// ScalarWave RHS with external faces closed by the SphericalRadiation ghost
// condition: kind 1 Sommerfeld, 2 Bayliss-Turkel (coords [E][3][Np]).
int sxo_sw_rhs_spherical_radiation(int num_elements, const int* ext, const int* nbr,
                                   const double* u, const double* gamma2, const double* inv_j,
                                   const double* coords, int kind, double* dt) {
  if (bad_extents(ext) || num_elements < 1) return -1;
  if (kind != sxo::kSwSommerfeld && kind != sxo::kSwBaylissTurkel) return -1;
  if (kind == sxo::kSwBaylissTurkel && coords == nullptr) return -1;
  sxo::sw_rhs(num_elements, ext, nbr, u, gamma2, inv_j, dt, coords, kind);
  return 0;
}

// ScalarWave RHS with external faces closed by the constraint-preserving
// spherical radiation condition: kind 1 Sommerfeld, 2 first-order and
// 3 second-order Bayliss-Turkel (coords [E][3][Np], not read for Sommerfeld).
int sxo_sw_rhs_cpsr(int num_elements, const int* ext, const int* nbr, const double* u,
                    const double* gamma2, const double* inv_j, const double* coords, int kind,
                    double* dt) {
  if (bad_extents(ext) || num_elements < 1 || kind < 1 || kind > 3) return -1;
  if (kind > 1 && coords == nullptr) return -1;
  sxo::sw_rhs(num_elements, ext, nbr, u, gamma2, inv_j, dt, coords,
              sxo::kSwCpsrSommerfeld + kind - 1);
  return 0;
}

// dt Pi correction of the constraint-preserving spherical radiation condition
// at n points: u [5][n], du [15][n], normal [3][n], gamma2 [n], coords [3][n].
int sxo_sw_cpsr_correction(int n, int kind, const double* u, const double* du,
                           const double* normal, const double* gamma2, const double* coords,
                           double* out) {
  if (n < 0 || kind < 1 || kind > 3) return -1;
  for (int p = 0; p < n; ++p) {
    double up[5], dup[15], nn[3];
    for (int c = 0; c < 5; ++c) up[c] = u[c * n + p];
    for (int c = 0; c < 15; ++c) dup[c] = du[c * n + p];
    for (int i = 0; i < 3; ++i) nn[i] = normal[i * n + p];
    out[p] = sxo::sw_cpsr_correction_point(sxo::kSwCpsrSommerfeld + kind - 1, up, dup, nn,
                                           gamma2[p], coords + p, static_cast<size_t>(n));
  }
  return 0;
}

int sxo_gh_rhs_dirichlet(int num_elements, const int* ext, const int* nbr, const double* u,
                         const double* gammas, const double* inv_j, const double* ghost,
                         double* dt) {
  if (bad_extents(ext) || num_elements < 1 || ghost == nullptr) return -1;
  sxo::gh_rhs(num_elements, ext, nbr, u, gammas, inv_j, dt, ghost);
  return 0;
}

int sxo_gh_rhs(int num_elements, const int* ext, const int* nbr, const double* u,
               const double* gammas, const double* inv_j, double* dt) {
  if (bad_extents(ext) || num_elements < 1) return -1;
  sxo::gh_rhs(num_elements, ext, nbr, u, gammas, inv_j, dt);
  return 0;
}

// Volume terms and the internal faces' corrections only: the part of the GH
// RHS that a time-derivative condition on the external faces (Bjorhus)
// leaves to the upwind scheme.
int sxo_gh_rhs_internal_faces(int num_elements, const int* ext, const int* nbr, const double* u,
                              const double* gammas, const double* inv_j, double* dt) {
  if (bad_extents(ext) || num_elements < 1) return -1;
  const size_t np = static_cast<size_t>(ext[0]) * ext[1] * ext[2];
  std::vector<sxo::GhFaceTemps<double>> temps(static_cast<size_t>(num_elements) * np);
  sxo::gh_rhs_volume(0, num_elements, ext, u, gammas, inv_j, dt, temps.data());
  sxo::gh_rhs_faces(0, num_elements, ext, nbr, u, inv_j, dt, temps.data(),
                    static_cast<const double*>(nullptr), true);
  return 0;
}

int sxo_sw_rhs_adjoint(int num_elements, const int* ext, const int* nbr, const double* gbar,
                       const double* gamma2, const double* inv_j, double* ubar) {
  if (bad_extents(ext) || num_elements < 1) return -1;
  sxo::sw_rhs_adjoint(num_elements, ext, nbr, gbar, gamma2, inv_j, ubar);
  return 0;
}

// ---- mixed p ---------------------------------------------------------------

int sxo_mortar_interpolation_matrix(int n_face, int n_mortar, double* out) {
  if (n_face < 2 || n_mortar > 20 || n_mortar < n_face) return -1;
  const auto m = sxo::mortar_interpolation_matrix(n_face, n_mortar);
  std::memcpy(out, m.data(), m.size() * sizeof(double));
  return 0;
}

int sxo_mortar_projection_matrix(int n_face, int n_mortar, double* out) {
  if (n_face < 2 || n_mortar > 20 || n_mortar < n_face) return -1;
  const auto m = sxo::mortar_projection_matrix(n_face, n_mortar);
  std::memcpy(out, m.data(), m.size() * sizeof(double));
  return 0;
}

int sxo_lgl_interpolation_matrix(int n, int num_targets, const double* targets, double* out) {
  if (n < 2 || n > 20 || num_targets < 1) return -1;
  const auto m = sxo::lgl_interpolation_matrix(n, std::vector<double>(targets, targets + num_targets));
  std::memcpy(out, m.data(), m.size() * sizeof(double));
  return 0;
}

static bool bad_ragged(int num_elements, const int* extents, const int* nbr,
                       bool external = false) {
  if (num_elements < 1) return true;
  for (int e = 0; e < num_elements; ++e) {
    if (bad_extents(extents + 3 * e)) return true;
    for (int k = 0; k < 6; ++k)
      if (nbr[6 * e + k] < (external ? -1 : 0) || nbr[6 * e + k] >= num_elements) return true;
  }
  return false;
}

// Mixed-p ScalarWave RHS with external faces (neighbour -1) closed by the
// ghost state `ghost` (ragged, 5 components).
int sxo_sw_rhs_mixed_dirichlet(int num_elements, const int* extents, const int* nbr,
                               const double* u, const double* gamma2, const double* inv_j,
                               const double* ghost, double* dt) {
  if (bad_ragged(num_elements, extents, nbr, true) || ghost == nullptr) return -1;
  sxo::sw_rhs_mixed(num_elements, extents, nbr, u, gamma2, inv_j, dt, ghost);
  return 0;
}

// Mixed-p ScalarWave RHS with external faces closed by SphericalRadiation:
// kind 1 Sommerfeld, 2 Bayliss-Turkel (coords ragged [3][Np_e], not read for
// Sommerfeld).
int sxo_sw_rhs_mixed_spherical_radiation(int num_elements, const int* extents, const int* nbr,
                                         const double* u, const double* gamma2,
                                         const double* inv_j, const double* coords, int kind,
                                         double* dt) {
  if (bad_ragged(num_elements, extents, nbr, true)) return -1;
  if (kind != sxo::kSwSommerfeld && kind != sxo::kSwBaylissTurkel) return -1;
  if (kind == sxo::kSwBaylissTurkel && coords == nullptr) return -1;
  sxo::sw_rhs_mixed(num_elements, extents, nbr, u, gamma2, inv_j, dt, coords, kind);
  return 0;
}

// Mixed-p ScalarWave RHS with external faces closed by the constraint-
// preserving spherical radiation condition: kind 1 Sommerfeld, 2 first-order
// and 3 second-order Bayliss-Turkel (coords ragged [3][Np_e], not read for
// Sommerfeld).
int sxo_sw_rhs_mixed_cpsr(int num_elements, const int* extents, const int* nbr,
                          const double* u, const double* gamma2, const double* inv_j,
                          const double* coords, int kind, double* dt) {
  if (bad_ragged(num_elements, extents, nbr, true)) return -1;
  if (kind < 1 || kind > 3 || (kind > 1 && coords == nullptr)) return -1;
  sxo::sw_rhs_mixed(num_elements, extents, nbr, u, gamma2, inv_j, dt, coords,
                    sxo::kSwCpsrSommerfeld + kind - 1);
  return 0;
}

int sxo_sw_rhs_mixed(int num_elements, const int* extents, const int* nbr, const double* u,
                     const double* gamma2, const double* inv_j, double* dt) {
  if (bad_ragged(num_elements, extents, nbr)) return -1;
  sxo::sw_rhs_mixed(num_elements, extents, nbr, u, gamma2, inv_j, dt);
  return 0;
}

// Mixed-p GH RHS with external faces (neighbour -1) closed by the ghost state
// `ghost` (ragged, 50 components).
int sxo_gh_rhs_mixed_dirichlet(int num_elements, const int* extents, const int* nbr,
                               const double* u, const double* gammas, const double* inv_j,
                               const double* ghost, double* dt) {
  if (bad_ragged(num_elements, extents, nbr, true) || ghost == nullptr) return -1;
  sxo::gh_rhs_mixed(num_elements, extents, nbr, u, gammas, inv_j, dt, ghost);
  return 0;
}

int sxo_gh_rhs_mixed_internal_faces(int num_elements, const int* extents, const int* nbr,
                                    const double* u, const double* gammas,
                                    const double* inv_j, double* dt) {
  if (bad_ragged(num_elements, extents, nbr, true)) return -1;
  sxo::gh_rhs_mixed(num_elements, extents, nbr, u, gammas, inv_j, dt,
                    static_cast<const double*>(nullptr), true);
  return 0;
}

int sxo_gh_rhs_mixed(int num_elements, const int* extents, const int* nbr, const double* u,
                     const double* gammas, const double* inv_j, double* dt) {
  if (bad_ragged(num_elements, extents, nbr)) return -1;
  sxo::gh_rhs_mixed(num_elements, extents, nbr, u, gammas, inv_j, dt);
  return 0;
}

// The same RHS with the elements split over `threads` host threads (element
// data parallelism, as the reference distributes elements over cores).  Used
// for the CPU baseline.  Volume terms per element, then the face terms per
// element; each element writes only its own dt block.
int sxo_gh_rhs_threads(int threads, int num_elements, const int* ext, const int* nbr,
                       const double* u, const double* gammas, const double* inv_j,
                       double* dt) {
  if (bad_extents(ext) || num_elements < 1) return -1;
  const size_t np = static_cast<size_t>(ext[0]) * ext[1] * ext[2];
  std::vector<sxo::GhFaceTemps<double>> temps(static_cast<size_t>(num_elements) * np);
  run_split(threads, num_elements, [&](int b, int e) {
    sxo::gh_rhs_volume(b, e, ext, u, gammas, inv_j, dt, temps.data());
  });
  run_split(threads, num_elements, [&](int b, int e) {
    sxo::gh_rhs_faces(b, e, ext, nbr, u, inv_j, dt, temps.data());
  });
  return 0;
}

// The same RHS with the volume terms arranged as the reference computes them
// on a CPU (sxo_refstyle.hpp: transposes and matrix products per dimension,
// materialised temporaries): the CPU baseline of the benchmark.
int sxo_gh_rhs_refstyle_threads(int threads, int num_elements, const int* ext, const int* nbr,
                                const double* u, const double* gammas, const double* inv_j,
                                double* dt) {
  if (bad_extents(ext) || num_elements < 1) return -1;
  const size_t np = static_cast<size_t>(ext[0]) * ext[1] * ext[2];
  std::vector<sxo::GhFaceTemps<double>> temps(static_cast<size_t>(num_elements) * np);
  run_split(threads, num_elements, [&](int b, int e) {
    sxo::refstyle::gh_rhs_volume(b, e, ext, u, gammas, inv_j, dt, temps.data());
  });
  run_split(threads, num_elements, [&](int b, int e) {
    sxo::gh_rhs_faces(b, e, ext, nbr, u, inv_j, dt, temps.data());
  });
  return 0;
}

int sxo_sw_rhs_threads(int threads, int num_elements, const int* ext, const int* nbr,
                       const double* u, const double* gamma2, const double* inv_j,
                       double* dt) {
  if (bad_extents(ext) || num_elements < 1) return -1;
  run_split(threads, num_elements, [&](int b, int e) {
    sxo::sw_rhs_volume(b, e, ext, u, gamma2, inv_j, dt);
  });
  run_split(threads, num_elements, [&](int b, int e) {
    sxo::sw_rhs_faces(b, e, ext, nbr, u, gamma2, inv_j, dt);
  });
  return 0;
}

int sxo_rk3_substep(int substep, long n, double dt_step, const double* u0,
                    const double* uk, const double* lk, double* out) {
  if (substep < 0 || substep > 2) return -1;
  sxo::rk3_substep(substep, static_cast<size_t>(n), dt_step, u0, uk, lk, out);
  return 0;
}

// Non-conforming mortars (sxo_mortar.hpp).
int sxo_half_interpolation_matrix(int n_face, int n_mortar, int upper, double* out) {
  if (n_face < 2 || n_mortar < n_face || n_mortar > 20) return -1;
  const std::vector<double> m = sxo::half_interpolation_matrix(n_face, n_mortar, upper != 0);
  std::memcpy(out, m.data(), m.size() * sizeof(double));
  return 0;
}

int sxo_half_projection_matrix(int n_face, int n_mortar, int upper, double* out) {
  if (n_face < 2 || n_mortar < n_face || n_mortar > 20) return -1;
  const std::vector<double> m = sxo::half_projection_matrix(n_face, n_mortar, upper != 0);
  std::memcpy(out, m.data(), m.size() * sizeof(double));
  return 0;
}

int sxo_oriented_offset_on_slice(int n1, int n2, int sliced_dim, const int* mapped, int* out) {
  if (n1 < 1 || n2 < 1 || sliced_dim < 0 || sliced_dim > 2) return -1;
  sxo::Orientation o{{mapped[0], mapped[1], mapped[2]}};
  const std::vector<int> v = sxo::oriented_offset_on_slice(n1, n2, sliced_dim, o);
  std::memcpy(out, v.data(), v.size() * sizeof(int));
  return 0;
}

int sxo_sw_rhs_mortars(int num_elements, const int* extents, int num_records, const int* records,
                       const double* u, const double* gamma2, const double* inv_j, double* dt) {
  for (int e = 0; e < num_elements; ++e)
    if (bad_extents(extents + 3 * e)) return -1;
  try {
    sxo::rhs_mortars<false>(num_elements, extents, num_records, records, u, gamma2, inv_j, dt);
  } catch (const std::exception&) {
    return -2;
  }
  return 0;
}

int sxo_gh_rhs_mortars(int num_elements, const int* extents, int num_records, const int* records,
                       const double* u, const double* gammas, const double* inv_j, double* dt) {
  for (int e = 0; e < num_elements; ++e)
    if (bad_extents(extents + 3 * e)) return -1;
  try {
    sxo::rhs_mortars<true>(num_elements, extents, num_records, records, u, gammas, inv_j, dt);
  } catch (const std::exception&) {
    return -2;
  }
  return 0;
}

}  // extern "C"
