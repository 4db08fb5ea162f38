// Test of the C++ shim (spectre_amd/cpp/spectre_amd.hpp) on a HIP device:
// the reference-shaped calls give the oracle's results.
//  * partial_derivatives of polynomials of degree < N on an anisotropic LGL
//    mesh with an affine map are exact (closed form), and equal the oracle;
//  * ScalarWave / GH TimeDerivative::apply equal the oracle bit for bit;
//  * dg::lift_flux scales by -N(N-1)/2 |n| exactly;
//  * invalid arguments throw.
// Links the oracle (oracle/_build/liboracle.so), test infrastructure only.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>
#include <vector>

#include "../../spectre_amd/cpp/spectre_amd.hpp"

extern "C" {
int sxo_lgl_points_and_weights(int n, double* x, double* w);
int sxo_partial_derivatives(const double* u, int nc, const int* ext, const double* inv_j,
                            double* du);
int sxo_sw_time_derivative(int npts, const double* u, const double* du, const double* gamma2,
                           double* dt);
int sxo_gh_time_derivative(int npts, const double* u, const double* du, const double* gammas,
                           double* dt, double* temps);
}

namespace sa = spectre_amd;

static int failures = 0;
#define EXPECT(cond)                                                 \
  do {                                                               \
    if (!(cond)) {                                                   \
      std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
      ++failures;                                                    \
    }                                                                \
  } while (0)

static void to_device(sa::DeviceVariables& v, const std::vector<double>& h) {
  if (hipMemcpy(v.data(), h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    throw std::runtime_error("copy to device");
}
static std::vector<double> to_host(const sa::DeviceVariables& v) {
  std::vector<double> h(v.size());
  if (hipMemcpy(h.data(), v.data(), h.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    throw std::runtime_error("copy to host");
  return h;
}

int main() {
  // ---- partial derivatives on an anisotropic mesh -------------------------
  const sa::Mesh<3> mesh({{5, 7, 4}});
  const size_t np = mesh.number_of_grid_points();
  const double lo[3] = {-1.0, 0.5, 2.0}, h[3] = {2.0, 0.5, 3.0};
  std::vector<double> x[3];
  for (int d = 0; d < 3; ++d) {
    std::vector<double> w(mesh.extents(d));
    x[d].resize(mesh.extents(d));
    sxo_lgl_points_and_weights(static_cast<int>(mesh.extents(d)), x[d].data(), w.data());
  }
  // u_0 = X^4 Y^2 Z, u_1 = X Y^6 + Z^3, u_2 = 1 (degrees below the extents)
  std::vector<double> u(3 * np), exact(9 * np, 0.0), ij(9 * np, 0.0);
  for (size_t k = 0; k < mesh.extents(2); ++k)
    for (size_t j = 0; j < mesh.extents(1); ++j)
      for (size_t i = 0; i < mesh.extents(0); ++i) {
        const size_t p = i + mesh.extents(0) * (j + mesh.extents(1) * k);
        const double X = lo[0] + 0.5 * h[0] * (x[0][i] + 1.0);
        const double Y = lo[1] + 0.5 * h[1] * (x[1][j] + 1.0);
        const double Z = lo[2] + 0.5 * h[2] * (x[2][k] + 1.0);
        u[p] = std::pow(X, 4) * Y * Y * Z;
        u[np + p] = X * std::pow(Y, 6) + Z * Z * Z;
        u[2 * np + p] = 1.0;
        exact[0 * np + p] = 4 * std::pow(X, 3) * Y * Y * Z;
        exact[1 * np + p] = 2 * std::pow(X, 4) * Y * Z;
        exact[2 * np + p] = std::pow(X, 4) * Y * Y;
        exact[3 * np + p] = std::pow(Y, 6);
        exact[4 * np + p] = 6 * X * std::pow(Y, 5);
        exact[5 * np + p] = 3 * Z * Z;
        for (int d = 0; d < 3; ++d) ij[(d + 3 * d) * np + p] = 2.0 / h[d];
      }
  sa::DeviceVariables du_dev, u_dev(3, np), ij_dev(9, np);
  to_device(u_dev, u);
  to_device(ij_dev, ij);
  sa::partial_derivatives(&du_dev, u_dev, mesh, ij_dev);
  EXPECT(du_dev.number_of_independent_components() == 9);
  const std::vector<double> du = to_host(du_dev);
  double scale = 0.0, err = 0.0;
  for (size_t q = 0; q < du.size(); ++q) {
    scale = std::fmax(scale, std::fabs(exact[q]));
    err = std::fmax(err, std::fabs(du[q] - exact[q]));
  }
  std::printf("partial_derivatives: max error %.3e (scale %.3e)\n", err, scale);
  EXPECT(err <= 1e-12 * scale);
  std::vector<double> du_oracle(9 * np);
  const int ext[3] = {5, 7, 4};
  sxo_partial_derivatives(u.data(), 3, ext, ij.data(), du_oracle.data());
  EXPECT(std::memcmp(du.data(), du_oracle.data(), du.size() * sizeof(double)) == 0);

  // ---- point-wise systems ---------------------------------------------------
  std::mt19937 gen(7);
  std::uniform_real_distribution<double> dist(-1.0, 1.0);
  const size_t n = 777;
  {
    std::vector<double> su(5 * n), sdu(15 * n), g2(n), ref(5 * n);
    for (auto& v : su) v = dist(gen);
    for (auto& v : sdu) v = dist(gen);
    for (auto& v : g2) v = 0.5 + 0.5 * dist(gen);
    sa::DeviceVariables u5(5, n), du15(15, n), gam(1, n), dt, g2tmp;
    to_device(u5, su);
    to_device(du15, sdu);
    to_device(gam, g2);
    sa::ScalarWave::TimeDerivative::apply(&dt, &g2tmp, du15, u5, gam);
    sxo_sw_time_derivative(static_cast<int>(n), su.data(), sdu.data(), g2.data(), ref.data());
    const std::vector<double> got = to_host(dt);
    EXPECT(std::memcmp(got.data(), ref.data(), ref.size() * sizeof(double)) == 0);
    EXPECT(std::memcmp(to_host(g2tmp).data(), g2.data(), n * sizeof(double)) == 0);
  }
  {
    // near-Minkowski GH state: g = eta + small, random Pi, Phi and derivatives
    std::vector<double> gu(50 * n), gdu(150 * n), gam(3 * n), ref(50 * n), tref(12 * n);
    const double eta[4] = {-1.0, 1.0, 1.0, 1.0};
    const int diag[4] = {0, 4, 7, 9};
    for (auto& v : gu) v = 0.05 * dist(gen);
    for (auto& v : gdu) v = 0.05 * dist(gen);
    for (size_t p = 0; p < n; ++p)
      for (int a = 0; a < 4; ++a) gu[diag[a] * n + p] += eta[a];
    for (size_t p = 0; p < n; ++p) {
      gam[p] = 1.0;
      gam[n + p] = -1.0;
      gam[2 * n + p] = 1.0;
    }
    sa::DeviceVariables u50(50, n), du150(150, n), g3(3, n), dt, temps;
    to_device(u50, gu);
    to_device(du150, gdu);
    to_device(g3, gam);
    sa::gh::TimeDerivative::apply(&dt, &temps, du150, u50, g3);
    sxo_gh_time_derivative(static_cast<int>(n), gu.data(), gdu.data(), gam.data(), ref.data(),
                           tref.data());
    EXPECT(std::memcmp(to_host(dt).data(), ref.data(), ref.size() * sizeof(double)) == 0);
    EXPECT(std::memcmp(to_host(temps).data(), tref.data(), tref.size() * sizeof(double)) == 0);
  }

  // ---- lift ---------------------------------------------------------------
  {
    std::vector<double> bc(5 * n), mag(n);
    for (auto& v : bc) v = dist(gen);
    for (auto& v : mag) v = 1.0 + std::fabs(dist(gen));
    sa::DeviceVariables bc_dev(5, n), mag_dev(1, n);
    to_device(bc_dev, bc);
    to_device(mag_dev, mag);
    sa::dg::lift_flux(&bc_dev, 6, mag_dev);
    const std::vector<double> got = to_host(bc_dev);
    bool same = true;
    for (size_t c = 0; c < 5; ++c)
      for (size_t p = 0; p < n; ++p)
        same = same && got[c * n + p] == bc[c * n + p] * ((-0.5 * 30.0) * mag[p]);
    EXPECT(same);
  }

  // ---- errors -------------------------------------------------------------
  {
    bool threw = false;
    try {
      sa::DeviceVariables small(3, 10), out;
      sa::partial_derivatives(&out, small, mesh, ij_dev);
    } catch (const std::invalid_argument&) {
      threw = true;
    }
    EXPECT(threw);
    threw = false;
    std::vector<double> host_u(np);
    try {
      const int32_t e3[3] = {5, 7, 4};
      sa::check(sxd_partial_derivatives(1, e3, 1, host_u.data(), ij_dev.data(), du_dev.data(),
                                        nullptr));
    } catch (const std::runtime_error& e) {
      threw = std::strstr(e.what(), "device") != nullptr;
    }
    EXPECT(threw);
  }
  if (hipDeviceSynchronize() != hipSuccess) ++failures;
  std::printf("%s (%d failures)\n", failures == 0 ? "OK" : "FAILED", failures);
  return failures == 0 ? 0 : 1;
}
