// C++ shim over the spectre_amd C ABI with the call shapes of SpECTRE's own
// functions: a Mesh<Dim> of LGL extents, Variables-like SoA device storage,
// and partial_derivatives / TimeDerivative::apply / dg::lift_flux taking their
// arguments in the reference's order (outputs first, through pointers).
//
// Data stay on the device: DeviceVariables owns hipMalloc'd memory of
// nc blocks of Np doubles (the Variables layout, DataStructures/
// Variables.hpp:94-247).  A failing call throws std::runtime_error with the
// library's message, where the reference calls ERROR.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

#include "../../include/spectre_amd.h"

namespace spectre_amd {

inline void check(int status) {
  if (status != SXD_OK) throw std::runtime_error(std::string("spectre_amd: ") + sxd_last_error());
}

// LGL mesh of extents N_0 .. N_{Dim-1} (Spectral/Mesh.hpp:53-243; basis
// Legendre and quadrature Gauss-Lobatto in every dimension).
template <size_t Dim>
class Mesh {
 public:
  Mesh() = default;
  explicit Mesh(const std::array<size_t, Dim>& extents) : extents_(extents) {}
  size_t extents(size_t d) const { return extents_.at(d); }
  const std::array<size_t, Dim>& extents() const { return extents_; }
  size_t number_of_grid_points() const {
    size_t n = 1;
    for (size_t e : extents_) n *= e;
    return n;
  }
  // Mesh::slice_away (Mesh.hpp:190): the face mesh normal to dimension d.
  Mesh<Dim - 1> slice_away(size_t d) const {
    std::array<size_t, Dim - 1> e{};
    for (size_t i = 0, k = 0; i < Dim; ++i)
      if (i != d) e[k++] = extents_[i];
    return Mesh<Dim - 1>(e);
  }

 private:
  std::array<size_t, Dim> extents_{};
};

// nc components x Np points of device memory in the Variables layout.
class DeviceVariables {
 public:
  DeviceVariables() = default;
  DeviceVariables(size_t number_of_independent_components, size_t number_of_grid_points) {
    initialize(number_of_independent_components, number_of_grid_points);
  }
  DeviceVariables(const DeviceVariables&) = delete;
  DeviceVariables& operator=(const DeviceVariables&) = delete;
  DeviceVariables(DeviceVariables&& o) noexcept { *this = std::move(o); }
  DeviceVariables& operator=(DeviceVariables&& o) noexcept {
    std::swap(data_, o.data_);
    std::swap(nc_, o.nc_);
    std::swap(np_, o.np_);
    return *this;
  }
  ~DeviceVariables() {
    if (data_ != nullptr) (void)hipFree(data_);
  }
  // Variables::initialize: (re)allocate for a new size, contents undefined.
  void initialize(size_t number_of_independent_components, size_t number_of_grid_points) {
    if (number_of_independent_components == nc_ && number_of_grid_points == np_) return;
    if (data_ != nullptr) (void)hipFree(data_);
    data_ = nullptr;
    nc_ = number_of_independent_components;
    np_ = number_of_grid_points;
    if (size() > 0 && hipMalloc(&data_, size() * sizeof(double)) != hipSuccess) {
      nc_ = np_ = 0;
      throw std::runtime_error("spectre_amd: device allocation failed");
    }
  }
  double* data() { return data_; }
  const double* data() const { return data_; }
  double* component(size_t c) { return data_ + c * np_; }
  const double* component(size_t c) const { return data_ + c * np_; }
  size_t number_of_independent_components() const { return nc_; }
  size_t number_of_grid_points() const { return np_; }
  size_t size() const { return nc_ * np_; }

 private:
  double* data_ = nullptr;
  size_t nc_ = 0, np_ = 0;
};

// InverseJacobian<DataVector, 3, Logical, Inertial>: 9 components, d xi^j /
// d x^i at storage index j + 3 i (PartialDerivatives.tpp:70-77).
using DeviceInverseJacobian = DeviceVariables;

namespace detail {
inline std::array<int32_t, 3> extents3(const Mesh<3>& mesh) {
  return {static_cast<int32_t>(mesh.extents(0)), static_cast<int32_t>(mesh.extents(1)),
          static_cast<int32_t>(mesh.extents(2))};
}
inline void require(bool ok, const char* what) {
  if (!ok) throw std::invalid_argument(std::string("spectre_amd: ") + what);
}
}  // namespace detail

// partial_derivatives (LinearOperators/PartialDerivatives.tpp:191-241):
// du gets 3 nc components, block 3 c + i = d_i u_c; resized like the
// reference's output Variables.
inline void partial_derivatives(DeviceVariables* du, const DeviceVariables& u, const Mesh<3>& mesh,
                                const DeviceInverseJacobian& inverse_jacobian,
                                hipStream_t stream = nullptr) {
  detail::require(u.number_of_grid_points() == mesh.number_of_grid_points(),
                  "u does not match the mesh");
  detail::require(inverse_jacobian.number_of_independent_components() == 9 &&
                      inverse_jacobian.number_of_grid_points() == mesh.number_of_grid_points(),
                  "inverse Jacobian must have 9 components on the mesh");
  du->initialize(3 * u.number_of_independent_components(), mesh.number_of_grid_points());
  const auto ext = detail::extents3(mesh);
  check(sxd_partial_derivatives(1, ext.data(), static_cast<int>(u.number_of_independent_components()),
                                u.data(), inverse_jacobian.data(), du->data(), stream));
}

// logical_partial_derivatives (PartialDerivatives.tpp:137-184): the three
// logical derivatives of every component, d/dxi first.
inline void logical_partial_derivatives(std::array<DeviceVariables, 3>* du,
                                        const DeviceVariables& u, const Mesh<3>& mesh,
                                        hipStream_t stream = nullptr) {
  detail::require(u.number_of_grid_points() == mesh.number_of_grid_points(),
                  "u does not match the mesh");
  const size_t n = u.size();
  DeviceVariables all(3 * u.number_of_independent_components(), mesh.number_of_grid_points());
  const auto ext = detail::extents3(mesh);
  check(sxd_logical_partial_derivatives(1, ext.data(),
                                        static_cast<int>(u.number_of_independent_components()),
                                        u.data(), all.data(), stream));
  for (size_t d = 0; d < 3; ++d) {
    (*du)[d].initialize(u.number_of_independent_components(), mesh.number_of_grid_points());
    if (hipMemcpyAsync((*du)[d].data(), all.data() + d * n, n * sizeof(double),
                       hipMemcpyDeviceToDevice, stream) != hipSuccess)
      throw std::runtime_error("spectre_amd: copy failed");
  }
  if (hipStreamSynchronize(stream) != hipSuccess)
    throw std::runtime_error("spectre_amd: stream synchronisation failed");
}

namespace ScalarWave {
// ScalarWave::TimeDerivative<3> (Evolution/Systems/ScalarWave/
// TimeDerivative.cpp:13-45).  dt: (psi, Pi, Phi_i); result_gamma2: the
// temporary copy of gamma2; du: (d psi, d Pi, d Phi) in the
// partial_derivatives layout; u: (psi, Pi, Phi_i).
struct TimeDerivative {
  static void apply(DeviceVariables* dt, DeviceVariables* result_gamma2, const DeviceVariables& du,
                    const DeviceVariables& u, const DeviceVariables& gamma2,
                    hipStream_t stream = nullptr) {
    const size_t n = u.number_of_grid_points();
    detail::require(u.number_of_independent_components() == 5 &&
                        du.number_of_independent_components() == 15 &&
                        du.number_of_grid_points() == n && gamma2.number_of_grid_points() == n,
                    "ScalarWave needs u [5], du [15] and gamma2 [1] on the same points");
    dt->initialize(5, n);
    if (result_gamma2 != nullptr) result_gamma2->initialize(1, n);
    check(sxd_scalar_wave_time_derivative(static_cast<int64_t>(n), u.data(), du.data(),
                                          gamma2.data(), dt->data(),
                                          result_gamma2 ? result_gamma2->data() : nullptr,
                                          stream));
  }
};
}  // namespace ScalarWave

namespace gh {
// gh::TimeDerivative<3> with the harmonic gauge (Evolution/Systems/
// GeneralizedHarmonic/TimeDerivative.cpp:30-407).  temps: gamma1, gamma2,
// lapse, shift^i, inverse spatial metric (the temporaries the boundary
// correction uses), may be null.
struct TimeDerivative {
  static void apply(DeviceVariables* dt, DeviceVariables* temps, const DeviceVariables& du,
                    const DeviceVariables& u, const DeviceVariables& gammas,
                    hipStream_t stream = nullptr) {
    const size_t n = u.number_of_grid_points();
    detail::require(u.number_of_independent_components() == 50 &&
                        du.number_of_independent_components() == 150 &&
                        gammas.number_of_independent_components() == 3 &&
                        du.number_of_grid_points() == n && gammas.number_of_grid_points() == n,
                    "GH needs u [50], du [150] and gammas [3] on the same points");
    dt->initialize(50, n);
    if (temps != nullptr) temps->initialize(12, n);
    check(sxd_gh_time_derivative(static_cast<int64_t>(n), u.data(), du.data(), gammas.data(),
                                 dt->data(), temps ? temps->data() : nullptr, stream));
  }
};
}  // namespace gh

namespace dg {
// dg::lift_flux for Gauss-Lobatto points, in place
// (NumericalAlgorithms/DiscontinuousGalerkin/LiftFlux.hpp:41-62).
inline void lift_flux(DeviceVariables* flux, size_t extent_perpendicular_to_boundary,
                      const DeviceVariables& magnitude_of_face_normal,
                      hipStream_t stream = nullptr) {
  detail::require(magnitude_of_face_normal.number_of_grid_points() == flux->number_of_grid_points(),
                  "normal magnitude does not match the flux");
  check(sxd_lift_flux(static_cast<int64_t>(flux->number_of_grid_points()),
                      static_cast<int>(flux->number_of_independent_components()), flux->data(),
                      static_cast<int>(extent_perpendicular_to_boundary),
                      magnitude_of_face_normal.data(), stream));
}
}  // namespace dg

}  // namespace spectre_amd
