// C ABI of spectre_amd: argument validation, error reporting, device tables
// and the domain handle.  Host code; the kernels live in kernels.hip and
// gh_fused.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/spectre_amd.h"
#include "device_common.hpp"
#include "kernels.hpp"
#include "spectral.hpp"

namespace sxd {
// Device work arrays of the composite, mixed-p and adjoint paths: the
// derivative array `du` and the per-point / per-face-slot temporaries.
// Grown on demand (hipFree, then hipMalloc: a growth waits for the device),
// or, when `fixed`, a caller-provided block that is never reallocated.
struct Scratch {
  double* du = nullptr;
  double* temps = nullptr;
  size_t du_doubles = 0, temps_doubles = 0;
  bool fixed = false;
  size_t fixed_doubles = 0;  // capacity of a fixed block
  // A partitioned mortar RHS between its begin and end calls: the system,
  // the time step id of the exchange, and which receive slots are filled;
  // the message status array of the unpack calls.
  bool begun = false;
  int phase_system = -1;
  int64_t step = 0;
  std::vector<uint8_t> filled;
  int32_t* d_status = nullptr;
  size_t status_n = 0;
  // A second stream of the calling stream's work (mortar organisation: the
  // face packages and boundary terms run beside the volume terms), with the
  // fork and join events; created on first use outside a graph capture.
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;

  // The side stream and its events; false where they cannot be created.
  bool ensure_side() {
    if (side != nullptr) return true;
    hipStream_t s = nullptr;
    hipEvent_t f = nullptr, j = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return false;
    if (hipEventCreateWithFlags(&f, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&j, hipEventDisableTiming) != hipSuccess) {
      if (f != nullptr) (void)hipEventDestroy(f);
      (void)hipStreamDestroy(s);
      return false;
    }
    side = s;
    fork = f;
    join = j;
    return true;
  }

  // Make room for du_n + temps_n doubles.  Returns SXD_OK,
  // SXD_ERR_OUT_OF_MEMORY, or SXD_ERR_UNSUPPORTED for a fixed block that is
  // too small.
  int reserve(size_t du_n, size_t temps_n) {
    if (fixed) {
      if (du_n > fixed_doubles || temps_n > fixed_doubles - du_n) return SXD_ERR_UNSUPPORTED;
      temps = du + du_n;
      du_doubles = du_n;
      temps_doubles = temps_n;
      return SXD_OK;
    }
    if (du_doubles < du_n) {
      if (du != nullptr) (void)hipFree(du);
      du = nullptr;
      du_doubles = 0;
      if (hipMalloc(&du, du_n * sizeof(double)) != hipSuccess) return SXD_ERR_OUT_OF_MEMORY;
      du_doubles = du_n;
    }
    if (temps_doubles < temps_n) {
      if (temps != nullptr) (void)hipFree(temps);
      temps = nullptr;
      temps_doubles = 0;
      if (hipMalloc(&temps, temps_n * sizeof(double)) != hipSuccess) return SXD_ERR_OUT_OF_MEMORY;
      temps_doubles = temps_n;
    }
    return SXD_OK;
  }
  void release() {
    if (!fixed) {
      if (du != nullptr) (void)hipFree(du);
      if (temps != nullptr) (void)hipFree(temps);
    }
    if (d_status != nullptr) (void)hipFree(d_status);
    d_status = nullptr;
    status_n = 0;
    if (side != nullptr) {
      (void)hipEventDestroy(fork);
      (void)hipEventDestroy(join);
      (void)hipStreamDestroy(side);
    }
    side = nullptr;
    fork = join = nullptr;
    du = temps = nullptr;
    du_doubles = temps_doubles = 0;
  }
};
}  // namespace sxd

struct sxd_domain {
  int device = 0;
  int num_elements = 0;
  int32_t ext[3] = {0, 0, 0};
  int32_t* d_nbr = nullptr;
  // Some neighbour is -1 (an external face closed by a ghost state).
  bool has_boundary = false;
  // Work arrays of the composite (non-fused) paths, one set per stream that
  // has called the domain, allocated on first use: calls on different
  // streams never share one.
  std::mutex scratch_mutex;
  std::vector<std::pair<void*, std::unique_ptr<sxd::Scratch>>> scratch;
  // Kernel choice per system (SXD_SYSTEM_*), SXD_KERNEL_AUTO until set, and
  // threads per workgroup of the mixed-p kernels.
  int kernel[2] = {SXD_KERNEL_AUTO, SXD_KERNEL_AUTO};
  int threads[2] = {128, 128};
  // Mixed p: per-element extents and offsets (ragged layout).
  bool mixed = false;
  int64_t total_points = 0;
  int np_max = 0, kf = 0, km = 0;
  sxd::MixedElement* d_el = nullptr;
  int32_t* d_list = nullptr;          // all elements, decreasing cost
  int32_t* d_buckets = nullptr;       // elements grouped by extents
  int32_t* d_point_el = nullptr;      // element of each grid point
  std::vector<std::pair<int, int>> buckets;  // (start, count) in d_buckets
  // Mortar domains (non-aligned neighbours, h-refined faces): the ragged
  // element set above (its neighbour table unused) and the mortar tables.
  bool mortars = false;
  // A periodic mixed-p domain also holds the mortar tables of its
  // neighbour table (SXD_KERNEL_MORTAR).
  bool conforming_plan = false;
  sxd::MortarPlan plan;
};

namespace sxd {
namespace {

thread_local std::string g_last_error;

int fail(int status, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return status;
}

int ok() {
  g_last_error.clear();
  return SXD_OK;
}

bool valid_extents(const int32_t* ext) {
  if (ext == nullptr) return false;
  for (int d = 0; d < 3; ++d)
    if (ext[d] < kMinPoints || ext[d] > kMaxPoints) return false;
  return true;
}

// True if p is device (or managed) memory of the current device's address
// space.  Host pointers passed by mistake are rejected before any launch.
bool is_device_pointer(const void* p) {
  if (p == nullptr) return false;
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
         attr.type == hipMemoryTypeUnified;
}

#define SXD_REQUIRE_DEVICE(ptr)                                                     \
  do {                                                                              \
    if (!is_device_pointer(ptr))                                                    \
      return fail(SXD_ERR_INVALID_ARGUMENT, "%s: '%s' is not a device pointer",     \
                  __func__, #ptr);                                                  \
  } while (0)

int finish(int rc, const char* where) {
  if (rc == SXD_OK) return ok();
  if (rc == SXD_ERR_DEVICE) {
    const hipError_t err = hipGetLastError();
    return fail(rc, "%s: kernel launch failed: %s", where, hipGetErrorString(err));
  }
  if (rc == SXD_ERR_UNSUPPORTED) return fail(rc, "%s: unsupported configuration", where);
  if (rc == SXD_ERR_OUT_OF_MEMORY)
    return fail(rc, "%s: device allocation of the work arrays failed", where);
  return fail(rc, "%s: failed with status %d", where, rc);
}

struct DeviceTables {
  std::mutex mutex;
  std::vector<double*> per_device;
};

DeviceTables& tables() {
  static DeviceTables t;
  return t;
}

DeviceTables& mortar_tables() {
  static DeviceTables t;
  return t;
}

// Upload `host` once per device; the copy lives until the process ends.
const double* device_table(DeviceTables& t, const std::vector<double>& host) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lock(t.mutex);
  if (static_cast<int>(t.per_device.size()) <= dev) t.per_device.resize(dev + 1, nullptr);
  if (t.per_device[dev] == nullptr) {
    double* d = nullptr;
    if (hipMalloc(&d, host.size() * sizeof(double)) != hipSuccess) return nullptr;
    if (hipMemcpy(d, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice) !=
        hipSuccess) {
      (void)hipFree(d);
      return nullptr;
    }
    t.per_device[dev] = d;
  }
  return t.per_device[dev];
}

}  // namespace

const double* device_diff_matrices() {
  return device_table(tables(), packed_differentiation_matrices());
}

const double* device_mortar_matrices() {
  return device_table(mortar_tables(), packed_mortar_matrices());
}

const double* device_half_mortar_matrices() {
  static DeviceTables t;
  return device_table(t, packed_half_mortar_matrices());
}

const double* device_weak_flux_matrices() {
  static DeviceTables t;
  return device_table(t, packed_weak_flux_matrices());
}

const double* device_gauss_tables() {
  static DeviceTables t;
  return device_table(t, packed_gauss_tables());
}

}  // namespace sxd

using namespace sxd;

extern "C" {

const char* sxd_last_error(void) { return g_last_error.c_str(); }

const char* sxd_version(void) { return "spectre_amd 0.1 (gfx950, FP64)"; }

int sxd_lgl_points_and_weights(int n, double* x, double* w) {
  if (n < kMinPoints || n > kMaxPoints)
    return fail(SXD_ERR_INVALID_ARGUMENT, "LGL quadrature needs 2..20 points, got %d", n);
  if (x == nullptr || w == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null output");
  lgl_points_and_weights(n, x, w);
  return ok();
}

int sxd_lgl_differentiation_matrix(int n, double* d) {
  if (n < kMinPoints || n > kMaxPoints)
    return fail(SXD_ERR_INVALID_ARGUMENT, "LGL quadrature needs 2..20 points, got %d", n);
  if (d == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null output");
  const std::vector<double>& t = packed_differentiation_matrices();
  std::memcpy(d, t.data() + packed_d_offset(n), sizeof(double) * n * n);
  return ok();
}

int sxd_lg_points_and_weights(int n, double* x, double* w) {
  if (n < kMinPoints || n > kMaxPoints)
    return fail(SXD_ERR_INVALID_ARGUMENT, "Gauss quadrature here takes 2..20 points, got %d", n);
  if (x == nullptr || w == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null output");
  lg_points_and_weights(n, x, w);
  return ok();
}

int sxd_lg_differentiation_matrix(int n, double* d) {
  if (n < kMinPoints || n > kMaxPoints)
    return fail(SXD_ERR_INVALID_ARGUMENT, "Gauss quadrature here takes 2..20 points, got %d", n);
  if (d == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null output");
  const std::vector<double>& t = packed_gauss_tables();
  std::memcpy(d, t.data() + packed_d_offset(n), sizeof(double) * n * n);
  return ok();
}

int sxd_projection_matrix_parent_to_child(int n_parent, int n_child, int child_size,
                                          double* out) {
  if (n_parent < kMinPoints || n_child < n_parent || n_child > kMaxPoints)
    return fail(SXD_ERR_INVALID_ARGUMENT, "need 2 <= n_parent <= n_child <= 20");
  if (child_size < SXD_CHILD_FULL || child_size > SXD_CHILD_LOWER_HALF)
    return fail(SXD_ERR_INVALID_ARGUMENT, "unknown child size %d", child_size);
  if (out == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null output");
  if (child_size == SXD_CHILD_FULL)
    mortar_interpolation_matrix(n_parent, n_child, out);
  else
    half_mortar_interpolation_matrix(n_parent, n_child, child_size == SXD_CHILD_UPPER_HALF, out);
  return ok();
}

int sxd_projection_matrix_child_to_parent(int n_child, int n_parent, int child_size,
                                          double* out) {
  if (n_parent < kMinPoints || n_child < n_parent || n_child > kMaxPoints)
    return fail(SXD_ERR_INVALID_ARGUMENT, "need 2 <= n_parent <= n_child <= 20");
  if (child_size < SXD_CHILD_FULL || child_size > SXD_CHILD_LOWER_HALF)
    return fail(SXD_ERR_INVALID_ARGUMENT, "unknown child size %d", child_size);
  if (out == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null output");
  if (child_size == SXD_CHILD_FULL)
    mortar_projection_matrix(n_parent, n_child, out);
  else
    half_mortar_projection_matrix(n_parent, n_child, child_size == SXD_CHILD_UPPER_HALF, out);
  return ok();
}

int sxd_lg_boundary_terms(int n, double* out) {
  if (n < kMinPoints || n > kMaxPoints)
    return fail(SXD_ERR_INVALID_ARGUMENT, "Gauss quadrature here takes 2..20 points, got %d", n);
  if (out == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null output");
  const std::vector<double>& t = packed_gauss_tables();
  std::memcpy(out, t.data() + kPackedDSize + packed_gauss_boundary_offset(n),
              sizeof(double) * 4 * n);
  return ok();
}

int sxd_logical_partial_derivatives(int num_elements, const int32_t extents[3], int nc,
                                    const double* u, double* dlogical, void* stream) {
  if (!valid_extents(extents))
    return fail(SXD_ERR_INVALID_ARGUMENT, "extents must be 2..20 in each dimension");
  if (num_elements < 0 || nc < 1)
    return fail(SXD_ERR_INVALID_ARGUMENT, "bad element or component count");
  if (num_elements == 0) return ok();
  SXD_REQUIRE_DEVICE(u);
  SXD_REQUIRE_DEVICE(dlogical);
  return finish(launch_partial_derivatives(num_elements, extents, nc, u, nullptr, nullptr,
                                           dlogical, static_cast<hipStream_t>(stream)),
                __func__);
}

int sxd_partial_derivatives(int num_elements, const int32_t extents[3], int nc,
                            const double* u, const double* inv_jac, double* du,
                            void* stream) {
  if (!valid_extents(extents))
    return fail(SXD_ERR_INVALID_ARGUMENT, "extents must be 2..20 in each dimension");
  if (num_elements < 0 || nc < 1)
    return fail(SXD_ERR_INVALID_ARGUMENT, "bad element or component count");
  if (num_elements == 0) return ok();
  SXD_REQUIRE_DEVICE(u);
  SXD_REQUIRE_DEVICE(inv_jac);
  SXD_REQUIRE_DEVICE(du);
  return finish(launch_partial_derivatives(num_elements, extents, nc, u, inv_jac, du,
                                           nullptr, static_cast<hipStream_t>(stream)),
                __func__);
}

int sxd_partial_derivatives_adjoint(int num_elements, const int32_t extents[3], int nc,
                                    const double* dubar, const double* inv_jac, double* ubar,
                                    void* stream) {
  if (!valid_extents(extents))
    return fail(SXD_ERR_INVALID_ARGUMENT, "extents must be 2..20 in each dimension");
  if (num_elements < 0 || nc < 1)
    return fail(SXD_ERR_INVALID_ARGUMENT, "bad element or component count");
  if (num_elements == 0) return ok();
  SXD_REQUIRE_DEVICE(dubar);
  SXD_REQUIRE_DEVICE(inv_jac);
  SXD_REQUIRE_DEVICE(ubar);
  if (ubar == dubar) return fail(SXD_ERR_INVALID_ARGUMENT, "ubar must not alias dubar");
  return finish(launch_partial_derivatives_adjoint(num_elements, extents, nc, dubar, inv_jac,
                                                   ubar, static_cast<hipStream_t>(stream)),
                __func__);
}

namespace {
int check_divergence_args(int num_elements, const int32_t* extents, int nc) {
  if (!valid_extents(extents))
    return fail(SXD_ERR_INVALID_ARGUMENT, "extents must be 2..20 in each dimension");
  if (num_elements < 0 || nc < 1)
    return fail(SXD_ERR_INVALID_ARGUMENT, "bad element or component count");
  if (flux_divergence_lds_bytes(extents) > 160 * 1024)
    return fail(SXD_ERR_UNSUPPORTED, "3 Np doubles of fluxes do not fit the LDS");
  return SXD_OK;
}
}  // namespace

int sxd_divergence(int num_elements, const int32_t extents[3], int nc, const double* fluxes,
                   const double* inv_jac, double* div, void* stream) {
  if (int rc = check_divergence_args(num_elements, extents, nc); rc != SXD_OK) return rc;
  if (num_elements == 0) return ok();
  SXD_REQUIRE_DEVICE(fluxes);
  SXD_REQUIRE_DEVICE(inv_jac);
  SXD_REQUIRE_DEVICE(div);
  return finish(launch_flux_divergence(num_elements, extents, nc, false,
                                       const_cast<double*>(fluxes), inv_jac, div, false,
                                       nullptr, nullptr, nullptr, nullptr,
                                       static_cast<hipStream_t>(stream)),
                __func__);
}

int sxd_weak_divergence(int num_elements, const int32_t extents[3], int nc,
                        const double* fluxes, const double* det_jac_times_inv_jac, double* div,
                        void* stream) {
  if (int rc = check_divergence_args(num_elements, extents, nc); rc != SXD_OK) return rc;
  if (num_elements == 0) return ok();
  SXD_REQUIRE_DEVICE(fluxes);
  SXD_REQUIRE_DEVICE(det_jac_times_inv_jac);
  SXD_REQUIRE_DEVICE(div);
  return finish(launch_flux_divergence(num_elements, extents, nc, true,
                                       const_cast<double*>(fluxes), det_jac_times_inv_jac, div,
                                       false, nullptr, nullptr, nullptr, nullptr,
                                       static_cast<hipStream_t>(stream)),
                __func__);
}

int sxd_flux_volume_terms(int num_elements, const int32_t extents[3], int nc, const double* u,
                          double* fluxes, const double* mesh_velocity,
                          const double* div_mesh_velocity, int formulation,
                          const double* inv_jac, const double* det_jac_times_inv_jac,
                          const double* det_inv_jac, double* dt, void* stream) {
  if (int rc = check_divergence_args(num_elements, extents, nc); rc != SXD_OK) return rc;
  if (formulation != SXD_DG_STRONG_INERTIAL && formulation != SXD_DG_WEAK_INERTIAL)
    return fail(SXD_ERR_INVALID_ARGUMENT, "unknown DG formulation %d", formulation);
  if (num_elements == 0) return ok();
  SXD_REQUIRE_DEVICE(fluxes);
  SXD_REQUIRE_DEVICE(dt);
  const bool weak = formulation == SXD_DG_WEAK_INERTIAL;
  if (weak) {
    SXD_REQUIRE_DEVICE(det_jac_times_inv_jac);
    SXD_REQUIRE_DEVICE(det_inv_jac);
  } else {
    SXD_REQUIRE_DEVICE(inv_jac);
  }
  if (mesh_velocity != nullptr) {
    SXD_REQUIRE_DEVICE(mesh_velocity);
    SXD_REQUIRE_DEVICE(div_mesh_velocity);
    SXD_REQUIRE_DEVICE(u);
  }
  return finish(launch_flux_divergence(num_elements, extents, nc, weak, fluxes,
                                       weak ? det_jac_times_inv_jac : inv_jac, dt, true, u,
                                       mesh_velocity, div_mesh_velocity, det_inv_jac,
                                       static_cast<hipStream_t>(stream)),
                __func__);
}

int sxd_mesh_velocity_nonconservative(int64_t num_points, int nc, const double* du,
                                      const double* mesh_velocity, double* dt, void* stream) {
  if (num_points < 0 || nc < 1)
    return fail(SXD_ERR_INVALID_ARGUMENT, "bad point or component count");
  if (num_points == 0) return ok();
  SXD_REQUIRE_DEVICE(du);
  SXD_REQUIRE_DEVICE(mesh_velocity);
  SXD_REQUIRE_DEVICE(dt);
  return finish(launch_mesh_velocity_nonconservative(num_points, nc, du, mesh_velocity, dt,
                                                     static_cast<hipStream_t>(stream)),
                __func__);
}

int sxd_scalar_wave_time_derivative(int64_t num_points, const double* u, const double* du,
                                    const double* gamma2, double* dt, double* gamma2_temp,
                                    void* stream) {
  if (num_points < 0) return fail(SXD_ERR_INVALID_ARGUMENT, "negative point count");
  if (num_points == 0) return ok();
  SXD_REQUIRE_DEVICE(u);
  SXD_REQUIRE_DEVICE(du);
  SXD_REQUIRE_DEVICE(gamma2);
  SXD_REQUIRE_DEVICE(dt);
  if (gamma2_temp != nullptr) SXD_REQUIRE_DEVICE(gamma2_temp);
  return finish(launch_sw_time_derivative(num_points, u, du, gamma2, dt, gamma2_temp,
                                          static_cast<hipStream_t>(stream)),
                __func__);
}

int sxd_gh_time_derivative(int64_t num_points, const double* u, const double* du,
                           const double* gammas, double* dt, double* temps, void* stream) {
  if (num_points < 0) return fail(SXD_ERR_INVALID_ARGUMENT, "negative point count");
  if (num_points == 0) return ok();
  SXD_REQUIRE_DEVICE(u);
  SXD_REQUIRE_DEVICE(du);
  SXD_REQUIRE_DEVICE(gammas);
  SXD_REQUIRE_DEVICE(dt);
  if (temps != nullptr) SXD_REQUIRE_DEVICE(temps);
  return finish(launch_gh_time_derivative(num_points, u, du, gammas, dt, temps,
                                          static_cast<hipStream_t>(stream)),
                __func__);
}

int sxd_scalar_wave_upwind_penalty(int64_t num_points, const double* u_int,
                                   const double* gamma2_int, const double* normal_int,
                                   const double* u_ext, const double* gamma2_ext,
                                   const double* normal_ext, double* bc, void* stream) {
  if (num_points < 0) return fail(SXD_ERR_INVALID_ARGUMENT, "negative point count");
  if (num_points == 0) return ok();
  SXD_REQUIRE_DEVICE(u_int);
  SXD_REQUIRE_DEVICE(gamma2_int);
  SXD_REQUIRE_DEVICE(normal_int);
  SXD_REQUIRE_DEVICE(u_ext);
  SXD_REQUIRE_DEVICE(gamma2_ext);
  SXD_REQUIRE_DEVICE(normal_ext);
  SXD_REQUIRE_DEVICE(bc);
  return finish(launch_sw_upwind(num_points, u_int, gamma2_int, normal_int, u_ext,
                                 gamma2_ext, normal_ext, bc,
                                 static_cast<hipStream_t>(stream)),
                __func__);
}

int sxd_gh_upwind_penalty(int64_t num_points, const double* u_int, const double* temps_int,
                          const double* normal_int, const double* u_ext,
                          const double* temps_ext, const double* normal_ext, double* bc,
                          void* stream) {
  if (num_points < 0) return fail(SXD_ERR_INVALID_ARGUMENT, "negative point count");
  if (num_points == 0) return ok();
  SXD_REQUIRE_DEVICE(u_int);
  SXD_REQUIRE_DEVICE(temps_int);
  SXD_REQUIRE_DEVICE(normal_int);
  SXD_REQUIRE_DEVICE(u_ext);
  SXD_REQUIRE_DEVICE(temps_ext);
  SXD_REQUIRE_DEVICE(normal_ext);
  SXD_REQUIRE_DEVICE(bc);
  return finish(launch_gh_upwind(num_points, u_int, temps_int, normal_int, u_ext, temps_ext,
                                 normal_ext, bc, static_cast<hipStream_t>(stream)),
                __func__);
}

int sxd_lift_flux(int64_t num_points, int nc, double* bc, int extent_perpendicular,
                  const double* magnitude_of_normal, void* stream) {
  if (num_points < 0 || nc < 1)
    return fail(SXD_ERR_INVALID_ARGUMENT, "bad point or component count");
  if (extent_perpendicular < kMinPoints || extent_perpendicular > kMaxPoints)
    return fail(SXD_ERR_INVALID_ARGUMENT, "perpendicular extent must be 2..20");
  if (num_points == 0) return ok();
  SXD_REQUIRE_DEVICE(bc);
  SXD_REQUIRE_DEVICE(magnitude_of_normal);
  return finish(launch_lift_flux(num_points, nc, bc, extent_perpendicular,
                                 magnitude_of_normal, static_cast<hipStream_t>(stream)),
                __func__);
}

int sxd_domain_create(sxd_domain** domain, int num_elements, const int32_t extents[3],
                      const int32_t* neighbors) {
  if (domain == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null domain output");
  *domain = nullptr;
  if (num_elements < 1) return fail(SXD_ERR_INVALID_ARGUMENT, "need at least one element");
  if (!valid_extents(extents))
    return fail(SXD_ERR_INVALID_ARGUMENT, "extents must be 2..20 in each dimension");
  if (neighbors == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null neighbour table");
  bool has_boundary = false;
  for (int64_t i = 0; i < 6LL * num_elements; ++i) {
    if (neighbors[i] < -1 || neighbors[i] >= num_elements)
      return fail(SXD_ERR_INVALID_ARGUMENT, "neighbour %d of element %d out of range",
                  neighbors[i], static_cast<int>(i / 6));
    has_boundary = has_boundary || neighbors[i] == -1;
  }
  sxd_domain* d = new (std::nothrow) sxd_domain;
  if (d == nullptr) return fail(SXD_ERR_OUT_OF_MEMORY, "host allocation failed");
  if (hipGetDevice(&d->device) != hipSuccess) {
    delete d;
    return fail(SXD_ERR_DEVICE, "no HIP device");
  }
  d->num_elements = num_elements;
  d->has_boundary = has_boundary;
  for (int i = 0; i < 3; ++i) d->ext[i] = extents[i];
  const size_t bytes = sizeof(int32_t) * 6 * static_cast<size_t>(num_elements);
  if (hipMalloc(&d->d_nbr, bytes) != hipSuccess) {
    delete d;
    return fail(SXD_ERR_OUT_OF_MEMORY, "device allocation of the neighbour table failed");
  }
  if (hipMemcpy(d->d_nbr, neighbors, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d->d_nbr);
    delete d;
    return fail(SXD_ERR_DEVICE, "copy of the neighbour table failed");
  }
  if (device_diff_matrices() == nullptr) {
    (void)hipFree(d->d_nbr);
    delete d;
    return fail(SXD_ERR_DEVICE, "upload of the spectral tables failed");
  }
  *domain = d;
  return ok();
}

namespace {
int create_mixed_domain(sxd_domain** domain, int num_elements, const int32_t* extents,
                        const int32_t* neighbors, bool conforming_plan);
}  // namespace

int sxd_domain_create_mixed(sxd_domain** domain, int num_elements, const int32_t* extents,
                            const int32_t* neighbors) {
  return create_mixed_domain(domain, num_elements, extents, neighbors, true);
}

namespace {
// conforming_plan: also build the mortar tables of the neighbour table
// (aligned, Full) when there is no external face, for SXD_KERNEL_MORTAR.
int create_mixed_domain(sxd_domain** domain, int num_elements, const int32_t* extents,
                        const int32_t* neighbors, bool conforming_plan) {
  if (domain == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null domain output");
  *domain = nullptr;
  if (num_elements < 1) return fail(SXD_ERR_INVALID_ARGUMENT, "need at least one element");
  if (extents == nullptr || neighbors == nullptr)
    return fail(SXD_ERR_INVALID_ARGUMENT, "null extents or neighbour table");
  for (int e = 0; e < num_elements; ++e)
    if (!valid_extents(extents + 3 * e))
      return fail(SXD_ERR_INVALID_ARGUMENT, "extents of element %d must be 2..20", e);
  bool has_boundary = false;
  for (int64_t i = 0; i < 6LL * num_elements; ++i) {
    if (neighbors[i] < -1 || neighbors[i] >= num_elements)
      return fail(SXD_ERR_INVALID_ARGUMENT, "neighbour %d of element %d out of range",
                  neighbors[i], static_cast<int>(i / 6));
    has_boundary = has_boundary || neighbors[i] == -1;
  }
  std::vector<MixedElement> el(num_elements);
  int64_t off = 0;
  int np_max = 0, kf = 0, km = 0;
  for (int e = 0; e < num_elements; ++e) {
    MixedElement& m = el[e];
    for (int d = 0; d < 3; ++d) m.n[d] = extents[3 * e + d];
    m.np = m.n[0] * m.n[1] * m.n[2];
    m.off = off;
    off += m.np;
    np_max = std::max(np_max, m.np);
  }
  for (int e = 0; e < num_elements; ++e)
    for (int dir = 0; dir < 6; ++dir) {
      const int d = dir / 2;
      const int32_t* a = extents + 3 * e;
      // an external face has no mortar: it sizes like a face onto itself
      const int32_t* b = neighbors[6 * e + dir] < 0 ? a : extents + 3 * neighbors[6 * e + dir];
      const int fa = (d == 0 ? a[1] : a[0]), fb = (d == 2 ? a[1] : a[2]);
      const int ga = (d == 0 ? b[1] : b[0]), gb = (d == 2 ? b[1] : b[2]);
      kf = std::max({kf, fa * fb, ga * gb});
      km = std::max(km, std::max(fa, ga) * std::max(fb, gb));
    }
  if (5 * np_max * 8 > 160 * 1024 || mixed_face_lds_bytes(kf, km) > 160 * 1024)
    return fail(SXD_ERR_UNSUPPORTED, "mixed-p elements or faces too large for the LDS plan");
  // work lists: decreasing cost; and grouped by extents (per-extents launches)
  std::vector<int32_t> list(num_elements), grouped(num_elements);
  for (int e = 0; e < num_elements; ++e) list[e] = e;
  std::stable_sort(list.begin(), list.end(),
                   [&](int32_t x, int32_t y) { return el[x].np > el[y].np; });
  grouped = list;
  auto key = [&](int32_t x) {
    return (static_cast<int64_t>(el[x].n[0]) << 40) | (static_cast<int64_t>(el[x].n[1]) << 20) |
           el[x].n[2];
  };
  std::stable_sort(grouped.begin(), grouped.end(),
                   [&](int32_t x, int32_t y) { return key(x) < key(y); });
  sxd_domain* d = new (std::nothrow) sxd_domain;
  if (d == nullptr) return fail(SXD_ERR_OUT_OF_MEMORY, "host allocation failed");
  for (int i = 0; i < num_elements;) {
    int j = i;
    while (j < num_elements && key(grouped[j]) == key(grouped[i])) ++j;
    d->buckets.emplace_back(i, j - i);
    i = j;
  }
  d->mixed = true;
  d->has_boundary = has_boundary;
  d->num_elements = num_elements;
  d->total_points = off;
  d->np_max = np_max;
  d->kf = kf;
  d->km = km;
  int rc = SXD_OK;
  if (hipGetDevice(&d->device) != hipSuccess) rc = SXD_ERR_DEVICE;
  auto upload = [&](auto** dst, const auto* src, size_t n) {
    if (rc != SXD_OK) return;
    if (hipMalloc(dst, n * sizeof(*src)) != hipSuccess) {
      rc = SXD_ERR_OUT_OF_MEMORY;
      return;
    }
    if (hipMemcpy(*dst, src, n * sizeof(*src), hipMemcpyHostToDevice) != hipSuccess)
      rc = SXD_ERR_DEVICE;
  };
  upload(&d->d_nbr, neighbors, 6 * static_cast<size_t>(num_elements));
  upload(&d->d_el, el.data(), el.size());
  {
    std::vector<int32_t> point_el(static_cast<size_t>(off));
    for (int e = 0; e < num_elements; ++e)
      std::fill(point_el.begin() + el[e].off, point_el.begin() + el[e].off + el[e].np, e);
    upload(&d->d_point_el, point_el.data(), point_el.size());
  }
  upload(&d->d_list, list.data(), list.size());
  upload(&d->d_buckets, grouped.data(), grouped.size());
  if (rc == SXD_OK && (device_diff_matrices() == nullptr || device_mortar_matrices() == nullptr))
    rc = SXD_ERR_DEVICE;
  if (rc != SXD_OK) {
    sxd_domain_destroy(d);
    return fail(rc, "mixed domain set-up failed");
  }
  if (conforming_plan && !has_boundary) {
    std::vector<int32_t> records(48 * static_cast<size_t>(num_elements));
    for (int e = 0; e < num_elements; ++e)
      for (int dir = 0; dir < 6; ++dir) {
        int32_t* q = records.data() + 8 * (6 * static_cast<size_t>(e) + dir);
        const int32_t rec[8] = {e, dir, neighbors[6 * e + dir], 1, 3, 5, SXD_CHILD_FULL,
                                SXD_CHILD_FULL};
        std::memcpy(q, rec, sizeof(rec));
      }
    std::string error;
    rc = build_mortar_plan(num_elements, extents, 6 * num_elements, records.data(), nullptr, 0,
                           &d->plan, &error);
    if (rc == SXD_OK && device_half_mortar_matrices() == nullptr) rc = SXD_ERR_DEVICE;
    if (rc != SXD_OK) {
      sxd_domain_destroy(d);
      return fail(rc, "mortar tables of the mixed domain: %s", error.c_str());
    }
    d->conforming_plan = true;
  }
  *domain = d;
  return ok();
}
}  // namespace

namespace {
int create_mortar_domain(sxd_domain** domain, int num_elements, const int32_t* extents,
                         int num_records, const int32_t* records, const int32_t* partition,
                         int rank) {
  if (domain == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null domain output");
  *domain = nullptr;
  if (num_elements < 1) return fail(SXD_ERR_INVALID_ARGUMENT, "need at least one element");
  if (extents == nullptr || records == nullptr)
    return fail(SXD_ERR_INVALID_ARGUMENT, "null extents or record list");
  for (int e = 0; e < num_elements; ++e)
    if (!valid_extents(extents + 3 * e))
      return fail(SXD_ERR_INVALID_ARGUMENT, "extents of element %d must be 2..20", e);
  // the ragged set of this rank's elements (all of them without a
  // partition), with a neighbour table of self-references that no mortar
  // path reads
  std::vector<int32_t> local_ext;
  for (int e = 0; e < num_elements; ++e)
    if (partition == nullptr || partition[e] == rank)
      local_ext.insert(local_ext.end(), extents + 3 * e, extents + 3 * e + 3);
  const int owned = static_cast<int>(local_ext.size() / 3);
  if (owned == 0) return fail(SXD_ERR_INVALID_ARGUMENT, "rank %d owns no element", rank);
  std::vector<int32_t> self(6 * static_cast<size_t>(owned));
  for (size_t i = 0; i < self.size(); ++i) self[i] = static_cast<int32_t>(i / 6);
  sxd_domain* d = nullptr;
  if (int rc = create_mixed_domain(&d, owned, local_ext.data(), self.data(), false);
      rc != SXD_OK)
    return rc;
  std::string error;
  if (int rc = build_mortar_plan(num_elements, extents, num_records, records, partition, rank,
                                 &d->plan, &error);
      rc != SXD_OK) {
    sxd_domain_destroy(d);
    return fail(rc, "%s", error.c_str());
  }
  d->mortars = true;
  if (device_half_mortar_matrices() == nullptr) {
    sxd_domain_destroy(d);
    return fail(SXD_ERR_DEVICE, "upload of the half-interval mortar tables failed");
  }
  *domain = d;
  return ok();
}
}  // namespace

int sxd_domain_create_mortars(sxd_domain** domain, int num_elements, const int32_t* extents,
                              int num_records, const int32_t* records) {
  return create_mortar_domain(domain, num_elements, extents, num_records, records, nullptr, 0);
}

int sxd_domain_create_mortars_partition(sxd_domain** domain, int num_elements,
                                        const int32_t* extents, int num_records,
                                        const int32_t* records, const int32_t* partition,
                                        int rank) {
  if (partition == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null partition");
  for (int e = 0; e < num_elements; ++e)
    if (partition[e] < 0)
      return fail(SXD_ERR_INVALID_ARGUMENT, "element %d has a negative rank", e);
  return create_mortar_domain(domain, num_elements, extents, num_records, records, partition,
                              rank);
}

int sxd_domain_num_points(const sxd_domain* domain, int64_t* points) {
  if (domain == nullptr || points == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null argument");
  *points = domain->mixed ? domain->total_points
                          : static_cast<int64_t>(domain->num_elements) * domain->ext[0] *
                                domain->ext[1] * domain->ext[2];
  return ok();
}

int sxd_domain_set_threads(sxd_domain* domain, int system, int threads) {
  if (domain == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null domain");
  if (system != SXD_SYSTEM_SCALAR_WAVE && system != SXD_SYSTEM_GH)
    return fail(SXD_ERR_INVALID_ARGUMENT, "unknown system %d", system);
  if (threads != 64 && threads != 128 && threads != 256 && threads != 512)
    return fail(SXD_ERR_INVALID_ARGUMENT, "threads per workgroup must be 64, 128, 256 or 512");
  if (threads == 512 && domain->mixed)
    return fail(SXD_ERR_UNSUPPORTED, "mixed-p kernels run 64, 128 or 256 threads");
  domain->threads[system] = threads;
  return ok();
}

int sxd_domain_get_threads(const sxd_domain* domain, int system, int* threads) {
  if (domain == nullptr || threads == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null argument");
  if (system != SXD_SYSTEM_SCALAR_WAVE && system != SXD_SYSTEM_GH)
    return fail(SXD_ERR_INVALID_ARGUMENT, "unknown system %d", system);
  *threads = domain->threads[system];
  return ok();
}

int sxd_domain_destroy(sxd_domain* domain) {
  if (domain == nullptr) return ok();
  if (domain->d_el != nullptr) (void)hipFree(domain->d_el);
  if (domain->d_point_el != nullptr) (void)hipFree(domain->d_point_el);
  if (domain->d_list != nullptr) (void)hipFree(domain->d_list);
  if (domain->d_buckets != nullptr) (void)hipFree(domain->d_buckets);
  if (domain->d_nbr != nullptr) (void)hipFree(domain->d_nbr);
  if (domain->mortars || domain->conforming_plan) destroy_mortar_plan(&domain->plan);
  if (!domain->scratch.empty()) {
    (void)hipDeviceSynchronize();
    for (auto& s : domain->scratch) s.second->release();
  }
  delete domain;
  return ok();
}

namespace {
// The work arrays of `stream` on domain d (created empty on first use).
extern "C++" Scratch& stream_scratch(sxd_domain* d, hipStream_t stream) {
  std::lock_guard<std::mutex> lock(d->scratch_mutex);
  for (auto& s : d->scratch)
    if (s.first == static_cast<void*>(stream)) return *s.second;
  d->scratch.emplace_back(static_cast<void*>(stream), std::make_unique<Scratch>());
  return *d->scratch.back().second;
}

int check_domain(const sxd_domain* d) {
  if (d == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null domain");
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != d->device)
    return fail(SXD_ERR_INVALID_ARGUMENT, "domain belongs to device %d, current is %d",
                d->device, dev);
  return SXD_OK;
}

// Paths that read every face's neighbour (all but the Dirichlet RHS calls)
// need a domain without external faces.
int check_periodic(const sxd_domain* d) {
  if (int rc = check_domain(d); rc != SXD_OK) return rc;
  if (d->has_boundary)
    return fail(SXD_ERR_UNSUPPORTED, "the domain has external faces (neighbour -1); only "
                                     "the *_rhs_dirichlet calls evaluate it");
  return SXD_OK;
}

int64_t domain_points(const sxd_domain* d) {
  return d->mixed ? d->total_points
                  : static_cast<int64_t>(d->num_elements) * d->ext[0] * d->ext[1] * d->ext[2];
}

// The whole-domain RHS kernels read neighbours' face values of u while other
// workgroups write dt, so the two arrays must not overlap.
int check_no_overlap(const sxd_domain* d, int nc, const double* u, const double* dt,
                     const char* where) {
  const int64_t n = static_cast<int64_t>(nc) * domain_points(d);
  const uintptr_t a = reinterpret_cast<uintptr_t>(u), b = reinterpret_cast<uintptr_t>(dt);
  const uintptr_t bytes = static_cast<uintptr_t>(n) * sizeof(double);
  if (a < b + bytes && b < a + bytes)
    return fail(SXD_ERR_INVALID_ARGUMENT, "%s: dt overlaps u (in-place evaluation is not "
                                          "supported)", where);
  return SXD_OK;
}
}  // namespace

namespace {
// Kernels a domain can run for a system, fastest expected first; returns
// their number (at most 3).
int candidate_kernels(const sxd_domain* d, int system, int* k) {
  int n = 0;
  if (d->mortars) {
    k[n++] = SXD_KERNEL_MORTAR;
    return n;
  }
  if (d->mixed) {
    if (d->conforming_plan) k[n++] = SXD_KERNEL_MORTAR;
    k[n++] = SXD_KERNEL_MIXED_BATCHED;
    k[n++] = SXD_KERNEL_MIXED_PER_EXTENTS;
    return n;
  }
  if (system == SXD_SYSTEM_SCALAR_WAVE) {
    if (sw_fused_supported(d->ext)) k[n++] = SXD_KERNEL_FUSED;
  } else {
    if (gh_fused_supported(d->ext)) {
      k[n++] = SXD_KERNEL_FUSED;
      k[n++] = SXD_KERNEL_FUSED_MFMA;
    }
  }
  k[n++] = SXD_KERNEL_COMPOSITE;
  return n;
}

int resolve_kernel(const sxd_domain* d, int system) {
  if (d->kernel[system] != SXD_KERNEL_AUTO) return d->kernel[system];
  int k[3];
  candidate_kernels(d, system, k);
  return k[0];
}

// Reserve work arrays (status only; finish() words the message).
int reserve_scratch(Scratch& sc, size_t du_doubles, size_t temps_doubles) {
  return sc.reserve(du_doubles, temps_doubles);
}

// Volume terms on the ragged element set, then the corrections of every
// mortar; work arrays: the derivatives (GH) and the mortar buffers.
int run_mortars(sxd_domain* d, Scratch& sc, bool gh, const double* u, const double* params,
                const double* inv_jac, double* dt, hipStream_t stream) {
  const size_t du = gh ? 150 * static_cast<size_t>(d->total_points) : 0;
  if (int rc = reserve_scratch(sc, du, mortar_work_doubles(gh, d->plan)); rc != SXD_OK) return rc;
  const int threads = d->threads[gh ? SXD_SYSTEM_GH : SXD_SYSTEM_SCALAR_WAVE];
  auto volume = [&](hipStream_t s) {
    return launch_mixed_volume(gh, d->num_elements, d->d_list, d->d_el, d->np_max, u, params,
                               inv_jac, dt, sc.du, threads, s);
  };
  // The packages and boundary terms read u only and write the work array:
  // they run on a second stream beside the volume terms, and the lift (which
  // adds into dt after them) waits for both.  Inside a graph capture the
  // same fork and join become graph edges; a stream first seen inside a
  // capture (no side stream yet) runs everything in order.
  hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &capture) != hipSuccess) return SXD_ERR_DEVICE;
  const bool fork = d->plan.num_remote == 0 &&
                    (sc.side != nullptr ||
                     (capture == hipStreamCaptureStatusNone && sc.ensure_side()));
  if (!fork) {
    if (int rc = volume(stream); rc != SXD_OK) return rc;
    return launch_mortar_faces(gh, d->plan, d->d_el, u, params, inv_jac, dt, sc.temps, stream);
  }
  if (hipEventRecord(sc.fork, stream) != hipSuccess ||
      hipStreamWaitEvent(sc.side, sc.fork, 0) != hipSuccess)
    return SXD_ERR_DEVICE;
  int rc = launch_mortar_packages(gh, d->plan, d->d_el, u, params, inv_jac, sc.temps, sc.side);
  if (rc == SXD_OK) rc = launch_mortar_boundary_terms(gh, d->plan, sc.temps, sc.side);
  const int rc_volume = volume(stream);
  // join in every case, so that no work of this call is left on the side
  // stream after it returns
  if (hipEventRecord(sc.join, sc.side) != hipSuccess ||
      hipStreamWaitEvent(stream, sc.join, 0) != hipSuccess)
    return SXD_ERR_DEVICE;
  if (rc != SXD_OK) return rc;
  if (rc_volume != SXD_OK) return rc_volume;
  return launch_mortar_lift(gh, d->plan, d->d_el, dt, sc.temps, stream);
}

int run_mixed(sxd_domain* d, Scratch& sc, bool gh, int kernel, const double* u,
              const double* params, const double* inv_jac, double* dt, hipStream_t stream,
              const double* ghost = nullptr, int bc_kind = kSwBcGhostArray) {
  if (d->mortars) {
    if (ghost != nullptr || bc_kind != kSwBcGhostArray) return SXD_ERR_UNSUPPORTED;
    return run_mortars(d, sc, gh, u, params, inv_jac, dt, stream);
  }
  if (kernel == SXD_KERNEL_MORTAR) {
    // external faces take the mixed-p kernels (no mortar plan is built then)
    if (!d->conforming_plan || ghost != nullptr || bc_kind != kSwBcGhostArray)
      kernel = SXD_KERNEL_MIXED_BATCHED;
    else
      return run_mortars(d, sc, gh, u, params, inv_jac, dt, stream);
  }
  double* scratch = nullptr;
  if (gh) {
    if (int rc = reserve_scratch(sc, 150 * static_cast<size_t>(d->total_points), 0); rc != SXD_OK)
      return rc;
    scratch = sc.du;
  }
  const int threads = d->threads[gh ? SXD_SYSTEM_GH : SXD_SYSTEM_SCALAR_WAVE];
  if (kernel == SXD_KERNEL_MIXED_BATCHED)
    return launch_mixed_rhs(gh, d->num_elements, d->d_list, d->d_el, d->d_nbr, d->np_max, d->kf,
                            d->km, u, params, inv_jac, dt, scratch, threads, stream, ghost,
                            bc_kind);
  for (const auto& b : d->buckets) {
    const int rc = launch_mixed_rhs(gh, b.second, d->d_buckets + b.first, d->d_el, d->d_nbr,
                                    d->np_max, d->kf, d->km, u, params, inv_jac, dt, scratch,
                                    threads, stream, ghost, bc_kind);
    if (rc != SXD_OK) return rc;
  }
  return SXD_OK;
}

int run_sw(sxd_domain* d, Scratch& sc, int kernel, const double* u, const double* gamma2,
           const double* inv_jac, double* dt, hipStream_t stream,
           const double* ghost = nullptr, int bc_kind = kSwBcGhostArray) {
  if (d->mixed) {
    return run_mixed(d, sc, false, kernel, u, gamma2, inv_jac, dt, stream, ghost, bc_kind);
  }
  double* scratch = nullptr;
  if (kernel == SXD_KERNEL_COMPOSITE) {
    const size_t np = static_cast<size_t>(d->ext[0]) * d->ext[1] * d->ext[2];
    if (int rc = reserve_scratch(sc, 15 * np * d->num_elements, 0); rc != SXD_OK) return rc;
    scratch = sc.du;
  }
  return launch_sw_rhs(d->num_elements, d->ext, d->d_nbr, u, gamma2, inv_jac, dt,
                       kernel == SXD_KERNEL_FUSED, scratch, ghost, bc_kind, stream);
}

int run_gh(sxd_domain* d, Scratch& sc, int kernel, const double* u, const double* gammas,
           const double* inv_jac, double* dt, hipStream_t stream,
           const double* ghost = nullptr) {
  if (d->mixed) return run_mixed(d, sc, true, kernel, u, gammas, inv_jac, dt, stream, ghost);
  if (kernel == SXD_KERNEL_FUSED || kernel == SXD_KERNEL_FUSED_MFMA)
    return launch_gh_rhs_fused(d->num_elements, d->ext, d->d_nbr, u, gammas, inv_jac, dt,
                               kernel == SXD_KERNEL_FUSED ? kGhFusedValu : kGhFusedMfma,
                               d->threads[SXD_SYSTEM_GH] == 512 ? 512 : 256, ghost, stream);
  const size_t np = static_cast<size_t>(d->ext[0]) * d->ext[1] * d->ext[2];
  if (int rc = reserve_scratch(sc, 150 * np * d->num_elements, 12 * np * d->num_elements);
      rc != SXD_OK)
    return rc;
  return launch_gh_rhs_composite(d->num_elements, d->ext, d->d_nbr, u, gammas, inv_jac, dt,
                                 sc.du, sc.temps, ghost, stream);
}

int run_system(sxd_domain* d, Scratch& sc, int system, int kernel, const double* u,
               const double* params, const double* inv_jac, double* dt, hipStream_t stream) {
  return system == SXD_SYSTEM_SCALAR_WAVE
             ? run_sw(d, sc, kernel, u, params, inv_jac, dt, stream)
             : run_gh(d, sc, kernel, u, params, inv_jac, dt, stream);
}
}  // namespace

int sxd_scalar_wave_rhs(const sxd_domain* domain, const double* u, const double* gamma2,
                        const double* inv_jac, double* dt, void* stream) {
  if (int rc = check_periodic(domain); rc != SXD_OK) return rc;
  SXD_REQUIRE_DEVICE(u);
  SXD_REQUIRE_DEVICE(gamma2);
  SXD_REQUIRE_DEVICE(inv_jac);
  SXD_REQUIRE_DEVICE(dt);
  sxd_domain* d = const_cast<sxd_domain*>(domain);
  if (int rc = check_no_overlap(d, 5, u, dt, __func__); rc != SXD_OK) return rc;
  Scratch& sc = stream_scratch(d, static_cast<hipStream_t>(stream));
  return finish(run_sw(d, sc, resolve_kernel(d, SXD_SYSTEM_SCALAR_WAVE), u, gamma2, inv_jac, dt,
                       static_cast<hipStream_t>(stream)),
                __func__);
}

int sxd_scalar_wave_rhs_dirichlet(const sxd_domain* domain, const double* u,
                                  const double* gamma2, const double* inv_jac,
                                  const double* ghost, double* dt, void* stream) {
  if (int rc = check_domain(domain); rc != SXD_OK) return rc;
  SXD_REQUIRE_DEVICE(u);
  SXD_REQUIRE_DEVICE(gamma2);
  SXD_REQUIRE_DEVICE(inv_jac);
  SXD_REQUIRE_DEVICE(ghost);
  SXD_REQUIRE_DEVICE(dt);
  sxd_domain* d = const_cast<sxd_domain*>(domain);
  if (int rc = check_no_overlap(d, 5, u, dt, __func__); rc != SXD_OK) return rc;
  Scratch& sc = stream_scratch(d, static_cast<hipStream_t>(stream));
  return finish(run_sw(d, sc, resolve_kernel(d, SXD_SYSTEM_SCALAR_WAVE), u, gamma2, inv_jac, dt,
                       static_cast<hipStream_t>(stream), ghost),
                __func__);
}

int sxd_scalar_wave_rhs_spherical_radiation(const sxd_domain* domain, const double* u,
                                            const double* gamma2, const double* inv_jac,
                                            const double* coords, int type, double* dt,
                                            void* stream) {
  if (int rc = check_domain(domain); rc != SXD_OK) return rc;
  SXD_REQUIRE_DEVICE(u);
  SXD_REQUIRE_DEVICE(gamma2);
  SXD_REQUIRE_DEVICE(inv_jac);
  SXD_REQUIRE_DEVICE(dt);
  if (type != SXD_SPHERICAL_RADIATION_SOMMERFELD &&
      type != SXD_SPHERICAL_RADIATION_BAYLISS_TURKEL)
    return fail(SXD_ERR_INVALID_ARGUMENT, "unknown spherical radiation type %d", type);
  if (type == SXD_SPHERICAL_RADIATION_BAYLISS_TURKEL) SXD_REQUIRE_DEVICE(coords);
  sxd_domain* d = const_cast<sxd_domain*>(domain);
  if (int rc = check_no_overlap(d, 5, u, dt, __func__); rc != SXD_OK) return rc;
  Scratch& sc = stream_scratch(d, static_cast<hipStream_t>(stream));
  const int kind = type == SXD_SPHERICAL_RADIATION_SOMMERFELD ? kSwBcSommerfeld
                                                              : kSwBcBaylissTurkel;
  return finish(run_sw(d, sc, resolve_kernel(d, SXD_SYSTEM_SCALAR_WAVE), u, gamma2, inv_jac, dt,
                       static_cast<hipStream_t>(stream),
                       kind == kSwBcBaylissTurkel ? coords : nullptr, kind),
                __func__);
}

int sxd_scalar_wave_rhs_constraint_preserving_spherical_radiation(
    const sxd_domain* domain, const double* u, const double* gamma2, const double* inv_jac,
    const double* coords, int type, double* dt, void* stream) {
  if (int rc = check_domain(domain); rc != SXD_OK) return rc;
  SXD_REQUIRE_DEVICE(u);
  SXD_REQUIRE_DEVICE(gamma2);
  SXD_REQUIRE_DEVICE(inv_jac);
  SXD_REQUIRE_DEVICE(dt);
  if (type != SXD_CPSR_SOMMERFELD && type != SXD_CPSR_FIRST_ORDER_BAYLISS_TURKEL &&
      type != SXD_CPSR_SECOND_ORDER_BAYLISS_TURKEL)
    return fail(SXD_ERR_INVALID_ARGUMENT,
                "unknown constraint-preserving spherical radiation type %d", type);
  if (type != SXD_CPSR_SOMMERFELD) SXD_REQUIRE_DEVICE(coords);
  sxd_domain* d = const_cast<sxd_domain*>(domain);
  if (int rc = check_no_overlap(d, 5, u, dt, __func__); rc != SXD_OK) return rc;
  Scratch& sc = stream_scratch(d, static_cast<hipStream_t>(stream));
  const int kind = kSwBcCpsrSommerfeld + (type - SXD_CPSR_SOMMERFELD);
  return finish(run_sw(d, sc, resolve_kernel(d, SXD_SYSTEM_SCALAR_WAVE), u, gamma2, inv_jac, dt,
                       static_cast<hipStream_t>(stream),
                       type == SXD_CPSR_SOMMERFELD ? nullptr : coords, kind),
                __func__);
}

int sxd_scalar_wave_rhs_gauss(const sxd_domain* domain, const double* u, const double* gamma2,
                              const double* inv_jac, const double* det_inv_jac,
                              const double* face_inv_jac, double* dt, void* stream) {
  if (int rc = check_periodic(domain); rc != SXD_OK) return rc;
  if (domain->mixed)
    return fail(SXD_ERR_UNSUPPORTED, "%s: Gauss meshes take uniform domains", __func__);
  SXD_REQUIRE_DEVICE(u);
  SXD_REQUIRE_DEVICE(gamma2);
  SXD_REQUIRE_DEVICE(inv_jac);
  SXD_REQUIRE_DEVICE(det_inv_jac);
  SXD_REQUIRE_DEVICE(face_inv_jac);
  SXD_REQUIRE_DEVICE(dt);
  sxd_domain* d = const_cast<sxd_domain*>(domain);
  if (int rc = check_no_overlap(d, 5, u, dt, __func__); rc != SXD_OK) return rc;
  if (device_gauss_tables() == nullptr)
    return fail(SXD_ERR_DEVICE, "%s: upload of the Gauss tables failed", __func__);
  Scratch& sc = stream_scratch(d, static_cast<hipStream_t>(stream));
  const size_t np = static_cast<size_t>(d->ext[0]) * d->ext[1] * d->ext[2];
  if (int rc = reserve_scratch(sc, 15 * np * d->num_elements, 0); rc != SXD_OK)
    return finish(rc, __func__);
  return finish(launch_sw_rhs_gauss(d->num_elements, d->ext, d->d_nbr, u, gamma2, inv_jac,
                                    det_inv_jac, face_inv_jac, dt, sc.du,
                                    static_cast<hipStream_t>(stream)),
                __func__);
}

int sxd_gh_rhs(const sxd_domain* domain, const double* u, const double* gammas,
               const double* inv_jac, double* dt, void* stream) {
  if (int rc = check_periodic(domain); rc != SXD_OK) return rc;
  SXD_REQUIRE_DEVICE(u);
  SXD_REQUIRE_DEVICE(gammas);
  SXD_REQUIRE_DEVICE(inv_jac);
  SXD_REQUIRE_DEVICE(dt);
  sxd_domain* d = const_cast<sxd_domain*>(domain);
  if (int rc = check_no_overlap(d, 50, u, dt, __func__); rc != SXD_OK) return rc;
  Scratch& sc = stream_scratch(d, static_cast<hipStream_t>(stream));
  return finish(run_gh(d, sc, resolve_kernel(d, SXD_SYSTEM_GH), u, gammas, inv_jac, dt,
                       static_cast<hipStream_t>(stream)),
                __func__);
}

int sxd_gh_rhs_noalloc(const sxd_domain* domain, const double* u, const double* gammas,
                       const double* inv_jac, double* dt, void* stream) {
  if (int rc = check_periodic(domain); rc != SXD_OK) return rc;
  if (domain->mortars)
    return fail(SXD_ERR_UNSUPPORTED, "%s: not built for mortar domains", __func__);
  sxd_domain* d = const_cast<sxd_domain*>(domain);
  int kernel = resolve_kernel(d, SXD_SYSTEM_GH);
  // the mortar organisation's package buffers do not fit the fixed array at
  // useful sizes: mixed-p domains run the batched mixed-p kernels here
  if (kernel == SXD_KERNEL_MORTAR) kernel = SXD_KERNEL_MIXED_BATCHED;
  // work arrays the kernel organisation needs: 150 doubles per point (mixed
  // p), 162 (composite), none (fused)
  const size_t points = static_cast<size_t>(domain_points(d));
  size_t need = 0;
  if (d->mixed)
    need = 150 * points;
  else if (kernel == SXD_KERNEL_COMPOSITE)
    need = 162 * points;
  if (need > noalloc_scratch_doubles())
    return fail(SXD_ERR_UNSUPPORTED,
                "%s: the domain's work arrays need %zu doubles, above the fixed array's %zu "
                "(16 Mi doubles, 128 MiB)",
                __func__, need, noalloc_scratch_doubles());
  SXD_REQUIRE_DEVICE(u);
  SXD_REQUIRE_DEVICE(gammas);
  SXD_REQUIRE_DEVICE(inv_jac);
  SXD_REQUIRE_DEVICE(dt);
  if (int rc = check_no_overlap(d, 50, u, dt, __func__); rc != SXD_OK) return rc;
  Scratch sc;
  sc.fixed = true;
  sc.fixed_doubles = noalloc_scratch_doubles();
  if (need > 0) {
    if (int rc = noalloc_scratch(&sc.du); rc != SXD_OK)
      return fail(rc, "%s: the fixed work array has no address on this device", __func__);
  }
  return finish(run_gh(d, sc, kernel, u, gammas, inv_jac, dt, static_cast<hipStream_t>(stream)),
                __func__);
}

int sxd_gh_rhs_dirichlet(const sxd_domain* domain, const double* u, const double* gammas,
                         const double* inv_jac, const double* ghost, double* dt,
                         void* stream) {
  if (int rc = check_domain(domain); rc != SXD_OK) return rc;
  SXD_REQUIRE_DEVICE(u);
  SXD_REQUIRE_DEVICE(gammas);
  SXD_REQUIRE_DEVICE(inv_jac);
  SXD_REQUIRE_DEVICE(ghost);
  SXD_REQUIRE_DEVICE(dt);
  sxd_domain* d = const_cast<sxd_domain*>(domain);
  if (int rc = check_no_overlap(d, 50, u, dt, __func__); rc != SXD_OK) return rc;
  Scratch& sc = stream_scratch(d, static_cast<hipStream_t>(stream));
  const int kernel = resolve_kernel(d, SXD_SYSTEM_GH);
  if (kernel == SXD_KERNEL_FUSED_MFMA)
    return fail(SXD_ERR_UNSUPPORTED, "external faces are not built for the MFMA engine; "
                                     "select the fused or composite kernel");
  return finish(run_gh(d, sc, kernel, u, gammas, inv_jac, dt, static_cast<hipStream_t>(stream),
                       ghost),
                __func__);
}

// ---- partitioned mortar RHS around a boundary-message exchange -------------

int sxd_boundary_messages_layout(const sxd_domain* domain, int system, int* num_peers,
                                 int32_t* ranks, int64_t* offsets, int64_t* bytes,
                                 int64_t* total_bytes) {
  if (int rc = check_domain(domain); rc != SXD_OK) return rc;
  if (!domain->mortars) return fail(SXD_ERR_UNSUPPORTED, "not a mortar domain");
  if (system != SXD_SYSTEM_SCALAR_WAVE && system != SXD_SYSTEM_GH)
    return fail(SXD_ERR_INVALID_ARGUMENT, "unknown system %d", system);
  if (num_peers == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null output");
  const auto& peers = domain->plan.peers;
  *num_peers = static_cast<int>(peers.size());
  if (total_bytes != nullptr) *total_bytes = domain->plan.message_bytes[system];
  for (size_t i = 0; i < peers.size(); ++i) {
    if (ranks != nullptr) ranks[i] = peers[i].rank;
    if (offsets != nullptr) offsets[i] = peers[i].offset[system];
    if (bytes != nullptr) bytes[i] = peers[i].bytes[system];
  }
  return ok();
}

int sxd_boundary_messages_expected(const sxd_domain* domain, int rank, int* count) {
  if (int rc = check_domain(domain); rc != SXD_OK) return rc;
  if (!domain->mortars) return fail(SXD_ERR_UNSUPPORTED, "not a mortar domain");
  if (count == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null output");
  const auto it = domain->plan.expected.find(rank);
  *count = it == domain->plan.expected.end() ? 0 : it->second;
  return ok();
}

int sxd_rhs_begin(const sxd_domain* domain, int system, const double* u, const double* params,
                  const double* inv_jac, double* dt, void* send, int64_t time_step_id,
                  void* stream) {
  if (int rc = check_domain(domain); rc != SXD_OK) return rc;
  if (!domain->mortars) return fail(SXD_ERR_UNSUPPORTED, "%s: not a mortar domain", __func__);
  if (system != SXD_SYSTEM_SCALAR_WAVE && system != SXD_SYSTEM_GH)
    return fail(SXD_ERR_INVALID_ARGUMENT, "unknown system %d", system);
  SXD_REQUIRE_DEVICE(u);
  SXD_REQUIRE_DEVICE(params);
  SXD_REQUIRE_DEVICE(inv_jac);
  SXD_REQUIRE_DEVICE(dt);
  if (domain->plan.message_bytes[system] > 0) SXD_REQUIRE_DEVICE(send);
  sxd_domain* d = const_cast<sxd_domain*>(domain);
  const bool gh = system == SXD_SYSTEM_GH;
  if (int rc = check_no_overlap(d, gh ? 50 : 5, u, dt, __func__); rc != SXD_OK) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  Scratch& sc = stream_scratch(d, s);
  const size_t du = gh ? 150 * static_cast<size_t>(d->total_points) : 0;
  int rc = reserve_scratch(sc, du, mortar_work_doubles(gh, d->plan));
  if (rc == SXD_OK)
    rc = launch_mixed_volume(gh, d->num_elements, d->d_list, d->d_el, d->np_max, u, params,
                             inv_jac, dt, sc.du, d->threads[system], s);
  if (rc == SXD_OK)
    rc = launch_mortar_packages(gh, d->plan, d->d_el, u, params, inv_jac, sc.temps, s);
  if (rc == SXD_OK) rc = launch_mortar_pack(gh, d->plan, sc.temps, send, time_step_id, s);
  if (rc != SXD_OK) return finish(rc, __func__);
  sc.begun = true;
  sc.phase_system = system;
  sc.step = time_step_id;
  sc.filled.assign(d->plan.num_remote, 0);
  return ok();
}

int sxd_boundary_messages_unpack(const sxd_domain* domain, int system, int sender_rank,
                                 const void* recv, int64_t bytes, void* stream) {
  if (int rc = check_domain(domain); rc != SXD_OK) return rc;
  if (!domain->mortars) return fail(SXD_ERR_UNSUPPORTED, "%s: not a mortar domain", __func__);
  sxd_domain* d = const_cast<sxd_domain*>(domain);
  hipStream_t s = static_cast<hipStream_t>(stream);
  Scratch& sc = stream_scratch(d, s);
  if (!sc.begun || sc.phase_system != system)
    return fail(SXD_ERR_INVALID_ARGUMENT, "%s: no sxd_rhs_begin of this system on this stream",
                __func__);
  const auto it = d->plan.expected.find(sender_rank);
  const int count = it == d->plan.expected.end() ? 0 : it->second;
  if (count == 0)
    return fail(SXD_ERR_INVALID_ARGUMENT, "%s: no message is expected from rank %d", __func__,
                sender_rank);
  SXD_REQUIRE_DEVICE(recv);
  if (bytes < directory_bytes(count))
    return fail(SXD_ERR_INVALID_ARGUMENT, "%s: %lld bytes cannot hold %d messages", __func__,
                static_cast<long long>(bytes), count);
  if (sc.status_n < 2 * static_cast<size_t>(count)) {
    if (sc.d_status != nullptr) (void)hipFree(sc.d_status);
    sc.d_status = nullptr;
    sc.status_n = 0;
    if (hipMalloc(&sc.d_status, 2 * static_cast<size_t>(count) * sizeof(int32_t)) != hipSuccess)
      return fail(SXD_ERR_OUT_OF_MEMORY, "%s: status array", __func__);
    sc.status_n = 2 * static_cast<size_t>(count);
  }
  int rc = launch_mortar_unpack(system == SXD_SYSTEM_GH, d->plan, sc.temps, recv, bytes, count,
                                sc.step, sc.d_status, s);
  if (rc != SXD_OK) return finish(rc, __func__);
  std::vector<int32_t> status(2 * static_cast<size_t>(count));
  if (hipStreamSynchronize(s) != hipSuccess ||
      hipMemcpy(status.data(), sc.d_status, status.size() * sizeof(int32_t),
                hipMemcpyDeviceToHost) != hipSuccess)
    return fail(SXD_ERR_DEVICE, "%s: reading the message status failed", __func__);
  static const char* const reasons[] = {"",
                                        "the group's message count is not the expected one",
                                        "a message extends past the buffer",
                                        "a header has no message magic",
                                        "a message is for another system",
                                        "a message carries another time step id",
                                        "a message is for a mortar this rank does not have",
                                        "a message's mortar mesh differs from the receiver's"};
  for (int i = 0; i < count; ++i) {
    const int code = status[2 * i], slot = status[2 * i + 1];
    if (code != 0)
      return fail(SXD_ERR_INVALID_ARGUMENT, "%s: message %d from rank %d: %s", __func__, i,
                  sender_rank, reasons[code]);
    if (sc.filled[slot] != 0)
      return fail(SXD_ERR_INVALID_ARGUMENT, "%s: message %d from rank %d: mortar already received",
                  __func__, i, sender_rank);
    sc.filled[slot] = 1;
  }
  return ok();
}

int sxd_rhs_end(const sxd_domain* domain, int system, double* dt, void* stream) {
  if (int rc = check_domain(domain); rc != SXD_OK) return rc;
  if (!domain->mortars) return fail(SXD_ERR_UNSUPPORTED, "%s: not a mortar domain", __func__);
  SXD_REQUIRE_DEVICE(dt);
  sxd_domain* d = const_cast<sxd_domain*>(domain);
  hipStream_t s = static_cast<hipStream_t>(stream);
  Scratch& sc = stream_scratch(d, s);
  if (!sc.begun || sc.phase_system != system)
    return fail(SXD_ERR_INVALID_ARGUMENT, "%s: no sxd_rhs_begin of this system on this stream",
                __func__);
  for (size_t i = 0; i < sc.filled.size(); ++i)
    if (sc.filled[i] == 0)
      return fail(SXD_ERR_INVALID_ARGUMENT, "%s: %zu remote mortars still lack their message",
                  __func__, static_cast<size_t>(std::count(sc.filled.begin(), sc.filled.end(), 0)));
  sc.begun = false;
  return finish(launch_mortar_corrections(system == SXD_SYSTEM_GH, d->plan, d->d_el, dt, sc.temps,
                                          s),
                __func__);
}

int sxd_gh_bjorhus_corrections(
    int bjorhus_type, int64_t num_points, const double* normal_covector, const double* spacetime_metric,
    const double* pi, const double* phi, const double* coords, const double* gamma1,
    const double* gamma2, const double* lapse, const double* shift,
    const double* inverse_spacetime_metric, const double* spacetime_unit_normal_vector,
    const double* three_index_constraint, const double* gauge_source,
    const double* spacetime_deriv_gauge_source, const double* dt_spacetime_metric,
    const double* dt_pi, const double* dt_phi, const double* d_pi, const double* d_phi,
    double* dt_spacetime_metric_correction, double* dt_pi_correction, double* dt_phi_correction,
    void* stream) {
  if (num_points < 0) return fail(SXD_ERR_INVALID_ARGUMENT, "negative point count");
  if (bjorhus_type != SXD_BJORHUS_CONSTRAINT_PRESERVING &&
      bjorhus_type != SXD_BJORHUS_CONSTRAINT_PRESERVING_PHYSICAL)
    return fail(SXD_ERR_INVALID_ARGUMENT, "%s: unknown Bjorhus type %d", __func__, bjorhus_type);
  const double* in[19] = {normal_covector, spacetime_metric, pi, phi, coords, gamma1, gamma2,
                          lapse, shift, inverse_spacetime_metric, spacetime_unit_normal_vector,
                          three_index_constraint, gauge_source, spacetime_deriv_gauge_source,
                          dt_spacetime_metric, dt_pi, dt_phi, d_pi, d_phi};
  double* out[3] = {dt_spacetime_metric_correction, dt_pi_correction, dt_phi_correction};
  for (const double* p : in) SXD_REQUIRE_DEVICE(p);
  for (double* p : out) SXD_REQUIRE_DEVICE(p);
  return finish(launch_gh_bjorhus_points(
                    num_points, bjorhus_type == SXD_BJORHUS_CONSTRAINT_PRESERVING_PHYSICAL, in,
                    out, static_cast<hipStream_t>(stream)),
                __func__);
}

int sxd_gh_rhs_bjorhus(const sxd_domain* domain, int bjorhus_type, const double* u,
                       const double* gammas, const double* inv_jac, const double* coords,
                       double* dt, void* stream) {
  if (int rc = check_domain(domain); rc != SXD_OK) return rc;
  if (bjorhus_type != SXD_BJORHUS_CONSTRAINT_PRESERVING &&
      bjorhus_type != SXD_BJORHUS_CONSTRAINT_PRESERVING_PHYSICAL)
    return fail(SXD_ERR_INVALID_ARGUMENT, "%s: unknown Bjorhus type %d", __func__, bjorhus_type);
  if (domain->mortars)
    return fail(SXD_ERR_UNSUPPORTED, "%s: mortar domains have no external faces", __func__);
  const bool physical = bjorhus_type == SXD_BJORHUS_CONSTRAINT_PRESERVING_PHYSICAL;
  SXD_REQUIRE_DEVICE(u);
  SXD_REQUIRE_DEVICE(gammas);
  SXD_REQUIRE_DEVICE(inv_jac);
  SXD_REQUIRE_DEVICE(coords);
  SXD_REQUIRE_DEVICE(dt);
  sxd_domain* d = const_cast<sxd_domain*>(domain);
  if (int rc = check_no_overlap(d, 50, u, dt, __func__); rc != SXD_OK) return rc;
  Scratch& sc = stream_scratch(d, static_cast<hipStream_t>(stream));
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (d->mixed) {
    // the mixed-p kernels with their face passes (the mortar organisation
    // has no external faces): volume terms, the condition on the external
    // faces, then the internal faces
    if (int rc = reserve_scratch(sc, 150 * static_cast<size_t>(d->total_points), 0); rc != SXD_OK)
      return finish(rc, __func__);
    const int threads = d->threads[SXD_SYSTEM_GH];
    int rc = launch_mixed_volume(true, d->num_elements, d->d_list, d->d_el, d->np_max, u, gammas,
                                 inv_jac, dt, sc.du, threads, s);
    if (rc == SXD_OK)
      rc = launch_gh_bjorhus_mixed_faces(d->num_elements, physical, d->d_list, d->d_el, d->d_nbr,
                                         u, sc.du, gammas, inv_jac, coords, dt, s);
    if (rc == SXD_OK)
      rc = launch_mixed_gh_face_passes(d->num_elements, d->d_list, d->d_el, d->d_nbr, d->kf,
                                       d->km, u, gammas, inv_jac, dt, threads, nullptr,
                                       kFacesInternal, s);
    return finish(rc, __func__);
  }
  const size_t np = static_cast<size_t>(d->ext[0]) * d->ext[1] * d->ext[2];
  if (int rc = reserve_scratch(sc, 150 * np * d->num_elements, 12 * np * d->num_elements);
      rc != SXD_OK)
    return finish(rc, __func__);
  return finish(launch_gh_rhs_bjorhus(d->num_elements, physical, d->ext, d->d_nbr, u, gammas,
                                      inv_jac, coords, dt, sc.du, sc.temps, s),
                __func__);
}

int sxd_domain_set_kernel(sxd_domain* domain, int system, int kernel) {
  if (int rc = check_domain(domain); rc != SXD_OK) return rc;
  if (system != SXD_SYSTEM_SCALAR_WAVE && system != SXD_SYSTEM_GH)
    return fail(SXD_ERR_INVALID_ARGUMENT, "unknown system %d", system);
  if (kernel != SXD_KERNEL_AUTO) {
    int k[3];
    const int n = candidate_kernels(domain, system, k);
    bool found = false;
    for (int i = 0; i < n; ++i) found = found || k[i] == kernel;
    if (!found)
      return fail(SXD_ERR_UNSUPPORTED, "kernel %d is not available for these extents", kernel);
  }
  domain->kernel[system] = kernel;
  return ok();
}

int sxd_domain_get_kernel(const sxd_domain* domain, int system, int* kernel) {
  if (int rc = check_domain(domain); rc != SXD_OK) return rc;
  if (system != SXD_SYSTEM_SCALAR_WAVE && system != SXD_SYSTEM_GH)
    return fail(SXD_ERR_INVALID_ARGUMENT, "unknown system %d", system);
  if (kernel == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null output");
  *kernel = resolve_kernel(domain, system);
  return ok();
}

int sxd_domain_autotune(sxd_domain* domain, int system, const double* u, const double* params,
                        const double* inv_jac, double* dt, int repetitions, void* stream,
                        int* chosen, float* milliseconds) {
  if (int rc = check_periodic(domain); rc != SXD_OK) return rc;
  if (system != SXD_SYSTEM_SCALAR_WAVE && system != SXD_SYSTEM_GH)
    return fail(SXD_ERR_INVALID_ARGUMENT, "unknown system %d", system);
  if (repetitions < 1) return fail(SXD_ERR_INVALID_ARGUMENT, "repetitions must be positive");
  SXD_REQUIRE_DEVICE(u);
  SXD_REQUIRE_DEVICE(params);
  SXD_REQUIRE_DEVICE(inv_jac);
  SXD_REQUIRE_DEVICE(dt);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (int rc = check_no_overlap(domain, system == SXD_SYSTEM_GH ? 50 : 5, u, dt, __func__);
      rc != SXD_OK)
    return rc;
  Scratch& sc = stream_scratch(domain, s);
  hipEvent_t ev[2];
  if (hipEventCreate(&ev[0]) != hipSuccess) return fail(SXD_ERR_DEVICE, "hipEventCreate failed");
  if (hipEventCreate(&ev[1]) != hipSuccess) {
    (void)hipEventDestroy(ev[0]);
    return fail(SXD_ERR_DEVICE, "hipEventCreate failed");
  }
  int best = -1;
  float best_ms = 0.0f;
  int rc = SXD_OK;
  int cand[3];
  const int ncand = candidate_kernels(domain, system, cand);
  const int saved_threads = domain->threads[system];
  int best_threads = saved_threads;
  // tile shapes: mixed p 64/128/256 threads; uniform GH fused 256/512
  const int mixed_tiles[3] = {64, 128, 256};
  const int fused_tiles[2] = {256, 512};
  const bool gh_fused_tiles = !domain->mixed && system == SXD_SYSTEM_GH && domain->ext[0] >= 7;
  const int ntiles = domain->mixed ? 3 : gh_fused_tiles ? 2 : 1;
  for (int cj = 0; cj < ncand * ntiles; ++cj) {
    const int ci = cj / ntiles;
    const int k = cand[ci];
    if (domain->mixed) domain->threads[system] = mixed_tiles[cj % ntiles];
    if (gh_fused_tiles) domain->threads[system] = fused_tiles[cj % ntiles];
    rc = run_system(domain, sc, system, k, u, params, inv_jac, dt, s);  // warm-up
    if (rc != SXD_OK) break;
    if (hipEventRecord(ev[0], s) != hipSuccess) { rc = SXD_ERR_DEVICE; break; }
    for (int r = 0; r < repetitions && rc == SXD_OK; ++r)
      rc = run_system(domain, sc, system, k, u, params, inv_jac, dt, s);
    if (rc != SXD_OK) break;
    float ms = 0.0f;
    if (hipEventRecord(ev[1], s) != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess ||
        hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess) {
      rc = SXD_ERR_DEVICE;
      break;
    }
    ms /= static_cast<float>(repetitions);
    if (best < 0 || ms < best_ms) {
      best = k;
      best_ms = ms;
      best_threads = domain->threads[system];
    }
  }
  domain->threads[system] = rc == SXD_OK ? best_threads : saved_threads;
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (rc != SXD_OK) return finish(rc, __func__);
  domain->kernel[system] = best;
  if (chosen != nullptr) *chosen = best;
  if (milliseconds != nullptr) *milliseconds = best_ms;
  return ok();
}

// ---- adjoints ------------------------------------------------------------------

int sxd_scalar_wave_rhs_adjoint(const sxd_domain* domain, const double* gbar, const double* gamma2,
                                const double* inv_jac, double* ubar, void* stream) {
  if (int rc = check_periodic(domain); rc != SXD_OK) return rc;
  SXD_REQUIRE_DEVICE(gbar);
  SXD_REQUIRE_DEVICE(gamma2);
  SXD_REQUIRE_DEVICE(inv_jac);
  SXD_REQUIRE_DEVICE(ubar);
  if (domain->mixed || !sw_adjoint_supported(domain->ext))
    return fail(SXD_ERR_UNSUPPORTED, "the ScalarWave adjoint needs a uniform domain with "
                                     "isotropic extents 2..10");
  if (gbar == ubar) return fail(SXD_ERR_INVALID_ARGUMENT, "gbar and ubar must not alias");
  return finish(launch_sw_rhs_adjoint(domain->num_elements, domain->ext, domain->d_nbr, gbar,
                                      gamma2, inv_jac, ubar, static_cast<hipStream_t>(stream)),
                __func__);
}

int sxd_gh_rhs_adjoint(const sxd_domain* domain, const double* u, const double* gammas,
                       const double* inv_jac, const double* gbar, double* ubar, void* stream) {
  if (int rc = check_periodic(domain); rc != SXD_OK) return rc;
  SXD_REQUIRE_DEVICE(u);
  SXD_REQUIRE_DEVICE(gammas);
  SXD_REQUIRE_DEVICE(inv_jac);
  SXD_REQUIRE_DEVICE(gbar);
  SXD_REQUIRE_DEVICE(ubar);
  if (gbar == ubar || u == ubar) return fail(SXD_ERR_INVALID_ARGUMENT, "ubar must not alias");
  if (domain->mortars)
    return fail(SXD_ERR_UNSUPPORTED, "%s: not built for mortar domains", __func__);
  sxd_domain* d = const_cast<sxd_domain*>(domain);
  Scratch& sc = stream_scratch(d, static_cast<hipStream_t>(stream));
  if (d->mixed) {
    if (int rc = reserve_scratch(sc, 150 * static_cast<size_t>(d->total_points), 0);
        rc != SXD_OK)
      return finish(rc, __func__);
    return finish(launch_mixed_gh_adjoint(d->num_elements, d->d_list, d->d_el, d->d_point_el,
                                          d->total_points, d->d_nbr,
                                          d->np_max, d->kf, d->km, u, gammas, inv_jac, gbar, ubar,
                                          sc.du, static_cast<hipStream_t>(stream)),
                  __func__);
  }
  const size_t np = static_cast<size_t>(d->ext[0]) * d->ext[1] * d->ext[2];
  const size_t face_doubles = 300 * static_cast<size_t>(gh_adjoint_face_slots(d->ext));
  if (int rc = reserve_scratch(sc, 150 * np * d->num_elements, face_doubles * d->num_elements);
      rc != SXD_OK)
    return finish(rc, __func__);
  return finish(launch_gh_rhs_adjoint(d->num_elements, d->ext, d->d_nbr, u, gammas, inv_jac, gbar,
                                      ubar, sc.du, sc.temps, static_cast<hipStream_t>(stream)),
                __func__);
}

int sxd_rk3_adjoint_substep(int substep, int64_t n, double dt, const double* lam,
                            const double* a, double* mu, double* lam_out, void* stream) {
  if (substep < 0 || substep > 2) return fail(SXD_ERR_INVALID_ARGUMENT, "substep must be 0..2");
  if (n < 0) return fail(SXD_ERR_INVALID_ARGUMENT, "negative size");
  SXD_REQUIRE_DEVICE(lam);
  SXD_REQUIRE_DEVICE(a);
  SXD_REQUIRE_DEVICE(mu);
  if (substep > 0) SXD_REQUIRE_DEVICE(lam_out);
  return finish(launch_rk3_adjoint_substep(substep, n, dt, lam, a, mu, lam_out,
                                           static_cast<hipStream_t>(stream)),
                __func__);
}

// ---- RK3 step captured in a HIP graph ---------------------------------------

struct sxd_rk3_stepper {
  sxd_domain* domain = nullptr;
  int system = 0;
  int64_t n = 0;                 // doubles per field
  double* u = nullptr;           // caller's state, updated in place
  double* d_stage = nullptr;     // u1, then u2 (in place)
  double* d_rhs = nullptr;       // L0, L1, L2
  double* d_dt = nullptr;        // step size read by the update kernels
  double* h_dt = nullptr;        // pinned host copy of the step size
  sxd::Scratch scratch;           // the stepper's own work arrays
  hipStream_t capture = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {
void destroy_stepper(sxd_rk3_stepper* s) {
  if (s->exec != nullptr) (void)hipGraphExecDestroy(s->exec);
  if (s->graph != nullptr) (void)hipGraphDestroy(s->graph);
  if (s->capture != nullptr) (void)hipStreamDestroy(s->capture);
  if (s->d_stage != nullptr) (void)hipFree(s->d_stage);
  if (s->d_rhs != nullptr) (void)hipFree(s->d_rhs);
  if (s->d_dt != nullptr) (void)hipFree(s->d_dt);
  if (s->h_dt != nullptr) (void)hipHostFree(s->h_dt);
  s->scratch.release();
  delete s;
}

// The three RHS evaluations and update_u substeps of one
// Rk3HesthavenSsp step (Time/TimeSteppers/Rk3HesthavenSsp.cpp:55-81):
//   L0 = RHS(u);  u1 = u + dt L0
//   L1 = RHS(u1); u2 = (3 u + u1 + dt L1) / 4     (written over u1)
//   L2 = RHS(u2); u  = (u + 2 u2 + 2 dt L2) / 3   (written over u)
int enqueue_rk3_step(sxd_rk3_stepper* s, const double* params, const double* inv_jac,
                     hipStream_t stream) {
  const int kernel = resolve_kernel(s->domain, s->system);
  const double* stage_in[3] = {s->u, s->d_stage, s->d_stage};
  double* stage_out[3] = {s->d_stage, s->d_stage, s->u};
  for (int k = 0; k < 3; ++k) {
    int rc = run_system(s->domain, s->scratch, s->system, kernel, stage_in[k], params, inv_jac,
                        s->d_rhs, stream);
    if (rc != SXD_OK) return rc;
    rc = launch_rk3_substep_dt(k, s->n, s->d_dt, s->u, s->d_stage, s->d_rhs, stage_out[k],
                               stream);
    if (rc != SXD_OK) return rc;
  }
  return SXD_OK;
}
}  // namespace

int sxd_rk3_stepper_create(sxd_rk3_stepper** stepper, sxd_domain* domain, int system, double* u,
                           const double* params, const double* inv_jac) {
  if (stepper == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null stepper output");
  *stepper = nullptr;
  if (int rc = check_periodic(domain); rc != SXD_OK) return rc;
  if (system != SXD_SYSTEM_SCALAR_WAVE && system != SXD_SYSTEM_GH)
    return fail(SXD_ERR_INVALID_ARGUMENT, "unknown system %d", system);
  SXD_REQUIRE_DEVICE(u);
  SXD_REQUIRE_DEVICE(params);
  SXD_REQUIRE_DEVICE(inv_jac);
  int64_t points = 0;
  sxd_domain_num_points(domain, &points);
  sxd_rk3_stepper* s = new (std::nothrow) sxd_rk3_stepper;
  if (s == nullptr) return fail(SXD_ERR_OUT_OF_MEMORY, "host allocation failed");
  s->domain = domain;
  s->system = system;
  s->n = points * (system == SXD_SYSTEM_GH ? 50 : 5);
  s->u = u;
  const size_t bytes = static_cast<size_t>(s->n) * sizeof(double);
  if (hipMalloc(&s->d_stage, bytes) != hipSuccess || hipMalloc(&s->d_rhs, bytes) != hipSuccess ||
      hipMalloc(&s->d_dt, sizeof(double)) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&s->h_dt), sizeof(double)) != hipSuccess) {
    destroy_stepper(s);
    return fail(SXD_ERR_OUT_OF_MEMORY, "stepper buffers");
  }
  if (hipStreamCreateWithFlags(&s->capture, hipStreamNonBlocking) != hipSuccess) {
    destroy_stepper(s);
    return fail(SXD_ERR_DEVICE, "stream creation failed");
  }
  // One RHS outside the capture allocates the stepper's own work arrays and
  // sets kernel attributes, so the captured region only enqueues kernels.
  // The graph refers to those arrays only: calls on the domain afterwards
  // (which may grow the domain's own work arrays) cannot move them.
  int rc = run_system(domain, s->scratch, system, resolve_kernel(domain, system), u, params,
                      inv_jac, s->d_rhs, s->capture);
  if (rc == SXD_OK && hipStreamSynchronize(s->capture) != hipSuccess) rc = SXD_ERR_DEVICE;
  if (rc != SXD_OK) {
    destroy_stepper(s);
    return finish(rc, __func__);
  }
  if (hipStreamBeginCapture(s->capture, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    destroy_stepper(s);
    return fail(SXD_ERR_DEVICE, "hipStreamBeginCapture failed");
  }
  rc = enqueue_rk3_step(s, params, inv_jac, s->capture);
  const hipError_t end = hipStreamEndCapture(s->capture, &s->graph);
  if (rc != SXD_OK || end != hipSuccess ||
      hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0) != hipSuccess) {
    destroy_stepper(s);
    return fail(rc != SXD_OK ? rc : SXD_ERR_DEVICE, "graph capture of the RK3 step failed");
  }
  *stepper = s;
  return ok();
}

int sxd_rk3_stepper_step(sxd_rk3_stepper* stepper, double dt, int num_steps, void* stream) {
  if (stepper == nullptr) return fail(SXD_ERR_INVALID_ARGUMENT, "null stepper");
  if (num_steps < 0) return fail(SXD_ERR_INVALID_ARGUMENT, "negative step count");
  hipStream_t st = static_cast<hipStream_t>(stream);
  // the update kernels read the step size from device memory
  if (hipStreamSynchronize(st) != hipSuccess) return fail(SXD_ERR_DEVICE, "stream sync failed");
  *stepper->h_dt = dt;
  if (hipMemcpyAsync(stepper->d_dt, stepper->h_dt, sizeof(double), hipMemcpyHostToDevice, st) !=
      hipSuccess)
    return fail(SXD_ERR_DEVICE, "step size upload failed");
  for (int k = 0; k < num_steps; ++k)
    if (hipGraphLaunch(stepper->exec, st) != hipSuccess)
      return fail(SXD_ERR_DEVICE, "hipGraphLaunch failed");
  return ok();
}

int sxd_rk3_stepper_destroy(sxd_rk3_stepper* stepper) {
  if (stepper != nullptr) {
    (void)hipDeviceSynchronize();
    destroy_stepper(stepper);
  }
  return ok();
}

int sxd_rk3_substep(int substep, int64_t n, double dt, const double* u0, const double* uk,
                    const double* lk, double* out, void* stream) {
  if (substep < 0 || substep > 2)
    return fail(SXD_ERR_INVALID_ARGUMENT, "substep must be 0, 1 or 2");
  if (n < 0) return fail(SXD_ERR_INVALID_ARGUMENT, "negative size");
  if (n == 0) return ok();
  SXD_REQUIRE_DEVICE(u0);
  if (substep > 0) SXD_REQUIRE_DEVICE(uk);
  SXD_REQUIRE_DEVICE(lk);
  SXD_REQUIRE_DEVICE(out);
  return finish(launch_rk3_substep(substep, n, dt, u0, substep > 0 ? uk : u0, lk, out,
                                   static_cast<hipStream_t>(stream)),
                __func__);
}

}  // extern "C"
