// Host-side launchers of the device kernels (defined in kernels.hip and
// gh_fused.hip).  Arguments are validated by the C ABI layer (capi.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/spectre_amd.h"

namespace sxd {

// Enables `bytes` of dynamic LDS for `kernel` on the current device (needed
// beyond 64 KiB; done once per device and kernel).  False if it cannot be.
bool allow_lds(const void* kernel, int bytes);
// dtab: packed 1-d differentiation matrices (packed_d_offset); null = the
// LGL table.
int launch_partial_derivatives(int num_elements, const int32_t ext[3], int nc,
                               const double* u, const double* inv_jac, double* du,
                               double* dlogical, hipStream_t stream,
                               const double* dtab = nullptr);
// ScalarWave point-wise terms of a batch ([E][5][Np] u and dt, [E][15][Np] du,
// [E][Np] gamma2; total = E Np).
int launch_sw_volume_batched(int64_t np, int64_t total, const double* u, const double* du,
                             const double* gamma2, double* dt, hipStream_t stream);
// ScalarWave RHS on Legendre-Gauss points (gauss.hip); scratch_du [E][15][Np];
// face_inv_jac per element: the six faces in direction order, each 3 rows of
// its points' d xi^d / d x^i.
const double* device_gauss_tables();
int launch_sw_rhs_gauss(int num_elements, const int32_t ext[3], const int32_t* nbr,
                        const double* u, const double* gamma2, const double* inv_jac,
                        const double* det_inv_jac, const double* face_inv_jac, double* dt,
                        double* scratch_du, hipStream_t stream);
// ubar [E][nc][Np] = transpose of the partial derivatives applied to dubar
// [E][3 nc][Np]; SXD_ERR_UNSUPPORTED if 3 Np doubles plus the D matrices
// exceed the LDS.
int launch_partial_derivatives_adjoint(int num_elements, const int32_t ext[3], int nc,
                                       const double* dubar, const double* inv_jac,
                                       double* ubar, hipStream_t stream);
int launch_sw_time_derivative(int64_t n, const double* u, const double* du,
                              const double* gamma2, double* dt, double* gamma2_temp,
                              hipStream_t stream);
int launch_gh_time_derivative(int64_t n, const double* u, const double* du,
                              const double* gammas, double* dt, double* temps,
                              hipStream_t stream);
int launch_sw_upwind(int64_t n, const double* u_int, const double* g2_int,
                     const double* n_int, const double* u_ext, const double* g2_ext,
                     const double* n_ext, double* bc, hipStream_t stream);
int launch_gh_upwind(int64_t n, const double* u_int, const double* t_int,
                     const double* n_int, const double* u_ext, const double* t_ext,
                     const double* n_ext, double* bc, hipStream_t stream);
int launch_lift_flux(int64_t n, int nc, double* bc, int n_perp, const double* mag,
                     hipStream_t stream);
bool sw_fused_supported(const int32_t ext[3]);
// Which faces a face-pass launch handles: all of them (no external faces), or
// the external or internal ones of a domain with external faces (external
// faces are corrected first).
constexpr int kFacesAll = 0, kFacesExternal = 1, kFacesInternal = 2;
// Exterior state of external ScalarWave faces: the caller's ghost array
// (DirichletAnalytic), or SphericalRadiation of the interior trace.
constexpr int kSwBcGhostArray = 0, kSwBcSommerfeld = 1, kSwBcBaylissTurkel = 2;
// ConstraintPreservingSphericalRadiation (a time-derivative condition: no
// upwind term on the face, dt Pi += correction); ghost = coordinates for the
// Bayliss-Turkel types.
constexpr int kSwBcCpsrSommerfeld = 3, kSwBcCpsrFirstOrderBt = 4, kSwBcCpsrSecondOrderBt = 5;
// GH constraint-preserving Bjorhus (a time-derivative condition, no upwind
// term on the external faces).
constexpr int kGhBcBjorhus = 10;
// fused: one launch per RHS (requires sw_fused_supported); otherwise the
// composite path, which needs scratch_du [E][15][Np].  Faces whose neighbour
// index is negative take their exterior state by bc_kind: kSwBcGhostArray
// ghost = [E][5][Np] state read at the face's own points; kSwBcSommerfeld
// ghost unused; kSwBcBaylissTurkel ghost = inertial coordinates [E][3][Np].
// ghost may be null when no neighbour is negative.
int launch_sw_rhs(int num_elements, const int32_t ext[3], const int32_t* nbr,
                  const double* u, const double* gamma2, const double* inv_jac, double* dt,
                  bool fused, double* scratch_du, const double* ghost, int bc_kind,
                  hipStream_t stream);
// ghost: as for launch_sw_rhs, with [E][50][Np].
int launch_gh_rhs_composite(int num_elements, const int32_t ext[3], const int32_t* nbr,
                            const double* u, const double* gammas, const double* inv_jac,
                            double* dt, double* scratch_du, double* scratch_temps,
                            const double* ghost, hipStream_t stream);
// Derivative engines of the fused GH kernel.
constexpr int kGhFusedValu = 0;  // LDS pencils on the vector ALUs
constexpr int kGhFusedMfma = 1;  // v_mfma_f64_16x16x4f64 tiles
bool gh_fused_supported(const int32_t ext[3]);
// threads: 256 or 512 per workgroup (512 applies to N = 7, 8).  ghost: as for
// launch_gh_rhs_composite; the MFMA engine takes none (SXD_ERR_UNSUPPORTED).
int launch_gh_rhs_fused(int num_elements, const int32_t ext[3], const int32_t* nbr,
                        const double* u, const double* gammas, const double* inv_jac,
                        double* dt, int variant, int threads, const double* ghost,
                        hipStream_t stream);
// Mixed p.  Per-element record of the ragged layout.
struct MixedElement {
  int32_t n[3];  // LGL extents
  int32_t np;    // n0 n1 n2
  int64_t off;   // point offset of the element
};
const double* device_mortar_matrices();
// LDS bytes of the face kernel for faces of <= kf and mortars of <= km points.
int mixed_face_lds_bytes(int kf, int km);
// Volume terms and mortar-coupled boundary corrections of the `count`
// elements in `list` (one workgroup each, `threads` threads: 64, 128, 256).
// gh: params = gammas [3][Np_e], scratch_du = 150 x total points; otherwise
// params = gamma2 [Np_e].  Faces whose neighbour is -1 take their exterior
// state by bc_kind: kSwBcGhostArray from ghost (ragged [nc][Np_e], read at the
// face's own points); for ScalarWave also kSwBcSommerfeld (ghost unused) and
// kSwBcBaylissTurkel (ghost = ragged [3][Np_e] coordinates).
int launch_mixed_rhs(bool gh, int count, const int32_t* list, const MixedElement* el,
                     const int32_t* nbr, int np_max, int kf, int km, const double* u,
                     const double* params, const double* inv_jac, double* dt, double* scratch_du,
                     int threads, hipStream_t stream, const double* ghost = nullptr,
                     int bc_kind = kSwBcGhostArray);
// The volume terms alone of launch_mixed_rhs (dt = point-wise terms).
int launch_mixed_volume(bool gh, int count, const int32_t* list, const MixedElement* el,
                        int np_max, const double* u, const double* params,
                        const double* inv_jac, double* dt, double* scratch_du, int threads,
                        hipStream_t stream);
// The GH face corrections alone of launch_mixed_rhs, one launch per dimension,
// for the faces that `phase` selects (kFacesAll, kFacesExternal with the ghost
// state, kFacesInternal).
int launch_mixed_gh_face_passes(int count, const int32_t* list, const MixedElement* el,
                                const int32_t* nbr, int kf, int km, const double* u,
                                const double* params, const double* inv_jac, double* dt,
                                int threads, const double* ghost, int phase,
                                hipStream_t stream);
// Non-conforming mortars (mortar.hip): rotated neighbours and h-refined faces.
// Spectral::ChildSize (NumericalAlgorithms/Spectral/Projection.hpp:22-30).
constexpr int kChildFull = 1, kChildUpperHalf = 2, kChildLowerHalf = 3;
const double* device_half_mortar_matrices();
// A face of an element with at least one mortar: its records are
// [rec_begin, rec_begin + rec_count) of the record table; pt_off is the
// offset of its points among all faces' points.
struct MortarFace {
  int32_t e, dir, nf, rec_begin, rec_count;
  int64_t pt_off;
};
// One side of a mortar.  na, nb: own face extents; ka, kb: mortar extents in
// the own frame; size_a, size_b: ChildSize of the mortar on the own face;
// mortar_off: offset of its own projected package among the projected
// mortars' points (-1: the mortar is the face); rec_off: offset of its
// corrections and gather indices among all records' mortar points.
// partner < 0: the partner is in another partition (rank); its package
// arrives as a boundary message into the receive slot -1 - partner, at
// recv_off among the received mortar points.
struct MortarRecord {
  int32_t face, partner, na, nb, ka, kb, size_a, size_b;
  int64_t mortar_off, rec_off, recv_off;
};
// Boundary messages (the BoundaryMessage of
// Evolution/DiscontinuousGalerkin/Messages/BoundaryMessage.hpp for DG mortar
// data).  The send buffer holds one group per receiving rank; a group is a
// directory (int64 count, int64 byte offset of each message within the
// group, padded to 64 bytes) followed by its messages; a message is a
// 64-byte header of 16 int32 and the packaged mortar data (doubles, padded to
// 64 bytes), oriented into the receiver's mortar point order.  Header:
//   0 magic, 1 system, 2 sender element, 3 receiver element,
//   4 receiver direction, 5 sender direction, 6-8 sender volume extents,
//   9-10 mortar extents in the receiver's frame, 11 packaged components,
//   12 data doubles, 13 reserved, 14-15 time step id (low, high 32 bits).
constexpr int32_t kMessageMagic = 0x53584442;  // "SXDB"
constexpr int64_t kMessageHeaderBytes = 64;
__host__ __device__ inline int64_t message_data_bytes(int64_t doubles) {
  return (8 * doubles + 63) / 64 * 64;
}
__host__ __device__ inline int64_t message_data_bytes(bool gh, int64_t km) {
  return message_data_bytes((gh ? 134 : 13) * km);
}
__host__ __device__ inline int64_t directory_bytes(int64_t count) {
  return (8 * (count + 1) + 63) / 64 * 64;
}
// status codes of a received message
constexpr int kMessageWrongCount = 1, kMessageTruncated = 2, kMessageBadMagic = 3,
              kMessageWrongSystem = 4, kMessageWrongStep = 5, kMessageUnknownMortar = 6,
              kMessageWrongMesh = 7;
struct MortarMessage {
  int32_t record, km, peer, index, group_count, pad;
  int64_t group_offset[2], offset[2];  // per system: SW, GH
  int32_t header[16];
};
struct MortarRecvKey {
  int32_t recv_e, recv_dir, send_e, slot, ka, kb, send_rank, pad;
  int64_t recv_off;
};
struct MortarPeer {
  int rank = 0, count = 0;
  int64_t offset[2] = {0, 0}, bytes[2] = {0, 0};
};
struct MortarPlan {
  int num_faces = 0, num_records = 0, num_projected = 0, num_remote = 0;
  // faces whose single mortar differs from the face (at most kDirectPoints
  // points) are packaged and projected in one kernel; the others are
  // packaged on the face, and their projected mortars (num_projected) are
  // projected from the face package
  int num_package_faces = 0, num_direct = 0;
  // largest face, ka nb, and ka na + kb nb (matrix doubles) among the direct faces
  int direct_nf = 0, direct_tmp = 1, direct_mat = 0;
  int64_t face_points = 0, projected_points = 0, record_points = 0, remote_points = 0;
  int64_t message_bytes[2] = {0, 0};  // the whole send buffer, per system
  std::vector<MortarPeer> peers;      // outgoing groups, ascending rank
  std::map<int, int> expected;        // messages expected from each rank
  int dim_begin[4] = {0, 0, 0, 0};  // faces of dimension d: [dim_begin[d], dim_begin[d+1])
  int kf = 0, km = 0;
  int tmp_max = 1;  // largest ka nb or na kb: doubles per component of a staged pass
  int mat_max = 0;  // largest ka na + kb nb: doubles of a record's two matrices
  MortarFace* d_faces = nullptr;
  MortarRecord* d_records = nullptr;
  int32_t* d_gather = nullptr;
  int32_t* d_projected = nullptr;
  int32_t* d_package_faces = nullptr;
  // boundary terms: one workgroup per local mortar (the record before its
  // partner) for both sides, one per record whose partner is remote
  int num_pairs = 0, num_remote_records = 0;
  int32_t* d_pairs = nullptr;
  int32_t* d_remote_records = nullptr;
  int32_t* d_direct = nullptr;
  MortarMessage* d_messages = nullptr;
  MortarRecvKey* d_keys = nullptr;
};
// Decode and check a record list ({e, dir, en, o0, o1, o2, size_a, size_b}
// per record) and upload its tables; SXD_ERR_INVALID_ARGUMENT with *error
// set on an inconsistent list.
// partition: the owning rank of every element (nullptr: all elements, rank
// ignored); the plan covers the elements of `rank`.
int build_mortar_plan(int num_elements, const int32_t* extents, int num_records,
                      const int32_t* records, const int32_t* partition, int rank,
                      MortarPlan* plan, std::string* error);
void destroy_mortar_plan(MortarPlan* plan);
size_t mortar_work_doubles(bool gh, const MortarPlan& plan);
// Boundary corrections of every mortar added into dt (after the volume
// terms); work: mortar_work_doubles(gh, plan) doubles.
int launch_mortar_faces(bool gh, const MortarPlan& plan, const MixedElement* el, const double* u,
                        const double* params, const double* inv_jac, double* dt, double* work,
                        hipStream_t stream);
// The two halves of launch_mortar_faces around a boundary-message exchange:
// the packages of every mortar, then (with the remote packages received) the
// corrections.
int launch_mortar_packages(bool gh, const MortarPlan& plan, const MixedElement* el,
                           const double* u, const double* params, const double* inv_jac,
                           double* work, hipStream_t stream);
int launch_mortar_corrections(bool gh, const MortarPlan& plan, const MixedElement* el,
                              double* dt, double* work, hipStream_t stream);
// The corrections in two steps: the boundary terms on every mortar (into the
// work array; they do not touch dt), then their projection and lift into dt.
int launch_mortar_boundary_terms(bool gh, const MortarPlan& plan, double* work,
                                 hipStream_t stream);
int launch_mortar_lift(bool gh, const MortarPlan& plan, const MixedElement* el, double* dt,
                       double* work, hipStream_t stream);
// Write every outgoing message into send (plan.message_bytes[gh] bytes).
int launch_mortar_pack(bool gh, const MortarPlan& plan, double* work, void* send, int64_t step,
                       hipStream_t stream);
// Unpack the `count` messages of one sending rank's group (bytes long) into
// the receive slots; status: 2 count int32 (code, slot) per message.
int launch_mortar_unpack(bool gh, const MortarPlan& plan, double* work, const void* recv,
                         int64_t bytes, int count, int64_t step, int32_t* status,
                         hipStream_t stream);
// Adjoint of the GH RHS on a mixed-p domain, linearised at u: ubar =
// (d dt / d u)^T gbar over the work list; scratch_du = 150 x total points.
int mixed_face_adjoint_lds_bytes(int kf, int km);
// point_el: the element of each of the total_points grid points.
int launch_mixed_gh_adjoint(int count, const int32_t* list, const MixedElement* el,
                            const int32_t* point_el, int64_t total_points,
                            const int32_t* nbr, int np_max, int kf, int km, const double* u,
                            const double* gammas, const double* inv_jac, const double* gbar,
                            double* ubar, double* scratch_du, hipStream_t stream);
// Adjoint (transpose) of the ScalarWave RHS on a uniform domain (isotropic
// extents 2..10) and of one RK3 substep.
bool sw_adjoint_supported(const int32_t ext[3]);
int launch_sw_rhs_adjoint(int num_elements, const int32_t ext[3], const int32_t* nbr,
                          const double* gbar, const double* gamma2, const double* inv_jac,
                          double* ubar, hipStream_t stream);
// Adjoint of the GH RHS linearised at u (uniform extents with
// 3 Np + N0^2 + N1^2 + N2^2 <= 20480):
// ubar = (d dt / d u)^T gbar; scratch_du: 150 x E x Np doubles; scratch_faces:
// 300 x E x gh_adjoint_face_slots(ext) doubles.
int gh_adjoint_face_slots(const int32_t ext[3]);
int launch_gh_rhs_adjoint(int num_elements, const int32_t ext[3], const int32_t* nbr,
                          const double* u, const double* gammas, const double* inv_jac,
                          const double* gbar, double* ubar, double* scratch_du,
                          double* scratch_faces, hipStream_t stream);
int launch_rk3_adjoint_substep(int substep, int64_t n, double dt, const double* lam,
                               const double* a, double* mu, double* lam_out, hipStream_t stream);
int launch_rk3_substep_dt(int substep, int64_t n, const double* dt, const double* u0,
                          const double* uk, const double* lk, double* out, hipStream_t stream);
int launch_rk3_substep(int substep, int64_t n, double dt, const double* u0, const double* uk,
                       const double* lk, double* out, hipStream_t stream);
// Flux systems (divergence.hip).  Strong (weak = false; jac = inverse
// Jacobian) or weak (jac = det J times the inverse Jacobian) divergence of
// fluxes [E][3 nc][Np] (block i + 3 c = F^i_c).  terms = false: out = div.
// terms = true: the volume terms of flux variables, on dt = out: with a mesh
// velocity v [E][3][Np] (may be null; div_v [E][Np] then required) the
// fluxes become F^i_c - u_c v^i (written back) and dt_c -= u_c div v, then
// dt -= div (strong) or dt += div det_inv_jac [E][Np] (weak).
// SXD_ERR_UNSUPPORTED if 3 Np doubles and the matrices exceed the LDS.
int launch_flux_divergence(int num_elements, const int32_t ext[3], int nc, bool weak,
                           double* fluxes, const double* jac, double* out, bool terms,
                           const double* u, const double* v, const double* div_v,
                           const double* det_inv_jac, hipStream_t stream);
int flux_divergence_lds_bytes(const int32_t ext[3]);
// dt_c += v^i d_i u_c on n points (du [3 nc][n], v [3][n], dt [nc][n]).
int launch_mesh_velocity_nonconservative(int64_t n, int nc, const double* du, const double* v,
                                         double* dt, hipStream_t stream);
const double* device_weak_flux_matrices();
// GH constraint-preserving Bjorhus condition (bjorhus.hip).  Point-wise: in =
// {normal [3], g [10], pi [10], phi [30], coords [3], gamma1, gamma2, lapse,
// shift [3], inverse spacetime metric [10], normal vector [4], three-index
// constraint [30], H [4], d H [16] (a + 4 b = d_a H_b), dt g [10], dt pi
// [10], dt phi [30], d pi [30], d phi [90]} (component blocks of n points),
// out = {dt g, dt pi, dt phi} corrections; every point is its own face.
// physical: the ConstraintPreservingPhysical type.
int launch_gh_bjorhus_points(int64_t n, bool physical, const double* const in[19],
                             double* const out[3], hipStream_t stream);
// The corrections of the external faces of a uniform domain added into dt
// (volume terms already there), one dimension after another.
int launch_gh_bjorhus_faces(int num_elements, bool physical, const int32_t ext[3],
                            const int32_t* nbr, const double* u, const double* du,
                            const double* gammas,
                            const double* inv_jac, const double* coords, double* dt,
                            hipStream_t stream);
// The corrections of the external faces of the `count` mixed-p elements in
// `list` (ragged arrays, du as launch_mixed_volume leaves it) added into dt,
// one dimension after another.
int launch_gh_bjorhus_mixed_faces(int count, bool physical, const int32_t* list,
                                  const MixedElement* el, const int32_t* nbr, const double* u,
                                  const double* du, const double* gammas, const double* inv_jac,
                                  const double* coords, double* dt, hipStream_t stream);
// The whole GH RHS of a uniform domain whose external faces carry the
// Bjorhus condition (composite kernels; scratch as launch_gh_rhs_composite).
int launch_gh_rhs_bjorhus(int num_elements, bool physical, const int32_t ext[3],
                          const int32_t* nbr, const double* u, const double* gammas,
                          const double* inv_jac,
                          const double* coords, double* dt, double* scratch_du,
                          double* scratch_temps, hipStream_t stream);
// The allocation-free entry point's fixed work array (noalloc.hip): its
// capacity in doubles, and its device address on the current device.
size_t noalloc_scratch_doubles();
int noalloc_scratch(double** ptr);

}  // namespace sxd
