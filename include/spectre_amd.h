/* spectre_amd: MI355X (gfx950) implementation of SpECTRE's DG volume
 * time-derivative path behind a C ABI.
 *
 * Conventions shared by every entry point
 *  - All field pointers are DEVICE pointers to FP64 data in SpECTRE's
 *    Variables layout: per element, nc contiguous blocks of Np doubles, grid
 *    index i + N0 (j + N1 k) (DataStructures/Index.hpp:177-183).  A batch of
 *    E elements of equal extents is E such buffers back to back.
 *  - Tensor components are in SpECTRE storage order
 *    (DataStructures/Tensor/Structure.hpp:98-195).
 *  - The inverse Jacobian is 9 blocks per element, block j + 3 i holding
 *    d xi^j / d x^i (NumericalAlgorithms/LinearOperators/PartialDerivatives.tpp:70-77).
 *  - `stream` is a hipStream_t passed as void* (NULL = default stream).  All
 *    device work is enqueued on it; calls return after enqueueing.
 *  - Return value: SXD_OK (0) or a negative status.  The reference reports
 *    the same conditions through ERROR/ASSERT; here the caller decides.
 *    sxd_last_error() gives a message for the calling thread's last failure.
 *  - Meshes are Legendre / Gauss-Lobatto in every dimension with 2..20
 *    points per dimension (Spectral.hpp:100-102).
 *  - A domain (sxd_domain) keeps one set of device work arrays per stream
 *    that calls it, so calls on different streams may overlap.  Calls that
 *    change its kernel choice must not overlap other calls on it.
 *  - The whole-domain RHS calls reject a dt that overlaps u
 *    (SXD_ERR_INVALID_ARGUMENT): neighbours' face values of u are read while
 *    dt is written.
 */
#ifndef SPECTRE_AMD_H
#define SPECTRE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SXD_OK = 0,
  SXD_ERR_INVALID_ARGUMENT = -1,
  SXD_ERR_UNSUPPORTED = -2,
  SXD_ERR_DEVICE = -3,
  SXD_ERR_OUT_OF_MEMORY = -4
};

/* Message describing the calling thread's most recent failure ("" if none). */
const char* sxd_last_error(void);

/* Library version string. */
const char* sxd_version(void);

/* ---- Spectral (host functions, no device needed) ------------------------ */

/* LGL collocation points and quadrature weights for n points.
 * Replaces Spectral::collocation_points / quadrature_weights for
 * Basis::Legendre, Quadrature::GaussLobatto
 * (NumericalAlgorithms/Spectral/Spectral.hpp:170-190, Legendre.cpp:187-232). */
int sxd_lgl_points_and_weights(int n, double* x, double* w);

/* LGL differentiation matrix, row-major d[i * n + j] = D_ij.
 * Replaces Spectral::differentiation_matrix(const Mesh<1>&)
 * (NumericalAlgorithms/Spectral/Spectral.cpp:113-150, 859-878). */
int sxd_lgl_differentiation_matrix(int n, double* d);

/* Legendre-Gauss collocation points and weights for n points (2..20).
 * Replaces Spectral::collocation_points / quadrature_weights for
 * Basis::Legendre, Quadrature::Gauss
 * (NumericalAlgorithms/Spectral/Legendre.cpp:106-148). */
int sxd_lg_points_and_weights(int n, double* x, double* w);

/* Differentiation matrix on the Legendre-Gauss points, row-major.
 * Replaces Spectral::differentiation_matrix for Quadrature::Gauss
 * (NumericalAlgorithms/Spectral/Spectral.cpp:113-150). */
int sxd_lg_differentiation_matrix(int n, double* d);

/* Boundary terms of the Gauss path, out [4][n]: the interpolation rows to
 * xi = -1 and xi = +1 (Spectral::boundary_interpolation_matrices,
 * Spectral.cpp:677-700) and the lifting terms l_j(-1) / w_j, l_j(+1) / w_j
 * (Spectral::boundary_lifting_term, Spectral.cpp:738-768). */
int sxd_lg_boundary_terms(int n, double* out);

/* The part of a parent mesh a child mesh (a mortar) covers
 * (Spectral::ChildSize, NumericalAlgorithms/Spectral/Projection.hpp:22-30). */
enum { SXD_CHILD_FULL = 1, SXD_CHILD_UPPER_HALF = 2, SXD_CHILD_LOWER_HALF = 3 };

/* LGL prolongation from n_parent to n_child >= n_parent points on the part
 * child_size of the parent interval, row-major [n_child][n_parent].
 * Replaces Spectral::projection_matrix_parent_to_child(const Mesh<1>&,
 * const Mesh<1>&, ChildSize) (Spectral/Projection.cpp:283-362). */
int sxd_projection_matrix_parent_to_child(int n_parent, int n_child, int child_size,
                                          double* out);

/* LGL restriction (non-massive operand) of data on n_child >= n_parent
 * points covering the part child_size of the parent interval, row-major
 * [n_parent][n_child].  Replaces Spectral::projection_matrix_child_to_parent
 * (Spectral/Projection.cpp:36-240; the mortar Vandermonde inverse is the
 * exact discrete one, w_j P_k(x_j) / gamma_k). */
int sxd_projection_matrix_child_to_parent(int n_child, int n_parent, int child_size,
                                          double* out);

/* ---- Linear operators ------------------------------------------------- */

/* Logical partial derivatives: dlogical = [3][E][nc][Np] (d/dxi, d/deta,
 * d/dzeta, each a batch of E Variables buffers).
 * Replaces logical_partial_derivatives
 * (NumericalAlgorithms/LinearOperators/PartialDerivatives.tpp:137-184). */
int sxd_logical_partial_derivatives(int num_elements, const int32_t extents[3],
                                    int nc, const double* u, double* dlogical,
                                    void* stream);

/* Partial derivatives with respect to inertial coordinates:
 * du = [E][3 nc][Np], block 3 c + i = d_i u_c.
 * Replaces partial_derivatives(not_null<Variables<deriv tags>*>,
 * const Variables<u tags>&, const Mesh<3>&, const InverseJacobian&)
 * (NumericalAlgorithms/LinearOperators/PartialDerivatives.tpp:191-241). */
int sxd_partial_derivatives(int num_elements, const int32_t extents[3], int nc,
                            const double* u, const double* inv_jac, double* du,
                            void* stream);

/* Transpose of sxd_partial_derivatives for a fixed inverse Jacobian (the
 * vector-Jacobian product of the linear map u -> du): dubar [E][3 nc][Np]
 * (block 3 c + i) -> ubar [E][nc][Np],
 *   ubar_c = sum_j D_j^T (sum_i d xi^j / d x^i dubar_{3c+i})
 * with D_j the LGL differentiation matrix applied along dimension j.  No
 * counterpart in the reference (which has no adjoints); it is the backward of
 * the torch.autograd path of the Python partial_derivative binding
 * (NumericalAlgorithms/LinearOperators/Python/PartialDerivatives.cpp:20-40).
 * SXD_ERR_UNSUPPORTED if 3 Np doubles exceed the LDS (Np above ~6500). */
int sxd_partial_derivatives_adjoint(int num_elements, const int32_t extents[3], int nc,
                                    const double* dubar, const double* inv_jac,
                                    double* ubar, void* stream);

/* ---- Flux systems: divergence and moving-mesh volume terms ------------- */

/* Fluxes use the layout of a Variables of flux tensors: per element 3 nc
 * blocks of Np, block i + 3 c = F^i of component c (the flux's spatial upper
 * index is its first, fastest index). */

/* Strong-form divergence div [E][nc][Np] of the fluxes,
 * div_c = sum_{i0} sum_d (d xi^d / d x^{i0}) d_d F^{i0}_c.
 * Replaces divergence(not_null<Variables<div tags>*>,
 * const Variables<flux tags>&, const Mesh<3>&, const InverseJacobian&)
 * (NumericalAlgorithms/LinearOperators/Divergence.tpp:50-94).
 * SXD_ERR_UNSUPPORTED if 3 Np doubles exceed the LDS (Np above ~6500). */
int sxd_divergence(int num_elements, const int32_t extents[3], int nc, const double* fluxes,
                   const double* inv_jac, double* div, void* stream);

/* Weak-form divergence: with G_j = sum_i (det J d xi^j / d x^i) F^i, the weak
 * flux differentiation matrix W_ab = D_ba w_b / w_a applied along each
 * logical direction j and summed.  det_jac_times_inv_jac [E][9][Np] in the
 * inverse-Jacobian layout.  Replaces weak_divergence
 * (NumericalAlgorithms/LinearOperators/WeakDivergence.hpp:40-231, 3-d
 * branch; matrix Spectral.cpp:462-485). */
int sxd_weak_divergence(int num_elements, const int32_t extents[3], int nc,
                        const double* fluxes, const double* det_jac_times_inv_jac, double* div,
                        void* stream);

/* The flux-variable part of the DG volume terms
 * (Evolution/DiscontinuousGalerkin/Actions/VolumeTermsImpl.tpp:156-200,
 * 264-291) for a system whose nc evolved components all have fluxes, on dt
 * [E][nc][Np] (which holds the system's source terms on entry):
 *  - moving mesh (mesh_velocity [E][3][Np] non-null): F^i_c -= u_c v^i
 *    (fluxes updated in place, as the reference updates its flux buffer) and
 *    dt_c -= u_c div v (div_mesh_velocity [E][Np]);
 *  - then dt -= div F (SXD_DG_STRONG_INERTIAL, inv_jac) or
 *    dt += det_inv_jac weak_div F (SXD_DG_WEAK_INERTIAL, det_jac_times_inv_jac
 *    and det_inv_jac [E][Np]).
 * Pointers a formulation does not use may be null. */
enum { SXD_DG_STRONG_INERTIAL = 0, SXD_DG_WEAK_INERTIAL = 1 };
int sxd_flux_volume_terms(int num_elements, const int32_t extents[3], int nc, const double* u,
                          double* fluxes, const double* mesh_velocity,
                          const double* div_mesh_velocity, int formulation,
                          const double* inv_jac, const double* det_jac_times_inv_jac,
                          const double* det_inv_jac, double* dt, void* stream);

/* Moving-mesh term of evolved variables without fluxes (ScalarWave, GH):
 * dt_c += v^i d_i u_c on num_points points; du [3 nc][n] in the
 * partial_derivatives layout, mesh_velocity [3][n], dt [nc][n]
 * (VolumeTermsImpl.tpp:235-262). */
int sxd_mesh_velocity_nonconservative(int64_t num_points, int nc, const double* du,
                                      const double* mesh_velocity, double* dt, void* stream);

/* ---- Point-wise system terms ------------------------------------------ */

/* ScalarWave::TimeDerivative<3>::apply over num_points points
 * (Evolution/Systems/ScalarWave/TimeDerivative.cpp:14-45).
 * u [5][n] (psi, Pi, Phi_i), du [15][n] (d psi, d Pi, d_i Phi_j at 6+i+3j),
 * gamma2 [n]; dt [5][n]; gamma2_temp [n] (the temporary copy, may be NULL). */
int sxd_scalar_wave_time_derivative(int64_t num_points, const double* u,
                                    const double* du, const double* gamma2,
                                    double* dt, double* gamma2_temp,
                                    void* stream);

/* gh::TimeDerivative<3>::apply, harmonic gauge, static mesh
 * (Evolution/Systems/GeneralizedHarmonic/TimeDerivative.cpp:30-407).
 * u [50][n] (g_ab, Pi_ab, Phi_iab), du [150][n] (partial_derivatives layout),
 * gammas [3][n] (gamma0, gamma1, gamma2); dt [50][n];
 * temps [12][n] (gamma1, gamma2, lapse, shift^i, inverse spatial metric), may
 * be NULL. */
int sxd_gh_time_derivative(int64_t num_points, const double* u,
                           const double* du, const double* gammas, double* dt,
                           double* temps, void* stream);

/* ---- Boundary correction and lift -------------------------------------- */

/* ScalarWave UpwindPenalty<3>: dg_package_data on both sides of n face
 * points followed by dg_boundary_terms
 * (Evolution/Systems/ScalarWave/BoundaryCorrections/UpwindPenalty.cpp:36-205).
 * u_int/u_ext [5][n], gamma2_int/ext [n], normal_int/ext [3][n] unit normal
 * covectors, each side pointing out of its own element; bc [5][n]. */
int sxd_scalar_wave_upwind_penalty(int64_t num_points, const double* u_int,
                                   const double* gamma2_int,
                                   const double* normal_int,
                                   const double* u_ext,
                                   const double* gamma2_ext,
                                   const double* normal_ext, double* bc,
                                   void* stream);

/* GH UpwindPenalty<3> on n face points
 * (Evolution/Systems/GeneralizedHarmonic/BoundaryCorrections/UpwindPenalty.cpp:36-275),
 * including the metric normalisation of the normal
 * (Evolution/DiscontinuousGalerkin/Actions/NormalCovectorAndMagnitude.hpp:47-81).
 * u_* [50][n]; temps_* [12][n] as written by sxd_gh_time_derivative;
 * normal_* [3][n] UNNORMALIZED covectors; bc [50][n]. */
int sxd_gh_upwind_penalty(int64_t num_points, const double* u_int,
                          const double* temps_int, const double* normal_int,
                          const double* u_ext, const double* temps_ext,
                          const double* normal_ext, double* bc, void* stream);

/* dg::lift_flux for Gauss-Lobatto points: bc *= -0.5 N (N - 1) |n|, in place,
 * nc components of n face points
 * (NumericalAlgorithms/DiscontinuousGalerkin/LiftFlux.hpp:41-62). */
int sxd_lift_flux(int64_t num_points, int nc, double* bc,
                  int extent_perpendicular, const double* magnitude_of_normal,
                  void* stream);

/* ---- Whole DG right-hand side on a periodic brick ----------------------- */

/* A set of conforming, aligned hexahedral elements with equal extents and a
 * neighbour table.  Replaces the per-element Mesh<3> + neighbour/mortar
 * bookkeeping of ComputeTimeDerivative / ApplyBoundaryCorrections
 * (Evolution/DiscontinuousGalerkin/Actions/ComputeTimeDerivative.hpp:383-650,
 * ApplyBoundaryCorrections.hpp:722-1050) for a single device.
 * neighbors: host array [6 E], entry 6 e + 2 d + side (side 0 = lower,
 * 1 = upper) is the element across that face, or -1 for an external face
 * (closed by a boundary condition, see sxd_scalar_wave_rhs_dirichlet,
 * sxd_scalar_wave_rhs_spherical_radiation,
 * sxd_scalar_wave_rhs_constraint_preserving_spherical_radiation and
 * sxd_gh_rhs_dirichlet; every other call on such a domain returns
 * SXD_ERR_UNSUPPORTED).  The table is copied to
 * the device; this call allocates device memory. */
typedef struct sxd_domain sxd_domain;
int sxd_domain_create(sxd_domain** domain, int num_elements,
                      const int32_t extents[3], const int32_t* neighbors);
int sxd_domain_destroy(sxd_domain* domain);

/* A mixed-p set of conforming, aligned hexahedral elements: element e has its
 * own LGL extents extents[3 e .. 3 e + 2] (2..20 each; p-refinement).  Faces
 * between elements of different extents are coupled through mortars with the
 * larger extents per dimension (dg::mortar_mesh,
 * NumericalAlgorithms/DiscontinuousGalerkin/MortarHelpers.cpp:22-45), packaged
 * data interpolated to the mortar and corrections projected back
 * (Spectral/Projection.cpp, ChildSize::Full), as in
 * ApplyBoundaryCorrections.hpp:797-907.  Fields passed to the RHS calls of
 * such a domain use the ragged Variables layout: element e's nc blocks of
 * Np_e points start at nc * offset_e, offset_e = sum_{e' < e} Np_{e'}
 * (damping parameters and inverse Jacobian likewise, with their component
 * counts).  SXD_ERR_UNSUPPORTED if an element or face is too large for the
 * kernels' LDS plan (element Np up to 4096 and faces up to 12 x 12 fit).
 * A neighbour of -1 marks an external face, evaluated by
 * sxd_scalar_wave_rhs_dirichlet, sxd_scalar_wave_rhs_spherical_radiation,
 * sxd_scalar_wave_rhs_constraint_preserving_spherical_radiation and
 * sxd_gh_rhs_dirichlet only. */
int sxd_domain_create_mixed(sxd_domain** domain, int num_elements,
                            const int32_t* extents, const int32_t* neighbors);
/* A set of elements coupled through explicitly listed mortars: neighbours
 * whose logical axes are oriented differently (OrientationMap,
 * Domain/Structure/OrientationMap.hpp) and h-refined faces covered by two or
 * four mortars (2:1 balance), on top of per-element extents as in
 * sxd_domain_create_mixed (ragged layout).  Replaces the mortar bookkeeping
 * of ComputeTimeDerivative.hpp:690-774 (orient_variables_on_slice,
 * Domain/Structure/OrientationMapHelpers.cpp:372-400) and
 * ApplyBoundaryCorrections.hpp:797-1043 with
 * projection_matrix_parent_to_child / _child_to_parent for the Full,
 * UpperHalf and LowerHalf mortar sizes (Spectral/Projection.cpp).
 * records: host array [num_records][8] of int32, one per (element,
 * direction, neighbour) -- the reference's DirectionalId keyed mortars:
 *   { e, dir, en, o0, o1, o2, size_a, size_b }
 * dir = 2 d + side of e's logical frame; o_j = the neighbour's logical
 * direction 2 k + side that e's upper logical axis j maps to (the mapped
 * directions of the OrientationMap; aligned = {1, 3, 5}); size_a, size_b =
 * the part of e's face the mortar covers along the face's two dimensions
 * (lower dimension first): SXD_CHILD_FULL, SXD_CHILD_UPPER_HALF,
 * SXD_CHILD_LOWER_HALF (Spectral::ChildSize; dg::mortar_size,
 * NumericalAlgorithms/DiscontinuousGalerkin/MortarHelpers.cpp:49-74).  Every
 * record needs the neighbour's record back (direction o(dir opposite)), with
 * the inverse orientation; the mortar mesh is the larger extents of the two
 * faces per dimension; the mortars of a face must cover it exactly once; no
 * external faces.  Corrections reach dt face by face in direction order and,
 * within a face, in the order of the records.  sxd_scalar_wave_rhs,
 * sxd_gh_rhs, sxd_domain_autotune and the RK3 stepper take such a domain
 * (kernel SXD_KERNEL_MORTAR); the boundary-condition, Gauss, adjoint and
 * allocation-free calls return SXD_ERR_UNSUPPORTED.  SXD_ERR_INVALID_ARGUMENT
 * (with the reason in sxd_last_error) for an inconsistent record list. */
int sxd_domain_create_mortars(sxd_domain** domain, int num_elements, const int32_t* extents,
                              int num_records, const int32_t* records);
/* One rank's part of a mortar domain, for a run split over devices or nodes:
 * the elements e with partition[e] == rank of the global element set and
 * record list above (the same arrays on every rank), in global order, with
 * the ragged layout of those elements only.  A mortar whose two sides live
 * on different ranks is coupled by boundary messages -- the
 * BoundaryData / BoundaryMessage exchange of ComputeTimeDerivative.hpp:
 * 690-774 (Evolution/DiscontinuousGalerkin/BoundaryData.hpp,
 * Messages/BoundaryMessage.hpp) -- so its RHS runs as sxd_rhs_begin, the
 * exchange, sxd_boundary_messages_unpack per sending rank, sxd_rhs_end;
 * sxd_scalar_wave_rhs / sxd_gh_rhs return SXD_ERR_UNSUPPORTED while such
 * mortars exist. */
int sxd_domain_create_mortars_partition(sxd_domain** domain, int num_elements,
                                        const int32_t* extents, int num_records,
                                        const int32_t* records, const int32_t* partition,
                                        int rank);
/* The send buffer of a partitioned mortar domain for `system`: one group per
 * receiving rank (ascending), each a directory (int64 message count, int64
 * byte offset of every message within the group; padded to 64 bytes) and
 * its messages.  A message is a 64-byte header of 16 int32 -- magic
 * 0x53584442, system, sender element, receiver element, receiver direction,
 * sender direction, sender extents (3), mortar extents in the receiver's
 * frame (2), packaged components (13 ScalarWave, 134 GH), data doubles,
 * reserved, time step id (low, high) -- and the sender's packaged mortar
 * data (dg_package_data projected to the mortar) in the receiver's point
 * order (orient_variables_on_slice), padded to 64 bytes.  Fills
 * *num_peers, the total size, and (when non-null, num_peers entries each)
 * the ranks, byte offsets and sizes of the groups: the ranges an RCCL
 * ncclSend to each rank would carry. */
int sxd_boundary_messages_layout(const sxd_domain* domain, int system, int* num_peers,
                                 int32_t* ranks, int64_t* offsets, int64_t* bytes,
                                 int64_t* total_bytes);
/* Number of messages this rank receives from `rank` per RHS. */
int sxd_boundary_messages_expected(const sxd_domain* domain, int rank, int* count);
/* First half of a partitioned RHS: volume terms into dt, the packages of
 * every mortar, and the outgoing messages into send (device, the layout's
 * total size; stamped with time_step_id). */
int sxd_rhs_begin(const sxd_domain* domain, int system, const double* u, const double* params,
                  const double* inv_jac, double* dt, void* send, int64_t time_step_id,
                  void* stream);
/* Take the group of messages received from sender_rank (device, bytes
 * long).  Every header is checked (magic, system, time step id, that the
 * mortar exists on this rank with the same mesh, no mortar twice); a failed
 * check returns SXD_ERR_INVALID_ARGUMENT naming the message and the reason.
 * Synchronizes the stream. */
int sxd_boundary_messages_unpack(const sxd_domain* domain, int system, int sender_rank,
                                 const void* recv, int64_t bytes, void* stream);
/* Second half: the corrections of every mortar into dt, once every remote
 * mortar's message has arrived (SXD_ERR_INVALID_ARGUMENT otherwise).  Same
 * stream as sxd_rhs_begin. */
int sxd_rhs_end(const sxd_domain* domain, int system, double* dt, void* stream);

/* Total grid points of the domain (sum of Np_e). */
int sxd_domain_num_points(const sxd_domain* domain, int64_t* points);

/* Full ScalarWave DG time derivative (volume terms + UpwindPenalty + LGL lift)
 * for every element of the domain: ComputeTimeDerivative followed by
 * ApplyBoundaryCorrectionsToTimeDerivative.
 * u, dt [E][5][Np]; gamma2 [E][Np]; inv_jac [E][9][Np]. */
int sxd_scalar_wave_rhs(const sxd_domain* domain, const double* u,
                        const double* gamma2, const double* inv_jac,
                        double* dt, void* stream);

/* sxd_scalar_wave_rhs on a domain with external faces (neighbour -1), closed
 * by the Dirichlet analytic ghost boundary condition: on an external face the
 * exterior UpwindPenalty state is ghost's fields at the face's own points,
 * with the interior gamma2 and the reversed interior unit normal, and the
 * correction is lifted with the interior normal magnitude
 * (ScalarWave/BoundaryConditions/DirichletAnalytic.cpp:72-100 dg_ghost, applied
 * by Evolution/DiscontinuousGalerkin/Actions/BoundaryConditionsImpl.hpp:427-586).
 * ghost [E][5][Np] (psi, pi, phi_i of the analytic solution at every point;
 * only face points are read).  Internal faces are as in sxd_scalar_wave_rhs.
 * Uniform and mixed-p domains (ghost then in the ragged layout). */
int sxd_scalar_wave_rhs_dirichlet(const sxd_domain* domain, const double* u,
                                  const double* gamma2, const double* inv_jac,
                                  const double* ghost, double* dt, void* stream);

/* sxd_scalar_wave_rhs on a domain with external faces closed by the
 * SphericalRadiation ghost condition
 * (ScalarWave/BoundaryConditions/SphericalRadiation.cpp:62-100 dg_ghost): the
 * exterior state keeps the interior psi and phi_i and sets
 * pi = n^i phi_i (outward interior unit normal), plus psi / |x| for the
 * Bayliss-Turkel type; interior gamma2 and reversed unit normal as in
 * sxd_scalar_wave_rhs_dirichlet.  coords [E][3][Np]: inertial coordinates
 * (Bayliss-Turkel only; may be null for Sommerfeld).  Uniform and mixed-p
 * domains (coords then in the ragged layout). */
enum {
  SXD_SPHERICAL_RADIATION_SOMMERFELD = 1,
  SXD_SPHERICAL_RADIATION_BAYLISS_TURKEL = 2
};
int sxd_scalar_wave_rhs_spherical_radiation(const sxd_domain* domain,
                                            const double* u,
                                            const double* gamma2,
                                            const double* inv_jac,
                                            const double* coords, int type,
                                            double* dt, void* stream);

/* sxd_scalar_wave_rhs on a domain with external faces closed by the
 * ConstraintPreservingSphericalRadiation time-derivative condition
 * (ScalarWave/BoundaryConditions/ConstraintPreservingSphericalRadiation.cpp:
 * 59-132 dg_time_derivative, static mesh; added to dt on the face as in
 * Actions/BoundaryConditionsImpl.hpp:636-648): external faces get no upwind
 * term; at their points dt Pi gains the correction built from the interior
 * state, its derivatives, the outward unit normal and gamma2 (Sommerfeld,
 * first- or second-order Bayliss-Turkel, the latter two with 1/|x|).
 * coords [E][3][Np] (Bayliss-Turkel types; may be null for Sommerfeld).
 * Uniform and mixed-p domains (coords then in the ragged layout). */
enum {
  SXD_CPSR_SOMMERFELD = 1,
  SXD_CPSR_FIRST_ORDER_BAYLISS_TURKEL = 2,
  SXD_CPSR_SECOND_ORDER_BAYLISS_TURKEL = 3
};
int sxd_scalar_wave_rhs_constraint_preserving_spherical_radiation(
    const sxd_domain* domain, const double* u, const double* gamma2,
    const double* inv_jac, const double* coords, int type, double* dt,
    void* stream);

/* Full ScalarWave DG time derivative on Legendre-Gauss points (the
 * Quadrature::Gauss branches of ComputeTimeDerivative and
 * ApplyBoundaryCorrections: face values interpolated along the normal pencil,
 * ProjectToBoundary.hpp:54-147; boundary corrections lifted onto whole
 * pencils, ApplyBoundaryCorrections.hpp:939-986 and
 * NumericalAlgorithms/DiscontinuousGalerkin/LiftFromBoundary.cpp:140-180).
 * Uniform periodic domains.  u, dt [E][5][Np]; gamma2, det_inv_jac [E][Np]
 * (det of the inverse Jacobian at the volume points); inv_jac [E][9][Np];
 * face_inv_jac: per element the six faces in direction order (lower xi,
 * upper xi, lower eta, ...), each 3 blocks of its Nf points (face-mesh order,
 * remaining dimensions lowest first) holding d xi^d / d x^i of the face's
 * dimension d, i = 0..2, as the map gives them on the face. */
int sxd_scalar_wave_rhs_gauss(const sxd_domain* domain, const double* u, const double* gamma2,
                              const double* inv_jac, const double* det_inv_jac,
                              const double* face_inv_jac, double* dt, void* stream);

/* Full generalized-harmonic DG time derivative (harmonic gauge).
 * u, dt [E][50][Np]; gammas [E][3][Np]; inv_jac [E][9][Np]. */
int sxd_gh_rhs(const sxd_domain* domain, const double* u, const double* gammas,
               const double* inv_jac, double* dt, void* stream);

/* sxd_gh_rhs without any device allocation: the composite and mixed-p kernel
 * organisations use a fixed work array of 16 Mi doubles (128 MiB) that the
 * library holds on each device it runs on until the process ends, instead of
 * the domain's per-stream work arrays.  No counterpart in the reference; it
 * is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The domain (and so its
 * device tables) must already exist.  SXD_ERR_UNSUPPORTED if the domain's
 * work arrays exceed the fixed array: 150 doubles per grid point on mixed-p
 * domains (about 111,000 points), 162 per point with the composite kernel,
 * none with the fused kernels (mixed-p domains run their batched mixed-p
 * kernels here, also where SXD_KERNEL_MORTAR is selected; mortar domains
 * return SXD_ERR_UNSUPPORTED).  Every call on one device shares that device's
 * array, which keeps the last call's contents: calls that overlap on one GPU
 * (other streams, other domains, other threads) race and must be serialised
 * by the caller.  Results equal sxd_gh_rhs bit for bit. */
int sxd_gh_rhs_noalloc(const sxd_domain* domain, const double* u, const double* gammas,
                       const double* inv_jac, double* dt, void* stream);

/* sxd_gh_rhs on a domain with external faces (neighbour -1), closed by the
 * generalized-harmonic Dirichlet analytic ghost condition: on an external face
 * the exterior UpwindPenalty state is ghost's g_ab, Pi_ab, Phi_iab at the
 * face's own points, with the interior gamma1 and gamma2, the lapse, shift and
 * inverse spatial metric of the ghost metric, and the reversed interior unit
 * normal renormalised with the ghost inverse spatial metric; the correction is
 * lifted with the interior normal magnitude
 * (GeneralizedHarmonic/BoundaryConditions/DirichletAnalytic.cpp:70-115
 * dg_ghost, applied by
 * Evolution/DiscontinuousGalerkin/Actions/BoundaryConditionsImpl.hpp:427-586).
 * ghost [E][50][Np] (only face points are read; ragged on mixed-p domains).
 * Runs the domain's fused (vector-ALU), composite or mixed-p kernels;
 * SXD_ERR_UNSUPPORTED if the fused MFMA kernel is selected. */
int sxd_gh_rhs_dirichlet(const sxd_domain* domain, const double* u,
                         const double* gammas, const double* inv_jac,
                         const double* ghost, double* dt, void* stream);

/* ---- Kernel selection (no counterpart in the reference) ----------------- */

/* The reference has one CPU implementation per call; here a whole-domain
 * right-hand side can run as one of several kernel organisations with
 * identical results (every choice is bit-for-bit the same): */
enum {
  SXD_KERNEL_AUTO = 0,      /* the library's default for the domain's extents */
  SXD_KERNEL_COMPOSITE = 1, /* derivative, point-wise, face and lift launches */
  SXD_KERNEL_FUSED = 2,     /* one launch per RHS, derivatives on the vector ALUs */
  SXD_KERNEL_FUSED_MFMA = 3, /* one launch per RHS, derivatives on FP64 MFMA (GH) */
  /* mixed-p domains: */
  SXD_KERNEL_MIXED_BATCHED = 4,    /* one launch per kernel over all elements,
                                      ordered by decreasing cost */
  SXD_KERNEL_MIXED_PER_EXTENTS = 5, /* one launch per kernel and group of equal extents */
  /* mortar domains, and periodic mixed-p domains (the default there): */
  SXD_KERNEL_MORTAR = 6 /* ragged volume kernels, then package, projection,
                           boundary-term and lift launches over the mortars (the
                           packages and boundary terms on an internal second
                           stream, joined to the caller's stream by an event) */
};
enum { SXD_SYSTEM_SCALAR_WAVE = 0, SXD_SYSTEM_GH = 1 };

/* Types of the GH constraint-preserving Bjorhus boundary condition
 * (detail::ConstraintPreservingBjorhusType, Evolution/Systems/
 * GeneralizedHarmonic/BoundaryConditions/Bjorhus.hpp): ConstraintPreserving,
 * and ConstraintPreservingPhysical, which also sets the incoming physical
 * (Weyl-tensor) modes of v_minus (BjorhusImpl.cpp:160-496). */
enum {
  SXD_BJORHUS_CONSTRAINT_PRESERVING = 0,
  SXD_BJORHUS_CONSTRAINT_PRESERVING_PHYSICAL = 1
};

/* GH constraint-preserving Bjorhus boundary condition of type bjorhus_type on
 * a static mesh: gh::BoundaryConditions::ConstraintPreservingBjorhus::
 * dg_time_derivative (Evolution/Systems/GeneralizedHarmonic/
 * BoundaryConditions/Bjorhus.cpp:67-383, BjorhusImpl.cpp:9-532).
 * Point-wise on num_points points, every argument a device array in the
 * reference's tensor storage (component blocks of num_points): the outward
 * unit normal covector [3], g_ab [10], Pi_ab [10], Phi_iab [30], inertial
 * coordinates [3], gamma1, gamma2, lapse, shift [3], inverse spacetime metric
 * [10], spacetime unit normal vector [4], three-index constraint [30], gauge
 * source H_a [4], its spacetime derivative d_a H_b [16, a + 4 b], the volume
 * time derivatives of g [10], Pi [10], Phi [30], and d_i Pi_ab [30],
 * d_j Phi_iab [90, j + 3 i + 9 ab]; outputs the corrections to dt g [10],
 * dt Pi [10], dt Phi [30].  Each point is treated as a face of its own (the
 * corrections vanish where no speed is negative).  SXD_ERR_INVALID_ARGUMENT
 * for an unknown type. */
int sxd_gh_bjorhus_corrections(
    int bjorhus_type, int64_t num_points, const double* normal_covector,
    const double* spacetime_metric, const double* pi, const double* phi, const double* coords, const double* gamma1,
    const double* gamma2, const double* lapse, const double* shift,
    const double* inverse_spacetime_metric, const double* spacetime_unit_normal_vector,
    const double* three_index_constraint, const double* gauge_source,
    const double* spacetime_deriv_gauge_source, const double* dt_spacetime_metric,
    const double* dt_pi, const double* dt_phi, const double* d_pi, const double* d_phi,
    double* dt_spacetime_metric_correction, double* dt_pi_correction, double* dt_phi_correction,
    void* stream);

/* Full GH DG time derivative of a uniform domain whose external faces
 * (neighbour -1) carry the ConstraintPreservingBjorhus condition of type
 * bjorhus_type (harmonic gauge: H_a = 0): volume terms, then on each external face, one dimension
 * after another, the Bjorhus correction of the face's points (from the
 * volume time derivatives and derivatives at the point; no correction on a
 * face without a negative characteristic speed), then the UpwindPenalty
 * corrections of the internal faces (Actions/BoundaryConditionsImpl.hpp:
 * 586-660, a time-derivative condition has no upwind term).  coords:
 * inertial coordinates (the gauge condition's 1/r), [E][3][Np] on uniform
 * domains, ragged [3][Np_e] on mixed-p ones.  Uniform domains run the
 * composite kernels, mixed-p domains the mixed-p volume and face-pass
 * kernels whatever kernel is set; SXD_ERR_UNSUPPORTED on mortar domains.  Results
 * agree with the oracle to 1e-12 (the corrections are tensor sums whose order
 * is not the reference's). */
int sxd_gh_rhs_bjorhus(const sxd_domain* domain, int bjorhus_type, const double* u,
                       const double* gammas, const double* inv_jac, const double* coords,
                       double* dt, void* stream);

/* Pin the kernel used by sxd_scalar_wave_rhs / sxd_gh_rhs on this domain.
 * SXD_ERR_UNSUPPORTED if the kernel cannot run these extents.  Not
 * thread-safe against concurrent RHS calls on the same domain. */
int sxd_domain_set_kernel(sxd_domain* domain, int system, int kernel);
/* The kernel the next RHS call on this domain will use. */
int sxd_domain_get_kernel(const sxd_domain* domain, int system, int* kernel);
/* Threads per workgroup (tile shape): mixed-p kernels 64, 128 or 256
 * (default 128); the fused GH kernel at 7^3 and 8^3 points 256 (one wave per
 * SIMD, two points per thread) or 512 (two waves per SIMD, one point per
 * thread).  Results do not depend on it. */
int sxd_domain_set_threads(sxd_domain* domain, int system, int threads);
int sxd_domain_get_threads(const sxd_domain* domain, int system, int* threads);
/* Autotuner: time every kernel available for (domain, system) on the given
 * inputs (one warm-up, then `repetitions` timed calls each, HIP events on
 * `stream`), keep the fastest for later calls and return it in *chosen and
 * its mean time in *milliseconds (either may be NULL).  params is gamma2
 * (ScalarWave) or gammas (GH); dt is overwritten.  Synchronises `stream`. */
int sxd_domain_autotune(sxd_domain* domain, int system, const double* u,
                        const double* params, const double* inv_jac, double* dt,
                        int repetitions, void* stream, int* chosen,
                        float* milliseconds);

/* ---- Time stepping ------------------------------------------------------ */

/* One substep of TimeSteppers::Rk3HesthavenSsp::update_u
 * (Time/TimeSteppers/Rk3HesthavenSsp.cpp:55-81) on n doubles:
 *   substep 0: out = u0 + dt lk
 *   substep 1: out = 0.25 (3 u0 + uk + dt lk)
 *   substep 2: out = (1/3) (u0 + 2 uk + 2 dt lk)               */
int sxd_rk3_substep(int substep, int64_t n, double dt, const double* u0,
                    const double* uk, const double* lk, double* out,
                    void* stream);

/* ---- Adjoints (no counterpart in the reference) ------------------------- */

/* Transpose of the ScalarWave DG RHS of sxd_scalar_wave_rhs, which is linear
 * in u for fixed gamma2 and inverse Jacobian: ubar = A^T gbar, i.e. the
 * gradient of a functional J with respect to u given gbar = dJ/d(dt).
 * Uniform domains with isotropic extents 2..10; gbar, ubar [E][5][Np] (must
 * not alias). */
int sxd_scalar_wave_rhs_adjoint(const sxd_domain* domain, const double* gbar,
                                const double* gamma2, const double* inv_jac,
                                double* ubar, void* stream);
/* Adjoint of the GH DG RHS of sxd_gh_rhs linearised at the state u:
 * ubar = (d dt / d u)^T gbar, the gradient with respect to u of a functional
 * J given gbar = dJ/d(dt).  Uniform domains (3 Np + N0^2 + N1^2 + N2^2 <=
 * 20480, e.g. 18^3): u, gbar,
 * ubar [E][50][Np]; mixed-p domains (sxd_domain_create_mixed): the ragged
 * layout.  ubar must not alias u or gbar.  Uses the domain's work arrays,
 * allocated on first use: 150 doubles per grid point, and on uniform domains
 * 300 doubles per element and point of its largest face.  No counterpart in
 * the reference (which has no adjoints). */
int sxd_gh_rhs_adjoint(const sxd_domain* domain, const double* u,
                       const double* gammas, const double* inv_jac,
                       const double* gbar, double* ubar, void* stream);
/* Transpose of one Rk3HesthavenSsp substep (sxd_rk3_substep), applied in
 * reverse substep order 2, 1, 0 on n doubles.  lam is the adjoint of the
 * substep's output, a = A^T lam (from the adjoint RHS), mu accumulates the
 * adjoint of the step's input state:
 *   substep 2: lam_out = 2/3 (lam + dt a),  mu  = lam / 3
 *   substep 1: lam_out = (lam + dt a) / 4,  mu += 3/4 lam
 *   substep 0:                              mu += lam + dt a   (lam_out unused) */
int sxd_rk3_adjoint_substep(int substep, int64_t n, double dt, const double* lam,
                            const double* a, double* mu, double* lam_out,
                            void* stream);

/* A full Rk3HesthavenSsp step (three right-hand sides and three update_u
 * substeps, Time/TimeSteppers/Rk3HesthavenSsp.cpp:55-81) of `system` on
 * `domain`, captured once into a HIP graph and replayed per step; replaces the
 * per-substep action loop (ComputeTimeDerivative, RecordTimeStepperData,
 * UpdateU: Evolution/Actions/..., Time/Actions/UpdateU.hpp:43-101) for a
 * fixed-step evolution on one device.  u (the state, updated in place),
 * params (gamma2 or gammas) and inv_jac are device arrays in the domain's
 * layout and must stay allocated while the stepper lives; the stepper owns
 * two field-sized work arrays.  The kernel organisation the domain has at
 * creation time (sxd_domain_set_kernel / sxd_domain_autotune) is the one
 * captured.  The stepper owns the work arrays its graph uses (sized at
 * creation), so later calls on the domain cannot move them; the domain must
 * outlive the stepper. */
typedef struct sxd_rk3_stepper sxd_rk3_stepper;
int sxd_rk3_stepper_create(sxd_rk3_stepper** stepper, sxd_domain* domain,
                           int system, double* u, const double* params,
                           const double* inv_jac);
/* Enqueue num_steps steps of size dt on `stream` (after waiting for work
 * already on it, to set the step size). */
int sxd_rk3_stepper_step(sxd_rk3_stepper* stepper, double dt, int num_steps,
                         void* stream);
int sxd_rk3_stepper_destroy(sxd_rk3_stepper* stepper);

#ifdef __cplusplus
}
#endif

#endif /* SPECTRE_AMD_H */
