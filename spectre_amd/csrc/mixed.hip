// Mixed-p (p-refinement) DG right-hand sides: elements with different LGL
// extents per dimension, coupled through mortars.
//
// Layout: the ragged Variables layout.  Element e has Np_e = N0 N1 N2 points
// and starts at point offset off_e; a field with nc components keeps element
// e's nc blocks of Np_e doubles at [nc off_e, nc (off_e + Np_e)).
//
// Launch organisation: every kernel takes a work list of element indices and
// runs one workgroup per entry.  The batched form is one launch per kernel
// over the whole domain with the list sorted by decreasing cost, so the
// dispatcher hands out the expensive elements first; the per-extents form
// launches the same kernels once per group of equal extents.
//
// Face coupling (ApplyBoundaryCorrections.hpp:797-907 with
// InternalMortarDataImpl.hpp:230-318), each element's faces added into dt in
// the order lower/upper xi, eta, zeta (ScalarWave: one workgroup per element
// runs its six faces; GH: one launch per dimension, one workgroup per face
// and symmetric pair):
//   1. packages both sides on their own face meshes (UpwindPenalty
//      dg_package_data; the neighbour's side from the neighbour's fields),
//   2. interpolates both packages to the mortar mesh (max extents per
//      dimension, dg::mortar_mesh) where a face mesh differs from it,
//      dimension a then dimension b (apply_matrices order),
//   3. evaluates dg_boundary_terms at the mortar points,
//   4. projects the correction back to its own face mesh (modal truncation)
//      where needed, and
//   5. lifts it into dt (dg::lift_flux, -N(N-1)/2 |n| on the LGL face).
// Matrix products are sums over the input index l = 0 .. n-1, first term a
// product and then fused multiply-adds, as in the oracle.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <vector>

#include "device_common.hpp"
#include "gh_adjoint.hpp"
#include "gh_point.hpp"
#include "kernels.hpp"
#include "sw_point.hpp"

namespace sxd {
namespace {

__device__ __forceinline__ int point_index(const MixedElement& m, int i, int j, int k) {
  return i + m.n[0] * (j + m.n[1] * k);
}

// Volume index of point (a, b) of the face in dimension d on `side`.
// ext_at(m, d) for a runtime d, by selects: a runtime index into the element's
// extents would place the element record in scratch memory.
__device__ __forceinline__ int ext_at(const MixedElement& m, int d) {
  return d == 0 ? m.n[0] : (d == 1 ? m.n[1] : m.n[2]);
}

__device__ __forceinline__ int face_point(const MixedElement& m, int d, int side, int a, int b) {
  const int fixed = side == 0 ? 0 : ext_at(m, d) - 1;
  if (d == 0) return point_index(m, fixed, a, b);
  if (d == 1) return point_index(m, a, fixed, b);
  return point_index(m, a, b, fixed);
}

// Logical derivatives of LDS component `sc` at point (i, j, k): sum over
// l = 0 .. n_d - 1 in order, first term a product.
__device__ __forceinline__ void logical_derivs(const double* sc, const MixedElement& m,
                                               const double* D0, const double* D1,
                                               const double* D2, int i, int j, int k,
                                               double* d) {
  const int n0 = m.n[0], n1 = m.n[1], n2 = m.n[2];
  double d0 = D0[i * n0] * sc[n0 * (j + n1 * k)];
  for (int l = 1; l < n0; ++l) d0 = fma(D0[i * n0 + l], sc[l + n0 * (j + n1 * k)], d0);
  double d1 = D1[j * n1] * sc[i + n0 * n1 * k];
  for (int l = 1; l < n1; ++l) d1 = fma(D1[j * n1 + l], sc[i + n0 * (l + n1 * k)], d1);
  double d2 = D2[k * n2] * sc[i + n0 * j];
  for (int l = 1; l < n2; ++l) d2 = fma(D2[k * n2 + l], sc[i + n0 * (j + n1 * l)], d2);
  d[0] = d0;
  d[1] = d1;
  d[2] = d2;
}

// ---------------------------------------------------------------------------
// ScalarWave volume terms.  Dynamic LDS: 5 Np_max doubles.  CPSR: external
// faces (neighbour -1) are closed by the constraint-preserving spherical
// radiation condition of type bc_kind (kSwBcCpsr*; coords = ragged [3][Np_e]
// for the Bayliss-Turkel types), a time-derivative condition: a point on such
// faces adds their dt Pi corrections, in direction order, to its volume terms
// (the face kernel then skips those faces).
template <bool CPSR>
__global__ void __launch_bounds__(256)
sw_mixed_volume_kernel(const MixedElement* __restrict__ el, const int32_t* __restrict__ list,
                       const double* __restrict__ u, const double* __restrict__ gamma2,
                       const double* __restrict__ inv_jac, double* __restrict__ dt,
                       const double* __restrict__ dtab, const int32_t* __restrict__ nbr,
                       int bc_kind, const double* __restrict__ coords) {
  extern __shared__ double s_u[];
  const int e = list[blockIdx.x];
  const MixedElement m = el[e];
  const int np = m.np;
  const double* ue = u + 5 * m.off;
  for (int p = threadIdx.x; p < 5 * np; p += blockDim.x) s_u[p] = ue[p];
  __syncthreads();
  const double* D0 = diff_matrix(dtab, m.n[0]);
  const double* D1 = diff_matrix(dtab, m.n[1]);
  const double* D2 = diff_matrix(dtab, m.n[2]);
  const double* ij = inv_jac + 9 * m.off;
  for (int p = threadIdx.x; p < np; p += blockDim.x) {
    const int i = p % m.n[0], j = (p / m.n[0]) % m.n[1], k = p / (m.n[0] * m.n[1]);
    double up[5], dup[15];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      up[c] = s_u[c * np + p];
      double d[3];
      logical_derivs(s_u + c * np, m, D0, D1, D2, i, j, k, d);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        double s = ij[ij3(0, a) * np + p] * d[0];
        s = fma(ij[ij3(1, a) * np + p], d[1], s);
        s = fma(ij[ij3(2, a) * np + p], d[2], s);
        dup[3 * c + a] = s;
      }
    }
    double dtp[5];
    sw_time_derivative(dtp, up, dup, gamma2[m.off + p]);
    if constexpr (CPSR) {
      const int at[3] = {i, j, k};
      for (int dir = 0; dir < 6; ++dir) {
        const int d = dir >> 1, side = dir & 1;
        if (at[d] != (side == 0 ? 0 : ext_at(m, d) - 1) || nbr[6 * e + dir] >= 0) continue;
        double n[3], mag;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          n[a] = ij[ij3(d, a) * np + p];
          if (side == 0) n[a] = n[a] * -1.0;
        }
        flat_normalize(n, &mag);
        dtp[1] = dtp[1] + sw_cpsr_correction(bc_kind, up, dup, n, gamma2[m.off + p],
                                             coords != nullptr ? coords + 3 * m.off + p
                                                               : nullptr,
                                             np);
      }
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) dt[5 * m.off + c * np + p] = dtp[c];
  }
}

// ---------------------------------------------------------------------------
// Partial derivatives of nc components (GH volume terms).  Grid: (list
// length, component chunks of kRaggedChunk); dynamic LDS: the element's D
// matrices (kStagedDiffDoubles), then Np_max doubles.
constexpr int kRaggedChunk = 5;

__global__ void __launch_bounds__(256)
ragged_partial_derivatives_kernel(const MixedElement* __restrict__ el,
                                  const int32_t* __restrict__ list, int nc,
                                  const double* __restrict__ u,
                                  const double* __restrict__ inv_jac, double* __restrict__ du,
                                  const double* __restrict__ dtab) {
  extern __shared__ double s_lds[];
  double* s_c = s_lds + kStagedDiffDoubles;
  const MixedElement m = el[list[blockIdx.x]];
  const int np = m.np;
  const double* Dm[3];
  stage_diff_matrices(dtab, m.n, s_lds, Dm);  // synchronised by the loop's first barrier
  const double* D0 = Dm[0];
  const double* D1 = Dm[1];
  const double* D2 = Dm[2];
  const double* ij = inv_jac + 9 * m.off;
  const int c_end = min(nc, static_cast<int>(blockIdx.y + 1) * kRaggedChunk);
  for (int c = blockIdx.y * kRaggedChunk; c < c_end; ++c) {
    __syncthreads();
    for (int p = threadIdx.x; p < np; p += blockDim.x) s_c[p] = u[nc * m.off + c * np + p];
    __syncthreads();
    for (int p = threadIdx.x; p < np; p += blockDim.x) {
      const int i = p % m.n[0], j = (p / m.n[0]) % m.n[1], k = p / (m.n[0] * m.n[1]);
      double d[3];
      logical_derivs(s_c, m, D0, D1, D2, i, j, k, d);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        double s = ij[ij3(0, a) * np + p] * d[0];
        s = fma(ij[ij3(1, a) * np + p], d[1], s);
        s = fma(ij[ij3(2, a) * np + p], d[2], s);
        du[3 * nc * m.off + (3 * c + a) * np + p] = s;
      }
    }
  }
}

// GH point-wise terms from u and du (ragged layouts).
__global__ void __launch_bounds__(256)
gh_mixed_pointwise_kernel(const MixedElement* __restrict__ el, const int32_t* __restrict__ list,
                          const double* __restrict__ u, const double* __restrict__ du,
                          const double* __restrict__ gammas, double* __restrict__ dt) {
  const MixedElement m = el[list[blockIdx.x]];
  const int np = m.np;
  for (int p = threadIdx.x; p < np; p += blockDim.x) {
    const double* ue = u + 50 * m.off + p;
    const double* due = du + 150 * m.off + p;
    double up[50];
#pragma unroll
    for (int c = 0; c < 50; ++c) up[c] = ue[c * np];
    const double* ga = gammas + 3 * m.off + p;
    GhPointState st;
    gh_point_setup(st, up, ga[0], ga[np], ga[2 * np]);
    double* dte = dt + 50 * m.off + p;
#pragma unroll
    for (int pr = 0; pr < 10; ++pr) {
      double phi[3], dg[3], dpi[3], dphi[9], out[5];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        phi[i] = up[20 + i + 3 * pr];
        dg[i] = due[(3 * pr + i) * np];
        dpi[i] = due[(30 + 3 * pr + i) * np];
      }
#pragma unroll
      for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int mm = 0; mm < 3; ++mm) dphi[mm + 3 * q] = due[(60 + 3 * (q + 3 * pr) + mm) * np];
      gh_point_pair(out, st, st.dt_pi_partial[pr], up[10 + pr], phi, dg, dpi, dphi);
      dte[pr * np] = out[0];
      dte[(10 + pr) * np] = out[1];
#pragma unroll
      for (int i = 0; i < 3; ++i) dte[(20 + i + 3 * pr) * np] = out[2 + i];
    }
  }
}

// ---------------------------------------------------------------------------
// Face coupling.
//
// Apply per-dimension matrices to NC components on a 2-d mesh, as the
// oracle's apply_face_matrices: in [NC][ia ib] -> out [NC][oa ob] through
// tmp [NC][oa ib]; Ma [oa][ia] (skipped when ia == oa), Mb [ob][ib]
// (skipped when ib == ob).  Each work item is one output point for a group
// of NC / G components (G groups): the matrix row is read once for the group
// and the index arithmetic done once.  Every output is the same sum over
// l = 0 .. n-1 (first term a product, then fused multiply-adds).  Includes
// the barriers between the steps.
// Dimension a of apply_face_matrices alone (no barrier): out [NC][oa ib] =
// Ma applied to in [NC][ia ib].
template <int NC, int G>
__device__ void face_matrix_a(int ia, int ib, int oa, const double* Ma, const double* in,
                              double* out) {
  constexpr int NG = (NC + G - 1) / G;
  for (int x = threadIdx.x; x < G * oa * ib; x += blockDim.x) {
    const int g = x / (oa * ib), y = x % (oa * ib), b = y / oa, r = y % oa;
    const double* v = in + g * NG * ia * ib + ia * b;
    const double* row = Ma + r * ia;
    double acc[NG];
#pragma unroll
    for (int c = 0; c < NG; ++c) acc[c] = g * NG + c < NC ? row[0] * v[c * ia * ib] : 0.0;
    for (int l = 1; l < ia; ++l) {
      const double m = row[l];
#pragma unroll
      for (int c = 0; c < NG; ++c)
        if (g * NG + c < NC) acc[c] = fma(m, v[c * ia * ib + l], acc[c]);
    }
#pragma unroll
    for (int c = 0; c < NG; ++c)
      if (g * NG + c < NC) out[(g * NG + c) * oa * ib + y] = acc[c];
  }
}

// Dimension b of apply_face_matrices at one output point (ra, rb): vals[c] =
// Mb applied to src [NC][oa ib] (src itself where ib == ob).
template <int NC>
__device__ __forceinline__ void face_matrix_b_point(int ib, int oa, int ob, const double* Mb,
                                                    const double* src, int ra, int rb,
                                                    double* vals) {
  const double* v = src + ra;
  if (ib == ob) {
#pragma unroll
    for (int c = 0; c < NC; ++c) vals[c] = v[c * oa * ib + oa * rb];
    return;
  }
  const double* row = Mb + rb * ib;
#pragma unroll
  for (int c = 0; c < NC; ++c) vals[c] = row[0] * v[c * oa * ib];
  for (int l = 1; l < ib; ++l) {
    const double m = row[l];
#pragma unroll
    for (int c = 0; c < NC; ++c) vals[c] = fma(m, v[c * oa * ib + oa * l], vals[c]);
  }
}

template <int NC, int G>
__device__ void apply_face_matrices(int ia, int ib, int oa, int ob, const double* Ma,
                                    const double* Mb, const double* in, double* tmp,
                                    double* out) {
  constexpr int NG = (NC + G - 1) / G;
  const int t = threadIdx.x, T = blockDim.x;
  const double* src = in;
  if (ia != oa) {
    face_matrix_a<NC, G>(ia, ib, oa, Ma, in, tmp);
    __syncthreads();
    src = tmp;
  }
  for (int x = t; x < G * oa * ob; x += T) {
    const int g = x / (oa * ob), y = x % (oa * ob), r = y / oa, a = y % oa;
    const double* v = src + g * NG * oa * ib + a;
    double acc[NG];
    if (ib == ob) {
#pragma unroll
      for (int c = 0; c < NG; ++c) acc[c] = g * NG + c < NC ? v[c * oa * ib + oa * r] : 0.0;
    } else {
      const double* row = Mb + r * ib;
#pragma unroll
      for (int c = 0; c < NG; ++c) acc[c] = g * NG + c < NC ? row[0] * v[c * oa * ib] : 0.0;
      for (int l = 1; l < ib; ++l) {
        const double m = row[l];
#pragma unroll
        for (int c = 0; c < NG; ++c)
          if (g * NG + c < NC) acc[c] = fma(m, v[c * oa * ib + oa * l], acc[c]);
      }
    }
#pragma unroll
    for (int c = 0; c < NG; ++c)
      if (g * NG + c < NC) out[(g * NG + c) * oa * ob + y] = acc[c];
  }
  __syncthreads();
}

// LDS plan of the face kernel for faces of at most kf points and mortars of
// at most km points (doubles).
struct FacePlan {
  int side_own, side_nb, pk_own, pk_nb, tmp, m_own, m_nb, lam_own, lam_nb, bc_m, bc, total;
  __host__ __device__ FacePlan(int kf, int km, int npk) {
    side_own = 0;                       // 12 per own face point
    side_nb = side_own + 12 * kf;       // 11 per neighbour face point
    pk_own = side_nb + 11 * kf;         // npk per own face point
    pk_nb = pk_own + npk * kf;          // npk per neighbour face point
    tmp = pk_nb + npk * kf;             // npk per mortar point
    m_own = tmp + npk * km;             // packages on the mortar
    m_nb = m_own + npk * km;
    lam_own = m_nb + npk * km;          // GH speeds on the mortar
    lam_nb = lam_own + 4 * km;
    bc_m = lam_nb + 4 * km;             // 5 corrections per mortar point
    bc = bc_m + 5 * km;                 // 5 per own face point
    total = bc + 5 * kf;
  }
};

// Side data of face point f: SW n[3], mag; GH n[3], nv[3], lam[4], gamma2 (+ mag).
template <bool GH>
__device__ void face_side(const MixedElement& m, int d, int side, bool flip, int f,
                          const double* __restrict__ u, const double* __restrict__ params,
                          const double* __restrict__ inv_jac, double* out, bool own) {
  const int na = d == 0 ? m.n[1] : m.n[0];
  const int p = face_point(m, d, side, f % na, f / na);
  const int np = m.np;
  double n[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    n[i] = inv_jac[9 * m.off + ij3(d, i) * np + p];
    if (flip) n[i] = n[i] * -1.0;
  }
  if constexpr (!GH) {
    double mag;
    flat_normalize(n, &mag);
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] = n[i];
    out[3] = mag;
  } else {
    double g[10];
#pragma unroll
    for (int c = 0; c < 10; ++c) g[c] = u[50 * m.off + c * np + p];
    double lapse, shift[3], inv_sp[6], mag;
    gh_lapse_shift(&lapse, shift, inv_sp, g);
    const double* ga = params + 3 * m.off + p;
    GhFaceSide s;
#pragma unroll
    for (int i = 0; i < 3; ++i) s.n[i] = n[i];
    gh_normalize(s.n, s.nv, &mag, inv_sp);
    gh_char_speeds(s.lam, ga[np], lapse, shift, s.n);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      out[i] = s.n[i];
      out[3 + i] = s.nv[i];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) out[6 + i] = s.lam[i];
    out[10] = ga[2 * np];
    if (own) out[11] = mag;
  }
}

// Package of pass `pr` (GH pair; SW: the single pass) at face point f into
// pk[c * nf + f].
template <bool GH>
__device__ void face_package(const MixedElement& m, int d, int side, int f, int nf, int pr,
                             const double* __restrict__ u, const double* __restrict__ params,
                             const double* sd, double* pk) {
  const int na = d == 0 ? m.n[1] : m.n[0];
  const int p = face_point(m, d, side, f % na, f / na);
  const int np = m.np;
  double out[13];
  if constexpr (!GH) {
    double up[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) up[c] = u[5 * m.off + c * np + p];
    sw_package(out, up, params[m.off + p], sd);
  } else {
    GhFaceSide s;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      s.n[i] = sd[i];
      s.nv[i] = sd[3 + i];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) s.lam[i] = sd[6 + i];
    s.gamma2 = sd[10];
    const double* ue = u + 50 * m.off + p;
    double phi[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) phi[i] = ue[(20 + i + 3 * pr) * np];
    gh_package_pair(out, ue[pr * np], ue[(10 + pr) * np], phi, s);
  }
#pragma unroll
  for (int c = 0; c < 13; ++c) pk[c * nf + f] = out[c];
}

// LDS plan of the ScalarWave face kernel (doubles): small enough for two
// workgroups per CU at 12 x 12 faces.  The neighbour side gets no mortar
// buffer (its dimension-b interpolation is formed per mortar point with the
// boundary terms), the mortar corrections overwrite the own side's mortar
// package, and the projected corrections reuse the own face package buffer.
struct SwFacePlan {
  int side_own, side_nb, pk_own, pk_nb, tmp, m_own, bc, total;
  __host__ __device__ SwFacePlan(int kf, int km) {
    side_own = 0;               // n[3], |n| per own face point
    side_nb = side_own + 4 * kf;  // the same per neighbour face point
    pk_own = side_nb + 4 * kf;  // 13 per own face point; later 5 corrections
    pk_nb = pk_own + 13 * kf;   // 13 per neighbour face point
    tmp = pk_nb + 13 * kf;      // 13 per mortar point
    m_own = tmp + 13 * km;      // own package on the mortar; then 5 corrections
    bc = pk_own;
    total = m_own + 13 * km;
  }
};

// ScalarWave faces: one workgroup per element, its six faces in direction
// order, external faces (neighbour -1) first.  An external face forms its
// exterior state at its own points by bc_kind (sw_ghost_state: the ghost
// array for the Dirichlet analytic condition, ragged [5][Np_e]; the interior
// trace for SphericalRadiation, ghost = ragged [3][Np_e] coordinates for
// Bayliss-Turkel) with the own gamma2 and the reversed unit normal, and
// needs no mortar; for the kSwBcCpsr* kinds it is skipped (the volume kernel
// applied the condition).  Dynamic LDS: SwFacePlan(kf, km).total doubles.
__global__ void __launch_bounds__(256)
sw_mixed_face_kernel(const MixedElement* __restrict__ el, const int32_t* __restrict__ list,
                     const int32_t* __restrict__ nbr, const double* __restrict__ u,
                     const double* __restrict__ params, const double* __restrict__ inv_jac,
                     double* __restrict__ dt, const double* __restrict__ mortar_tab, int kf,
                     int km, const double* __restrict__ ghost, int bc_kind) {
  extern __shared__ double sm[];
  const SwFacePlan L(kf, km);
  const int ghost_blocks = bc_kind == kSwBcGhostArray ? 5 : 3;
  const int t = threadIdx.x, T = blockDim.x;
  const int e = list[blockIdx.x];
  const MixedElement m = el[e];
  for (int step = 0; step < 12; ++step) {
    const int dir = step % 6;
    const int en = nbr[6 * e + dir];
    if ((en < 0) != (step < 6)) continue;  // the same for the whole workgroup
    const int d = dir >> 1, side = dir & 1;
    if (en < 0) {
      // the constraint-preserving condition was applied with the volume terms
      if (bc_kind >= kSwBcCpsrSommerfeld) continue;  // uniform over the workgroup
      __syncthreads();  // the previous face's dt updates are done
      const int na = d == 0 ? m.n[1] : m.n[0], nb = d == 2 ? m.n[1] : m.n[2];
      for (int f = t; f < na * nb; f += T) {
        const int p = face_point(m, d, side, f % na, f / na);
        double ui[5], ug[5], n[3], ng[3], mag, pki[13], pke[13], bc[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) ui[c] = u[5 * m.off + c * m.np + p];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          n[i] = inv_jac[9 * m.off + ij3(d, i) * m.np + p];
          if (side == 0) n[i] = n[i] * -1.0;
        }
        flat_normalize(n, &mag);
        sw_ghost_state(ug, bc_kind, ui, n,
                       ghost != nullptr ? ghost + ghost_blocks * m.off + p : nullptr, m.np);
        const double g2 = params[m.off + p];
        sw_package(pki, ui, g2, n);
#pragma unroll
        for (int i = 0; i < 3; ++i) ng[i] = n[i] * -1.0;
        sw_package(pke, ug, g2, ng);
        sw_boundary_terms(bc, pki, pke);
        const double fac = lift_factor(ext_at(m, d), mag);
#pragma unroll
        for (int c = 0; c < 5; ++c) {
          double& target = dt[5 * m.off + c * m.np + p];
          target = target + bc[c] * fac;
        }
      }
      continue;
    }
    const MixedElement mn = el[en];
    const int na = d == 0 ? m.n[1] : m.n[0], nb = d == 2 ? m.n[1] : m.n[2];
    const int ma = d == 0 ? mn.n[1] : mn.n[0], mb = d == 2 ? mn.n[1] : mn.n[2];
    const int ka = max(na, ma), kb = max(nb, mb);
    const int nf = na * nb, mf = ma * mb, kmf = ka * kb;
    const double* Ia_own = mortar_tab + mortar_table_offset(na, ka);
    const double* Ib_own = mortar_tab + mortar_table_offset(nb, kb);
    const double* Ia_nb = mortar_tab + mortar_table_offset(ma, ka);
    const double* Ib_nb = mortar_tab + mortar_table_offset(mb, kb);
    const double* Pa = Ia_own + na * ka;  // projection block follows the interpolation
    const double* Pb = Ib_own + nb * kb;
    __syncthreads();  // previous face done with the LDS
    for (int f = t; f < nf; f += T)
      face_side<false>(m, d, side, side == 0, f, u, params, inv_jac, sm + L.side_own + 4 * f,
                       true);
    for (int f = t; f < mf; f += T)
      face_side<false>(mn, d, 1 - side, side == 1, f, u, params, inv_jac,
                       sm + L.side_nb + 4 * f, false);
    __syncthreads();
    for (int f = t; f < nf; f += T)
      face_package<false>(m, d, side, f, nf, 0, u, params, sm + L.side_own + 4 * f,
                          sm + L.pk_own);
    for (int f = t; f < mf; f += T)
      face_package<false>(mn, d, 1 - side, f, mf, 0, u, params, sm + L.side_nb + 4 * f,
                          sm + L.pk_nb);
    __syncthreads();
    apply_face_matrices<13, 2>(na, nb, ka, kb, Ia_own, Ib_own, sm + L.pk_own, sm + L.tmp,
                        sm + L.m_own);
    // the neighbour side on the mortar: dimension a through tmp, then
    // dimension b for each mortar point where its boundary terms are formed
    // (apply_face_matrices' sums, without a mortar buffer for this side)
    const double* src_nb = sm + L.pk_nb;
    if (ma != ka) {
      face_matrix_a<13, 2>(ma, mb, ka, Ia_nb, sm + L.pk_nb, sm + L.tmp);
      __syncthreads();
      src_nb = sm + L.tmp;
    }
    // boundary terms at the mortar points, written over the own side's slots
    for (int k = t; k < kmf; k += T) {
      double in[13], ex[13], bc[5];
      face_matrix_b_point<13>(mb, ka, kb, Ib_nb, src_nb, k % ka, k / ka, ex);
#pragma unroll
      for (int c = 0; c < 13; ++c) in[c] = sm[L.m_own + c * kmf + k];
      sw_boundary_terms(bc, in, ex);
#pragma unroll
      for (int c = 0; c < 5; ++c) sm[L.m_own + c * kmf + k] = bc[c];
    }
    __syncthreads();
    apply_face_matrices<5, 1>(ka, kb, na, nb, Pa, Pb, sm + L.m_own, sm + L.tmp, sm + L.bc);
    // lift into dt
    for (int f = t; f < nf; f += T) {
      const int p = face_point(m, d, side, f % na, f / na);
      const double fac = lift_factor(ext_at(m, d), sm[L.side_own + 4 * f + 3]);
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        double& target = dt[5 * m.off + c * m.np + p];
        target = target + sm[L.bc + c * nf + f] * fac;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Generalized harmonic adjoint on mixed-p domains (the RHS linearised at u):
// the kernels of the uniform adjoint with per-element extents, and a face
// kernel that transposes the mortar coupling.
//
// Point-wise adjoint, one thread per grid point (point_el gives its
// element): ubar = (df/du)^T gbar, and the logical-derivative adjoints xbar
// written over du (read first, same point).
__global__ void __launch_bounds__(128)
gh_mixed_adjoint_pointwise_kernel(int64_t total, const MixedElement* __restrict__ el,
                                  const int32_t* __restrict__ point_el,
                                  const double* __restrict__ u, double* du_xbar,
                                  const double* __restrict__ gammas,
                                  const double* __restrict__ inv_jac,
                                  const double* __restrict__ gbar, double* __restrict__ ubar) {
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  const MixedElement m = el[point_el[gid]];
  const int np = m.np;
  const int p = static_cast<int>(gid - m.off);
  const double* gp = gbar + 50 * m.off + p;
  const double* ga = gammas + 3 * m.off + p;
  const double gamma1 = ga[np], gamma2 = ga[2 * np];
  double ub[50];
  GhVjpPoint pt;
  gh_time_derivative_vjp(u + 50 * m.off + p, np, du_xbar + 150 * m.off + p, np, ga[0], gamma1,
                         gamma2, gp, np, ub, &pt);
#pragma unroll
  for (int c = 0; c < 50; ++c) ubar[50 * m.off + c * np + p] = ub[c];
  double ij[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) ij[c] = inv_jac[9 * m.off + c * np + p];
  for (int c = 0; c < 50; ++c) {
    double dub[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
      dub[i] = gh_time_derivative_dubar(pt, gamma1, gamma2, gp, np, 3 * c + i);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i) s += ij[ij3(j, i)] * dub[i];
      du_xbar[150 * m.off + (3 * c + j) * np + p] = s;
    }
  }
}

// ubar_c += sum_j D_j^T xbar_cj per element.  Grid: (list length, component
// chunks of kRaggedChunk); dynamic LDS: the element's D matrices
// (kStagedDiffDoubles), then 3 Np_max doubles.
__global__ void __launch_bounds__(256)
ragged_transposed_sweeps_kernel(const MixedElement* __restrict__ el,
                                const int32_t* __restrict__ list, int nc,
                                const double* __restrict__ xbar, double* __restrict__ ubar,
                                const double* __restrict__ dtab) {
  extern __shared__ double s_lds[];
  double* s_x = s_lds + kStagedDiffDoubles;
  const MixedElement m = el[list[blockIdx.x]];
  const int np = m.np;
  const int stride[3] = {1, m.n[0], m.n[0] * m.n[1]};
  const double* D[3];
  stage_diff_matrices(dtab, m.n, s_lds, D);  // synchronised by the loop's first barrier
  const int c_end = min(nc, static_cast<int>(blockIdx.y + 1) * kRaggedChunk);
  for (int c = blockIdx.y * kRaggedChunk; c < c_end; ++c) {
    __syncthreads();
    for (int x = threadIdx.x; x < 3 * np; x += blockDim.x)
      s_x[x] = xbar[3 * nc * m.off + 3 * c * np + x];
    __syncthreads();
    for (int p = threadIdx.x; p < np; p += blockDim.x) {
      const int idx[3] = {p % m.n[0], (p / m.n[0]) % m.n[1], p / (m.n[0] * m.n[1])};
      double acc = ubar[nc * m.off + c * np + p];
      for (int j = 0; j < 3; ++j) {
        const int base = p - idx[j] * stride[j];
        for (int l = 0; l < m.n[j]; ++l)
          acc = fma(D[j][l * m.n[j] + idx[j]], s_x[j * np + base + l * stride[j]], acc);
      }
      ubar[nc * m.off + c * np + p] = acc;
    }
  }
}

// Transpose of apply_face_matrices: in [NC][oa ob] -> out [NC][ia ib] with
// Ma [oa][ia] (skipped when ia == oa) and Mb [ob][ib] (skipped when
// ib == ob), dimension b first, through tmp [NC][oa ib]; work items as in
// apply_face_matrices.  Includes the barriers between the steps.
template <int NC, int G>
__device__ void apply_face_matrices_t(int oa, int ob, int ia, int ib, const double* Ma,
                                      const double* Mb, const double* in, double* tmp,
                                      double* out) {
  constexpr int NG = (NC + G - 1) / G;
  const int t = threadIdx.x, T = blockDim.x;
  const double* src = in;
  if (ib != ob) {
    for (int x = t; x < G * oa * ib; x += T) {
      const int g = x / (oa * ib), y = x % (oa * ib), lb = y / oa, ra = y % oa;
      const double* v = in + g * NG * oa * ob + ra;
      double acc[NG];
#pragma unroll
      for (int c = 0; c < NG; ++c) acc[c] = g * NG + c < NC ? Mb[lb] * v[c * oa * ob] : 0.0;
      for (int r = 1; r < ob; ++r) {
        const double m = Mb[r * ib + lb];
#pragma unroll
        for (int c = 0; c < NG; ++c)
          if (g * NG + c < NC) acc[c] = fma(m, v[c * oa * ob + oa * r], acc[c]);
      }
#pragma unroll
      for (int c = 0; c < NG; ++c)
        if (g * NG + c < NC) tmp[(g * NG + c) * oa * ib + y] = acc[c];
    }
    __syncthreads();
    src = tmp;
  }
  for (int x = t; x < G * ia * ib; x += T) {
    const int g = x / (ia * ib), y = x % (ia * ib), lb = y / ia, la = y % ia;
    const double* v = src + g * NG * oa * ib + oa * lb;
    double acc[NG];
    if (ia == oa) {
#pragma unroll
      for (int c = 0; c < NG; ++c) acc[c] = g * NG + c < NC ? v[c * oa * ib + la] : 0.0;
    } else {
#pragma unroll
      for (int c = 0; c < NG; ++c) acc[c] = g * NG + c < NC ? Ma[la] * v[c * oa * ib] : 0.0;
      for (int r = 1; r < oa; ++r) {
        const double m = Ma[r * ia + la];
#pragma unroll
        for (int c = 0; c < NG; ++c)
          if (g * NG + c < NC) acc[c] = fma(m, v[c * oa * ib + r], acc[c]);
      }
    }
#pragma unroll
    for (int c = 0; c < NG; ++c)
      if (g * NG + c < NC) out[(g * NG + c) * ia * ib + y] = acc[c];
  }
  __syncthreads();
}

// Side values of face point f of element m's face (d, side); the raw normal
// is negated when `flip`.
__device__ void face_side_adjoint(const MixedElement& m, int d, int side, bool flip, int f,
                                  const double* __restrict__ u, const double* __restrict__ params,
                                  const double* __restrict__ inv_jac, GhSideFwd& s, int* point) {
  const int na = d == 0 ? m.n[1] : m.n[0];
  const int p = face_point(m, d, side, f % na, f / na);
  double n[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    n[i] = inv_jac[9 * m.off + ij3(d, i) * m.np + p];
    if (flip) n[i] = n[i] * -1.0;
  }
  gh_side_forward(s, u + 50 * m.off + p, m.np, params[3 * m.off + m.np + p], n);
  *point = p;
}

// Side values of every face point of the faces normal to dimension d, both
// sides of every element of the list, as face_side_adjoint forms them (the
// normal taken outward): n, nv, lam[4], gamma2, |n| at sides[12 (off + p)].
// Grid: (list length, 2 sides).
__global__ void __launch_bounds__(128)
gh_mixed_face_sides_kernel(int d, const MixedElement* __restrict__ el,
                           const int32_t* __restrict__ list, const double* __restrict__ u,
                           const double* __restrict__ params,
                           const double* __restrict__ inv_jac, double* __restrict__ sides) {
  const int side = blockIdx.y;
  const MixedElement m = el[list[blockIdx.x]];
  const int na = d == 0 ? m.n[1] : m.n[0], nb = d == 2 ? m.n[1] : m.n[2];
  for (int f = threadIdx.x; f < na * nb; f += blockDim.x) {
    GhSideFwd s;
    int p;
    face_side_adjoint(m, d, side, side == 0, f, u, params, inv_jac, s, &p);
    double* o = sides + 12 * (m.off + p);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      o[i] = s.n[i];
      o[3 + i] = s.nv[i];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) o[6 + k] = s.lam[k];
    o[10] = params[3 * m.off + 2 * m.np + p];
    o[11] = s.mag;
  }
}

// GH faces, one pass per workgroup: grid (list length, 2 sides, 10 pairs) for
// the faces normal to dimension d, so every (face, pair) runs concurrently.
// Each workgroup packages its pair on both face meshes together with the
// speeds (17 components), interpolates them to the mortar, evaluates the
// boundary terms there, projects back and lifts into dt.  The lower and
// upper faces of one dimension hold different points and the pairs write
// different components, so no two workgroups of a launch touch the same dt
// value; launching d = 0, 1, 2 in order adds each point's corrections in
// direction order.  Every value is computed with the arithmetic of the
// oracle's mortar coupling.
// kFusedNb: the neighbour side's dimension-b interpolation is done per mortar
// point where its boundary terms are formed, without a mortar buffer for that
// side (less LDS, so more workgroups per CU, but one thread per mortar point
// in that phase); otherwise as apply_face_matrices into its own buffer.
template <bool kFusedNb>
struct FacePassPlan {
  int face, mag, tmp, m_own, m_nb, total;
  __host__ __device__ FacePassPlan(int kf, int km) {
    face = 0;               // 17 per face point: speeds, package; then 5 corrections
    mag = face + 17 * kf;   // |n| per own face point
    tmp = mag + kf;         // 17 per mortar point
    m_own = tmp + 17 * km;  // own side on the mortar; then the mortar corrections
    m_nb = m_own + 17 * km;
    total = m_nb + (kFusedNb ? 0 : 17 * km);
  }
};

template <bool kFusedNb>
__global__ void __launch_bounds__(256, 3)
gh_mixed_face_pass_kernel(int d, const MixedElement* __restrict__ el,
                          const int32_t* __restrict__ list, const int32_t* __restrict__ nbr,
                          const double* __restrict__ u, const double* __restrict__ params,
                          const double* __restrict__ inv_jac, double* __restrict__ dt,
                          const double* __restrict__ mortar_tab, int kf, int km,
                          const double* __restrict__ ghost, int phase) {
  extern __shared__ double sm[];
  const FacePassPlan<kFusedNb> L(kf, km);
  const int t = threadIdx.x, T = blockDim.x;
  const int side = blockIdx.y, pr = blockIdx.z;
  const int e = list[blockIdx.x];
  const MixedElement m = el[e];
  const int en = nbr[6 * e + 2 * d + side];
  // the same for the whole workgroup
  if ((phase == kFacesExternal && en >= 0) || (phase == kFacesInternal && en < 0)) return;
  if (en < 0) {
    // external face (Dirichlet analytic): the ghost state at the face's own
    // points, the element's damping parameters and the reversed unit normal
    // renormalised with the ghost inverse spatial metric; no mortar
    const int na = d == 0 ? m.n[1] : m.n[0], nb = d == 2 ? m.n[1] : m.n[2];
    const int np = m.np;
    for (int f = t; f < na * nb; f += T) {
      const int p = face_point(m, d, side, f % na, f / na);
      double sd[12];
      face_side<true>(m, d, side, side == 0, f, u, params, inv_jac, sd, true);
      GhFaceSide si, se;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        si.n[i] = sd[i];
        si.nv[i] = sd[3 + i];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) si.lam[i] = sd[6 + i];
      si.gamma2 = sd[10];
      const double* ux = ghost + 50 * m.off + p;
      double g[10], lapse, shift[3], inv_sp[6];
#pragma unroll
      for (int c = 0; c < 10; ++c) g[c] = ux[c * np];
      gh_lapse_shift(&lapse, shift, inv_sp, g);
#pragma unroll
      for (int i = 0; i < 3; ++i) se.n[i] = si.n[i] * -1.0;
      gh_ghost_normal(se.n, se.nv, inv_sp);
      gh_char_speeds(se.lam, params[3 * m.off + np + p], lapse, shift, se.n);
      se.gamma2 = si.gamma2;
      const double* ue = u + 50 * m.off + p;
      double phi_i[3], phi_e[3], pki[13], pke[13], bc[5];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        phi_i[i] = ue[(20 + i + 3 * pr) * np];
        phi_e[i] = ux[(20 + i + 3 * pr) * np];
      }
      gh_package_pair(pki, ue[pr * np], ue[(10 + pr) * np], phi_i, si);
      gh_package_pair(pke, ux[pr * np], ux[(10 + pr) * np], phi_e, se);
      gh_boundary_terms_pair(bc, pki, si.lam, pke, se.lam);
      const double fac = lift_factor(ext_at(m, d), sd[11]);
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        const int comp = c == 0 ? pr : c == 1 ? 10 + pr : 18 + c + 3 * pr;
        double& target = dt[50 * m.off + comp * np + p];
        target = target + bc[c] * fac;
      }
    }
    return;
  }
  const MixedElement mn = el[en];
  const int na = d == 0 ? m.n[1] : m.n[0], nb = d == 2 ? m.n[1] : m.n[2];
  const int ma = d == 0 ? mn.n[1] : mn.n[0], mb = d == 2 ? mn.n[1] : mn.n[2];
  const int ka = max(na, ma), kb = max(nb, mb);
  const int nf = na * nb, mf = ma * mb, kmf = ka * kb;
  const double* Ia_own = mortar_tab + mortar_table_offset(na, ka);
  const double* Ib_own = mortar_tab + mortar_table_offset(nb, kb);
  const double* Ia_nb = mortar_tab + mortar_table_offset(ma, ka);
  const double* Ib_nb = mortar_tab + mortar_table_offset(mb, kb);
  const double* Pa = Ia_own + na * ka;  // projection block follows the interpolation
  const double* Pb = Ib_own + nb * kb;
  double* face = sm + L.face;
  // own side: speeds and package of the pair on the face mesh
  for (int f = t; f < nf; f += T) {
    double sd[12];
    face_side<true>(m, d, side, side == 0, f, u, params, inv_jac, sd, true);
#pragma unroll
    for (int c = 0; c < 4; ++c) face[c * nf + f] = sd[6 + c];
    sm[L.mag + f] = sd[11];
    face_package<true>(m, d, side, f, nf, pr, u, params, sd, face + 4 * nf);
  }
  __syncthreads();
  apply_face_matrices<17, 2>(na, nb, ka, kb, Ia_own, Ib_own, face, sm + L.tmp, sm + L.m_own);
  // neighbour side
  for (int f = t; f < mf; f += T) {
    double sd[12];
    face_side<true>(mn, d, 1 - side, side == 1, f, u, params, inv_jac, sd, false);
#pragma unroll
    for (int c = 0; c < 4; ++c) face[c * mf + f] = sd[6 + c];
    face_package<true>(mn, d, 1 - side, f, mf, pr, u, params, sd, face + 4 * mf);
  }
  __syncthreads();
  if constexpr (!kFusedNb) {
    apply_face_matrices<17, 2>(ma, mb, ka, kb, Ia_nb, Ib_nb, face, sm + L.tmp, sm + L.m_nb);
    // boundary terms at the mortar points, written over the own side's slots
    for (int k = t; k < kmf; k += T) {
      double in[13], ex[13], li[4], le[4], bc[5];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        li[c] = sm[L.m_own + c * kmf + k];
        le[c] = sm[L.m_nb + c * kmf + k];
      }
#pragma unroll
      for (int c = 0; c < 13; ++c) {
        in[c] = sm[L.m_own + (4 + c) * kmf + k];
        ex[c] = sm[L.m_nb + (4 + c) * kmf + k];
      }
      gh_boundary_terms_pair(bc, in, li, ex, le);
#pragma unroll
      for (int c = 0; c < 5; ++c) sm[L.m_own + c * kmf + k] = bc[c];
    }
    __syncthreads();
  } else {
    // the neighbour side on the mortar: dimension a through tmp, then
    // dimension b for each mortar point where its boundary terms are formed
    // (apply_face_matrices' sums, without a mortar buffer for this side)
    const double* src_nb = face;
    if (ma != ka) {
      face_matrix_a<17, 2>(ma, mb, ka, Ia_nb, face, sm + L.tmp);
      __syncthreads();
      src_nb = sm + L.tmp;
    }
    // boundary terms at the mortar points, written over the own side's slots
    for (int k = t; k < kmf; k += T) {
      double nbv[17];
      face_matrix_b_point<17>(mb, ka, kb, Ib_nb, src_nb, k % ka, k / ka, nbv);
      double in[13], li[4], bc[5];
#pragma unroll
      for (int c = 0; c < 4; ++c) li[c] = sm[L.m_own + c * kmf + k];
#pragma unroll
      for (int c = 0; c < 13; ++c) in[c] = sm[L.m_own + (4 + c) * kmf + k];
      gh_boundary_terms_pair(bc, in, li, nbv + 4, nbv);
#pragma unroll
      for (int c = 0; c < 5; ++c) sm[L.m_own + c * kmf + k] = bc[c];
    }
    __syncthreads();
  }
  apply_face_matrices<5, 1>(ka, kb, na, nb, Pa, Pb, sm + L.m_own, sm + L.tmp, face);
  // lift into dt: components g, Pi, Phi_0..2 of the pair
  for (int f = t; f < nf; f += T) {
    const int p = face_point(m, d, side, f % na, f / na);
    const double fac = lift_factor(ext_at(m, d), sm[L.mag + f]);
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      const int comp = c == 0 ? pr : c == 1 ? 10 + pr : 18 + c + 3 * pr;
      double& target = dt[50 * m.off + comp * m.np + p];
      target = target + face[c * nf + f] * fac;
    }
  }
}

// GH face adjoint, one pass per workgroup: grid (list length, 2 sides, 10
// pairs) for the faces normal to dimension d, as gh_mixed_face_pass_kernel.
// Workgroup (e, side, pr) adds to e's points on face (d, side) the adjoint,
// for pair pr, of e's own lifted correction and of the neighbour's (in which
// e is the exterior side): the forward is recomputed up to the mortar
// corrections, then
//   seeds   gbar * lift factor on both faces (the neighbour's with its own),
//   P^T     each seed projected back onto the mortar (transposed projection),
//   weights the adjoint of e's package on the mortar, interior for its own
//           correction and exterior for the neighbour's, from the upwind
//           weights of the mortar speeds,
//   I^T     transposed interpolation onto e's face,
//   package reverse at e's face points into ubar (the pair's components).
// The adjoints of e's normals, speeds and |n| from this pair go to side_adj
// (11 per pair at each face point, indexed by the point's volume index);
// gh_mixed_side_adjoint_kernel sums them over the pairs and takes them to g.
// Dynamic LDS: FaceAdjPassPlan(kf, km).total doubles.
// kFusedNb as in FacePassPlan: without a mortar buffer for the neighbour side.
template <bool kFusedNb>
struct FaceAdjPassPlan {
  int side, mag_nb, face, tmp, m_own, m_nb, total;
  __host__ __device__ FaceAdjPassPlan(int kf, int km) {
    side = 0;                // own face point: n, nv, lam, gamma2, |n|
    mag_nb = side + 12 * kf;  // |n| of the neighbour's face points
    face = mag_nb + kf;       // 17 per face point (speeds, package); then
                              // corrections, seeds, package adjoint
    tmp = face + 17 * kf;
    m_own = tmp + 17 * km;   // own speeds, package; then corrections, seeds on
                             // the mortar, package adjoint (all in place)
    m_nb = m_own + 17 * km;  // the neighbour side on the mortar
    total = m_nb + (kFusedNb ? 0 : 17 * km);
  }
};

template <bool kFusedNb>
__global__ void __launch_bounds__(256, 3)
gh_mixed_face_adjoint_pass_kernel(int d, const MixedElement* __restrict__ el,
                                  const int32_t* __restrict__ list,
                                  const int32_t* __restrict__ nbr, const double* __restrict__ u,
                                  const double* __restrict__ params,
                                  const double* __restrict__ inv_jac,
                                  const double* __restrict__ gbar, double* __restrict__ ubar,
                                  double* __restrict__ side_adj,
                                  const double* __restrict__ sides,
                                  const double* __restrict__ mortar_tab, int kf, int km) {
  extern __shared__ double sm[];
  const FaceAdjPassPlan<kFusedNb> L(kf, km);
  const int t = threadIdx.x, T = blockDim.x;
  const int side = blockIdx.y, pr = blockIdx.z;
  const int e = list[blockIdx.x];
  const MixedElement m = el[e];
  const MixedElement mn = el[nbr[6 * e + 2 * d + side]];
  const int na = d == 0 ? m.n[1] : m.n[0], nb = d == 2 ? m.n[1] : m.n[2];
  const int ma = d == 0 ? mn.n[1] : mn.n[0], mb = d == 2 ? mn.n[1] : mn.n[2];
  const int ka = max(na, ma), kb = max(nb, mb);
  const int nf = na * nb, mf = ma * mb, kmf = ka * kb;
  const double* Ia_own = mortar_tab + mortar_table_offset(na, ka);
  const double* Ib_own = mortar_tab + mortar_table_offset(nb, kb);
  const double* Ia_nb = mortar_tab + mortar_table_offset(ma, ka);
  const double* Ib_nb = mortar_tab + mortar_table_offset(mb, kb);
  const double* Pa = Ia_own + na * ka;  // projection block follows the interpolation
  const double* Pb = Ib_own + nb * kb;
  const double* Pa_nb = Ia_nb + ma * ka;
  const double* Pb_nb = Ib_nb + mb * kb;
  const double lift_own = lift_factor(ext_at(m, d), 1.0), lift_nb = lift_factor(ext_at(mn, d), 1.0);
  const int comp[5] = {pr, 10 + pr, 20 + 3 * pr, 21 + 3 * pr, 22 + 3 * pr};
  double* face = sm + L.face;
  // own side: side values (gh_mixed_face_sides_kernel), speeds and package
  // of the pair
  for (int f = t; f < nf; f += T) {
    const int p = face_point(m, d, side, f % na, f / na);
    const double* sv = sides + 12 * (m.off + p);
    double* o = sm + L.side + 12 * f;
#pragma unroll
    for (int c = 0; c < 12; ++c) o[c] = sv[c];
#pragma unroll
    for (int k = 0; k < 4; ++k) face[k * nf + f] = o[6 + k];
    double pk[13];
    gh_package_fwd(o, o + 3, o + 6, o[10], u + 50 * m.off + p, m.np, pr, pk);
#pragma unroll
    for (int c = 0; c < 13; ++c) face[(4 + c) * nf + f] = pk[c];
  }
  __syncthreads();
  apply_face_matrices<17, 2>(na, nb, ka, kb, Ia_own, Ib_own, face, sm + L.tmp, sm + L.m_own);
  for (int f = t; f < mf; f += T) {
    const int q = face_point(mn, d, 1 - side, f % ma, f / ma);
    double sv[12];
#pragma unroll
    for (int c = 0; c < 12; ++c) sv[c] = sides[12 * (mn.off + q) + c];
#pragma unroll
    for (int k = 0; k < 4; ++k) face[k * mf + f] = sv[6 + k];
    sm[L.mag_nb + f] = sv[11];
    double pk[13];
    gh_package_fwd(sv, sv + 3, sv + 6, sv[10], u + 50 * mn.off + q, mn.np, pr, pk);
#pragma unroll
    for (int c = 0; c < 13; ++c) face[(4 + c) * mf + f] = pk[c];
  }
  __syncthreads();
  // the neighbour side on the mortar (as apply_face_matrices; with kFusedNb
  // its dimension-b sums are formed per mortar point below), and e's
  // corrections there, over e's package slots
  const double* src_nb = face;
  if constexpr (!kFusedNb) {
    apply_face_matrices<17, 2>(ma, mb, ka, kb, Ia_nb, Ib_nb, face, sm + L.tmp, sm + L.m_nb);
  } else if (ma != ka) {
    face_matrix_a<17, 2>(ma, mb, ka, Ia_nb, face, sm + L.tmp);
    __syncthreads();
    src_nb = sm + L.tmp;
  }
  for (int k = t; k < kmf; k += T) {
    double in[13], li[4], nbv[17], bc[5];  // nbv: neighbour speeds, package
#pragma unroll
    for (int c = 0; c < 4; ++c) li[c] = sm[L.m_own + c * kmf + k];
#pragma unroll
    for (int c = 0; c < 13; ++c) in[c] = sm[L.m_own + (4 + c) * kmf + k];
    if constexpr (!kFusedNb) {
#pragma unroll
      for (int c = 0; c < 17; ++c) nbv[c] = sm[L.m_nb + c * kmf + k];
    } else {
      face_matrix_b_point<17>(mb, ka, kb, Ib_nb, src_nb, k % ka, k / ka, nbv);
    }
    gh_boundary_fwd(in, li, nbv + 4, nbv, bc);
#pragma unroll
    for (int c = 0; c < 5; ++c) sm[L.m_own + (4 + c) * kmf + k] = bc[c];
  }
  __syncthreads();
  apply_face_matrices<5, 1>(ka, kb, na, nb, Pa, Pb, sm + L.m_own + 4 * kmf, sm + L.tmp, face);
  // seeds on both faces; the pair's part of the |n| adjoint of e's lift factor
  double* seed_own = face + 5 * kf;
  double* seed_nb = face + 10 * kf;
  for (int f = t; f < nf; f += T) {
    const int p = face_point(m, d, side, f % na, f / na);
    const double fac = lift_own * sm[L.side + 12 * f + 11];
    double fac_b = 0.0;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      const double gb = gbar[50 * m.off + comp[c] * m.np + p];
      fac_b += gb * face[c * nf + f];
      seed_own[c * nf + f] = gb * fac;
    }
    side_adj[110 * (m.off + p) + 11 * pr + 10] = lift_own * fac_b;
  }
  for (int f = t; f < mf; f += T) {
    const int q = face_point(mn, d, 1 - side, f % ma, f / ma);
    const double fac = lift_nb * sm[L.mag_nb + f];
#pragma unroll
    for (int c = 0; c < 5; ++c)
      seed_nb[c * mf + f] = gbar[50 * mn.off + comp[c] * mn.np + q] * fac;
  }
  __syncthreads();
  apply_face_matrices_t<5, 1>(na, nb, ka, kb, Pa, Pb, seed_own, sm + L.tmp,
                              sm + L.m_own + 4 * kmf);
  apply_face_matrices_t<5, 1>(ma, mb, ka, kb, Pa_nb, Pb_nb, seed_nb, sm + L.tmp,
                        sm + L.m_own + 9 * kmf);
  for (int k = t; k < kmf; k += T) {
    double li[4], bi[5], bx[5], B[13];
#pragma unroll
    for (int c = 0; c < 4; ++c) li[c] = sm[L.m_own + c * kmf + k];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      bi[c] = sm[L.m_own + (4 + c) * kmf + k];
      bx[c] = sm[L.m_own + (9 + c) * kmf + k];
    }
    gh_package_seed(li, bi, bx, B);
#pragma unroll
    for (int c = 0; c < 13; ++c) sm[L.m_own + c * kmf + k] = B[c];
  }
  __syncthreads();
  apply_face_matrices_t<13, 2>(ka, kb, na, nb, Ia_own, Ib_own, sm + L.m_own, sm + L.tmp, face);
  for (int f = t; f < nf; f += T) {
    const double* o = sm + L.side + 12 * f;
    const int p = face_point(m, d, side, f % na, f / na);
    double B[13], a[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 13; ++c) B[c] = face[c * nf + f];
    gh_package_vjp(o, o + 3, o + 6, o[10], u + 50 * m.off + p, m.np, pr, B,
                   ubar + 50 * m.off + p, m.np, a, a + 3, a + 6);
    double* out = side_adj + 110 * (m.off + p) + 11 * pr;
#pragma unroll
    for (int c = 0; c < 10; ++c) out[c] = a[c];
  }
}

// The adjoints of e's normals, speeds and |n| on the faces normal to
// dimension d, summed over the ten pairs in order, into g.  Grid: (list
// length, 2 sides).
__global__ void __launch_bounds__(256)
gh_mixed_side_adjoint_kernel(int d, const MixedElement* __restrict__ el,
                             const int32_t* __restrict__ list, const double* __restrict__ u,
                             const double* __restrict__ params,
                             const double* __restrict__ inv_jac,
                             const double* __restrict__ side_adj, double* __restrict__ ubar) {
  const int side = blockIdx.y;
  const MixedElement m = el[list[blockIdx.x]];
  const int na = d == 0 ? m.n[1] : m.n[0], nb = d == 2 ? m.n[1] : m.n[2];
  for (int f = threadIdx.x; f < na * nb; f += blockDim.x) {
    GhSideFwd s;
    int p;
    face_side_adjoint(m, d, side, side == 0, f, u, params, inv_jac, s, &p);
    const double* in = side_adj + 110 * (m.off + p);
    double a[11];
#pragma unroll
    for (int c = 0; c < 11; ++c) a[c] = in[c];
    for (int pr = 1; pr < 10; ++pr)
#pragma unroll
      for (int c = 0; c < 11; ++c) a[c] += in[11 * pr + c];
    double gm_b[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    gh_side_vjp(s, params[3 * m.off + m.np + p], a, a + 3, a + 6, a[10], gm_b);
#pragma unroll
    for (int c = 0; c < 10; ++c) ubar[50 * m.off + c * m.np + p] += gm_b[c];
  }
}

// Dynamic LDS beyond the default 64 KiB has to be enabled per kernel; done
// once per device (before any graph capture of these launches).
bool enable_large_lds() {
  static std::mutex mutex;
  static std::vector<bool> done;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return false;
  std::lock_guard<std::mutex> lock(mutex);
  if (static_cast<int>(done.size()) <= dev) done.resize(dev + 1, false);
  if (done[dev]) return true;
  const int lds_max = 160 * 1024;
  if (hipFuncSetAttribute(reinterpret_cast<const void*>(&sw_mixed_volume_kernel<false>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, lds_max) != hipSuccess ||
      hipFuncSetAttribute(reinterpret_cast<const void*>(&sw_mixed_volume_kernel<true>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, lds_max) != hipSuccess ||
      hipFuncSetAttribute(reinterpret_cast<const void*>(&sw_mixed_face_kernel),
                          hipFuncAttributeMaxDynamicSharedMemorySize, lds_max) != hipSuccess ||
      hipFuncSetAttribute(reinterpret_cast<const void*>(&gh_mixed_face_pass_kernel<false>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, lds_max) != hipSuccess ||
      hipFuncSetAttribute(reinterpret_cast<const void*>(&gh_mixed_face_pass_kernel<true>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, lds_max) != hipSuccess ||
      hipFuncSetAttribute(reinterpret_cast<const void*>(&gh_mixed_face_adjoint_pass_kernel<false>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, lds_max) != hipSuccess ||
      hipFuncSetAttribute(reinterpret_cast<const void*>(&gh_mixed_face_adjoint_pass_kernel<true>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, lds_max) != hipSuccess ||
      hipFuncSetAttribute(reinterpret_cast<const void*>(&ragged_transposed_sweeps_kernel),
                          hipFuncAttributeMaxDynamicSharedMemorySize, lds_max) != hipSuccess)
    return false;
  done[dev] = true;
  return true;
}

}  // namespace

int mixed_face_lds_bytes(int kf, int km) { return FacePlan(kf, km, 13).total * 8; }

int launch_mixed_volume(bool gh, int count, const int32_t* list, const MixedElement* el,
                        int np_max, const double* u, const double* params,
                        const double* inv_jac, double* dt, double* scratch_du, int threads,
                        hipStream_t stream) {
  if (count == 0) return SXD_OK;
  const double* dtab = device_diff_matrices();
  if (dtab == nullptr) return SXD_ERR_DEVICE;
  if (threads != 64 && threads != 128 && threads != 256) return SXD_ERR_INVALID_ARGUMENT;
  if (!enable_large_lds()) return SXD_ERR_DEVICE;
  if (!gh) {
    if (5 * np_max * 8 > 160 * 1024) return SXD_ERR_UNSUPPORTED;
    hipLaunchKernelGGL(sw_mixed_volume_kernel<false>, dim3(count), dim3(threads),
                       5 * np_max * 8, stream, el, list, u, params, inv_jac, dt, dtab, nullptr,
                       kSwBcGhostArray, nullptr);
    return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
  }
  if (scratch_du == nullptr) return SXD_ERR_INVALID_ARGUMENT;
  hipLaunchKernelGGL(ragged_partial_derivatives_kernel,
                     dim3(count, (50 + kRaggedChunk - 1) / kRaggedChunk), dim3(threads),
                     (kStagedDiffDoubles + np_max) * 8, stream, el, list, 50, u, inv_jac,
                     scratch_du, dtab);
  if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  hipLaunchKernelGGL(gh_mixed_pointwise_kernel, dim3(count), dim3(threads), 0, stream, el, list,
                     u, scratch_du, params, dt);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

int launch_mixed_rhs(bool gh, int count, const int32_t* list, const MixedElement* el,
                     const int32_t* nbr, int np_max, int kf, int km, const double* u,
                     const double* params, const double* inv_jac, double* dt, double* scratch_du,
                     int threads, hipStream_t stream, const double* ghost, int bc_kind) {
  if (count == 0) return SXD_OK;
  if (gh && bc_kind != kSwBcGhostArray) return SXD_ERR_INVALID_ARGUMENT;
  if (bc_kind < kSwBcGhostArray || bc_kind > kSwBcCpsrSecondOrderBt) return SXD_ERR_UNSUPPORTED;
  const double* dtab = device_diff_matrices();
  const double* mtab = device_mortar_matrices();
  if (dtab == nullptr || mtab == nullptr) return SXD_ERR_DEVICE;
  if (threads != 64 && threads != 128 && threads != 256) return SXD_ERR_INVALID_ARGUMENT;
  const int face_lds = mixed_face_lds_bytes(kf, km);
  if (face_lds > 160 * 1024) return SXD_ERR_UNSUPPORTED;
  if (!enable_large_lds()) return SXD_ERR_DEVICE;
  if (!gh) {
    const int sw_face_lds = SwFacePlan(kf, km).total * 8;
    if (5 * np_max * 8 > 160 * 1024) return SXD_ERR_UNSUPPORTED;
    if (bc_kind >= kSwBcCpsrSommerfeld)
      hipLaunchKernelGGL(sw_mixed_volume_kernel<true>, dim3(count), dim3(threads),
                         5 * np_max * 8, stream, el, list, u, params, inv_jac, dt, dtab, nbr,
                         bc_kind, ghost);
    else
      hipLaunchKernelGGL(sw_mixed_volume_kernel<false>, dim3(count), dim3(threads),
                         5 * np_max * 8, stream, el, list, u, params, inv_jac, dt, dtab, nbr,
                         bc_kind, nullptr);
    if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
    hipLaunchKernelGGL(sw_mixed_face_kernel, dim3(count), dim3(threads), sw_face_lds, stream,
                       el, list, nbr, u, params, inv_jac, dt, mtab, kf, km, ghost, bc_kind);
    return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
  }
  if (scratch_du == nullptr) return SXD_ERR_INVALID_ARGUMENT;
  hipLaunchKernelGGL(ragged_partial_derivatives_kernel,
                     dim3(count, (50 + kRaggedChunk - 1) / kRaggedChunk), dim3(threads),
                     (kStagedDiffDoubles + np_max) * 8, stream, el, list, 50, u, inv_jac,
                     scratch_du, dtab);
  if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  hipLaunchKernelGGL(gh_mixed_pointwise_kernel, dim3(count), dim3(threads), 0, stream, el, list,
                     u, scratch_du, params, dt);
  if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  // with external faces: a sweep over the external faces, then one over the
  // internal ones (the order in which the corrections reach dt)
  if (ghost != nullptr) {
    const int rc = launch_mixed_gh_face_passes(count, list, el, nbr, kf, km, u, params, inv_jac,
                                               dt, threads, ghost, kFacesExternal, stream);
    if (rc != SXD_OK) return rc;
    return launch_mixed_gh_face_passes(count, list, el, nbr, kf, km, u, params, inv_jac, dt,
                                       threads, ghost, kFacesInternal, stream);
  }
  return launch_mixed_gh_face_passes(count, list, el, nbr, kf, km, u, params, inv_jac, dt,
                                     threads, nullptr, kFacesAll, stream);
}

int launch_mixed_gh_face_passes(int count, const int32_t* list, const MixedElement* el,
                                const int32_t* nbr, int kf, int km, const double* u,
                                const double* params, const double* inv_jac, double* dt,
                                int threads, const double* ghost, int phase,
                                hipStream_t stream) {
  if (count == 0) return SXD_OK;
  const double* mtab = device_mortar_matrices();
  if (mtab == nullptr) return SXD_ERR_DEVICE;
  if (threads != 64 && threads != 128 && threads != 256) return SXD_ERR_INVALID_ARGUMENT;
  if (mixed_face_lds_bytes(kf, km) > 160 * 1024) return SXD_ERR_UNSUPPORTED;
  if (!enable_large_lds()) return SXD_ERR_DEVICE;
  // the fused neighbour interpolation only where its smaller LDS plan fits
  // more workgroups on a CU
  const int lds_split = FacePassPlan<false>(kf, km).total * 8;
  const int lds_fused = FacePassPlan<true>(kf, km).total * 8;
  const bool fused_nb = 160 * 1024 / lds_fused > 160 * 1024 / lds_split;
  for (int d = 0; d < 3; ++d) {
    if (fused_nb)
      hipLaunchKernelGGL(gh_mixed_face_pass_kernel<true>, dim3(count, 2, 10), dim3(threads),
                         lds_fused, stream, d, el, list, nbr, u, params, inv_jac, dt, mtab, kf,
                         km, ghost, phase);
    else
      hipLaunchKernelGGL(gh_mixed_face_pass_kernel<false>, dim3(count, 2, 10), dim3(threads),
                         lds_split, stream, d, el, list, nbr, u, params, inv_jac, dt, mtab, kf,
                         km, ghost, phase);
    if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  }
  return SXD_OK;
}

int mixed_face_adjoint_lds_bytes(int kf, int km) {
  return FaceAdjPassPlan<false>(kf, km).total * 8;
}

int launch_mixed_gh_adjoint(int count, const int32_t* list, const MixedElement* el,
                            const int32_t* point_el, int64_t total_points,
                            const int32_t* nbr, int np_max, int kf, int km, const double* u,
                            const double* gammas, const double* inv_jac, const double* gbar,
                            double* ubar, double* scratch_du, hipStream_t stream) {
  if (count == 0) return SXD_OK;
  const double* dtab = device_diff_matrices();
  const double* mtab = device_mortar_matrices();
  if (dtab == nullptr || mtab == nullptr) return SXD_ERR_DEVICE;
  if (mixed_face_adjoint_lds_bytes(kf, km) > 160 * 1024 ||
      (kStagedDiffDoubles + 3 * np_max) * 8 > 160 * 1024)
    return SXD_ERR_UNSUPPORTED;
  if (scratch_du == nullptr) return SXD_ERR_INVALID_ARGUMENT;
  if (!enable_large_lds()) return SXD_ERR_DEVICE;
  const dim3 chunks(count, (50 + kRaggedChunk - 1) / kRaggedChunk);
  hipLaunchKernelGGL(ragged_partial_derivatives_kernel, chunks, dim3(256),
                     (kStagedDiffDoubles + np_max) * 8, stream, el, list, 50, u, inv_jac,
                     scratch_du, dtab);
  if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  hipLaunchKernelGGL(gh_mixed_adjoint_pointwise_kernel,
                     dim3(static_cast<unsigned>((total_points + 127) / 128)), dim3(128), 0, stream,
                     total_points, el, point_el, u, scratch_du, gammas, inv_jac, gbar, ubar);
  if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  hipLaunchKernelGGL(ragged_transposed_sweeps_kernel, chunks, dim3(256),
                     (kStagedDiffDoubles + 3 * np_max) * 8, stream, el, list, 50, scratch_du, ubar,
                     dtab);
  if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  // faces: the derivative work array is free again and holds the side
  // adjoints (110 doubles per point) and the side values of the current
  // dimension's face points (12 per point), under the 150 it has
  const int lds_split = FaceAdjPassPlan<false>(kf, km).total * 8;
  const int lds_fused = FaceAdjPassPlan<true>(kf, km).total * 8;
  const bool fused_nb = 160 * 1024 / lds_fused > 160 * 1024 / lds_split;
  double* sides = scratch_du + 110 * total_points;
  for (int d = 0; d < 3; ++d) {
    hipLaunchKernelGGL(gh_mixed_face_sides_kernel, dim3(count, 2), dim3(128), 0, stream, d, el,
                       list, u, gammas, inv_jac, sides);
    if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
    if (fused_nb)
      hipLaunchKernelGGL(gh_mixed_face_adjoint_pass_kernel<true>, dim3(count, 2, 10), dim3(256),
                         lds_fused, stream, d, el, list, nbr, u, gammas, inv_jac, gbar, ubar,
                         scratch_du, sides, mtab, kf, km);
    else
      hipLaunchKernelGGL(gh_mixed_face_adjoint_pass_kernel<false>, dim3(count, 2, 10),
                         dim3(256), lds_split, stream, d, el, list, nbr, u, gammas, inv_jac,
                         gbar, ubar, scratch_du, sides, mtab, kf, km);
    if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
    hipLaunchKernelGGL(gh_mixed_side_adjoint_kernel, dim3(count, 2), dim3(128), 0, stream, d, el,
                       list, u, gammas, inv_jac, scratch_du, ubar);
    if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  }
  return SXD_OK;
}

}  // namespace sxd
