// Non-conforming mortars (CDNA4, FP64): faces whose neighbour has a different
// orientation of its logical axes, and h-refined (2:1) faces covered by two or
// four mortars, for the ScalarWave and GH DG right-hand sides on ragged
// (mixed-p) element sets.
//
// The reference's data flow (ComputeTimeDerivative.hpp:690-774,
// InternalMortarDataImpl.hpp:230-318, ApplyBoundaryCorrections.hpp:797-1043):
// each element packages its face data (dg_package_data) on its own face mesh,
// projects the package to every mortar of the face with its own mortar sizes
// (projection_matrix_parent_to_child, dimension a then b; identity where the
// size is Full and the extents agree) and sends it to the neighbour oriented
// into the neighbour's point order (orient_variables_on_slice,
// Domain/Structure/OrientationMapHelpers.cpp:372-400).  The receiver forms
// dg_boundary_terms at each mortar point, projects the correction back to
// its face (projection_matrix_child_to_parent) and lifts it (dg::lift_flux),
// one mortar after another.
//
// Here every element's packages stay in HBM and the receiver gathers the
// partner's package through a per-record index table (the partner's
// oriented offsets, inverted on the host), so no data is reordered:
//   1. mortar_package_kernel: one workgroup per face, the package of every
//      face point and |n| (SW 13 values, GH 10 pairs x 13 + 4 speeds);
//      mortar_direct_kernel instead for a face whose single mortar differs
//      from it: each group of package values formed in LDS and projected to
//      the mortar at once (steps 1 and 2 without the face package in HBM);
//   2. mortar_project_kernel: the other records whose mortar differs from
//      the face,
//      one workgroup per (record, component chunk): the dimension-a sums
//      staged in LDS, then one output (mortar point, component) per work item;
//   3. mortar_boundary_pairs_kernel: one workgroup per mortar with both
//      sides local, the corrections of both records at every mortar point
//      from the two packages (each read once); mortar_boundary_terms_kernel
//      for a record whose partner's package arrived in a message;
//   4. mortar_lift_kernel: one launch per dimension (both sides; their points
//      differ), one workgroup per face: each of its mortars in record order
//      is projected back (dimension-a sums staged in LDS, then one work item
//      per (component, face point)) and lifted, so every dt point sees its
//      faces in direction order and a face's mortars in record order.
// Matrix products are sums over the input index, first term a product, then
// fused multiply-adds, dimension a before b, as in the oracle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "device_common.hpp"
#include "gh_point.hpp"
#include "kernels.hpp"
#include "sw_point.hpp"

namespace sxd {
namespace {

__device__ __forceinline__ int face_point_index(const MixedElement& m, int d, int side, int a,
                                                int b) {
  const int n0 = m.n[0], n1 = m.n[1];
  const int nd = d == 0 ? n0 : (d == 1 ? n1 : m.n[2]);
  const int fixed = side == 0 ? 0 : nd - 1;
  if (d == 0) return fixed + n0 * (a + n1 * b);
  if (d == 1) return a + n0 * (fixed + n1 * b);
  return a + n0 * (b + n1 * fixed);
}

// Face -> mortar matrix of one dimension ([n_mortar][n_face]) or nullptr for
// the identity.
__device__ __forceinline__ const double* to_mortar(const double* full, const double* half, int nf,
                                                   int nm, int size) {
  if (size == kChildFull)
    return nf == nm ? nullptr : full + mortar_table_offset(nf, nm);
  return half + (size == kChildUpperHalf ? 0 : kPackedMortarSize) + mortar_table_offset(nf, nm);
}

// Mortar -> face matrix of one dimension ([n_face][n_mortar]) or nullptr.
__device__ __forceinline__ const double* from_mortar(const double* full, const double* half,
                                                     int nf, int nm, int size) {
  if (size == kChildFull)
    return nf == nm ? nullptr : full + mortar_table_offset(nf, nm) + nf * nm;
  return half + (size == kChildUpperHalf ? 0 : kPackedMortarSize) + mortar_table_offset(nf, nm) +
         nf * nm;
}

// Face points of a face packaged and projected in one workgroup (one per
// thread).  The mixed-p LDS plan already keeps faces below this (at most 193
// points), so every face with a single differing mortar qualifies.
constexpr int kDirectPoints = 256;

// One face point's package inputs: GH keeps the side's normal, speeds and
// gamma2 and the point's u (g_ab is read again per pair); ScalarWave keeps the
// normalised normal, u and the parameter.
template <bool GH>
struct PackagePoint;
template <>
struct PackagePoint<true> {
  GhFaceSide s;
  const double* ue;  // u at the point, component stride np
  int np;
};
template <>
struct PackagePoint<false> {
  double n[3], up[5], param;
};

// Package component groups: GH 10 symmetric pairs of 13 values, then the 4
// speeds; ScalarWave one group of 13.
template <bool GH>
constexpr int kPackageGroups = GH ? 11 : 1;
__device__ __forceinline__ int package_group_begin(bool gh, int gr) { return gh ? 13 * gr : 0; }
__device__ __forceinline__ int package_group_size(bool gh, int gr) {
  return gh && gr == 10 ? 4 : 13;
}

// Face point f of face fr (element m): fills pt and returns |n|.
template <bool GH>
__device__ __forceinline__ double package_setup(const MixedElement& m, const MortarFace& fr, int f,
                                                const double* u, const double* params,
                                                const double* inv_jac, PackagePoint<GH>& pt) {
  const int d = fr.dir >> 1, side = fr.dir & 1;
  const int na = d == 0 ? m.n[1] : m.n[0];
  const int np = m.np;
  const int p = face_point_index(m, d, side, f % na, f / na);
  double n[3], mg;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    n[i] = inv_jac[9 * m.off + ij3(d, i) * np + p];
    if (side == 0) n[i] = n[i] * -1.0;
  }
  if constexpr (!GH) {
    flat_normalize(n, &mg);
#pragma unroll
    for (int i = 0; i < 3; ++i) pt.n[i] = n[i];
#pragma unroll
    for (int c = 0; c < 5; ++c) pt.up[c] = u[5 * m.off + c * np + p];
    pt.param = params[m.off + p];
  } else {
    pt.ue = u + 50 * m.off + p;
    pt.np = np;
    double g[10];
#pragma unroll
    for (int c = 0; c < 10; ++c) g[c] = pt.ue[c * np];
    double lapse, shift[3], inv_sp[6];
    gh_lapse_shift(&lapse, shift, inv_sp, g);
    const double* ga = params + 3 * m.off + p;
#pragma unroll
    for (int i = 0; i < 3; ++i) pt.s.n[i] = n[i];
    gh_normalize(pt.s.n, pt.s.nv, &mg, inv_sp);
    gh_char_speeds(pt.s.lam, ga[np], lapse, shift, pt.s.n);
    pt.s.gamma2 = ga[2 * np];
  }
  return mg;
}

// Package group gr of a point (package_group_size values into pk).
template <bool GH>
__device__ __forceinline__ void package_group(const PackagePoint<GH>& pt, int gr, double pk[13]) {
  if constexpr (!GH) {
    sw_package(pk, pt.up, pt.param, pt.n);
  } else if (gr < 10) {
    const int np = pt.np;
    double phi[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) phi[i] = pt.ue[(20 + i + 3 * gr) * np];
    gh_package_pair(pk, pt.ue[gr * np], pt.ue[(10 + gr) * np], phi, pt.s);
  } else {
#pragma unroll
    for (int c = 0; c < 4; ++c) pk[c] = pt.s.lam[c];
  }
}

// Speed c of a GH point (package component 130 + c).
__device__ __forceinline__ double package_speed(const PackagePoint<true>& pt, int c) {
  return pt.s.lam[c];
}
__device__ __forceinline__ double package_speed(const PackagePoint<false>&, int) { return 0.0; }

// Package of every point of the faces in `list` (one workgroup per face) and
// |n| of each point.
template <bool GH>
__global__ void __launch_bounds__(128)
mortar_package_kernel(const MixedElement* __restrict__ el, const MortarFace* __restrict__ faces,
                      const int32_t* __restrict__ list, const double* __restrict__ u,
                      const double* __restrict__ params, const double* __restrict__ inv_jac,
                      double* __restrict__ pkf, double* __restrict__ mag) {
  constexpr int NPK = GH ? 134 : 13;
  const MortarFace fr = faces[list[blockIdx.x]];
  const MixedElement m = el[fr.e];
  const int nf = fr.nf;
  double* out = pkf + static_cast<int64_t>(NPK) * fr.pt_off;
  for (int f = threadIdx.x; f < nf; f += blockDim.x) {
    PackagePoint<GH> pt;
    mag[fr.pt_off + f] = package_setup<GH>(m, fr, f, u, params, inv_jac, pt);
#pragma unroll
    for (int gr = 0; gr < kPackageGroups<GH>; ++gr) {
      double pk[13];
      package_group<GH>(pt, gr, pk);
      const int c0 = package_group_begin(GH, gr), cn = package_group_size(GH, gr);
#pragma unroll
      for (int c = 0; c < 13; ++c)
        if (c < cn) out[(c0 + c) * nf + f] = pk[c];
    }
  }
}

// Components staged at once: at most `cap`, and no more than fit 32 KiB of
// LDS at the largest staged pass (several workgroups per CU).
constexpr int kProjectCap = 16, kLiftCap = 25;
inline int staged_chunk(int nc, int tmp_max, int cap) {
  return std::max(1, std::min({nc, cap, 32768 / (8 * tmp_max)}));
}

// Dimension-a pass of apply_2d for a chunk of cn components, into LDS:
// tmp[c][lb][ra] (oa x ib per component) = sum_l Ma[ra][l] in[c][l + ia lb]
// (in itself where Ma is the identity).  in: component stride `stride`.
__device__ __forceinline__ void stage_pass_a(const double* in, int stride, int cn, int ia, int ib,
                                             int oa, const double* Ma, double* tmp) {
  for (int x = threadIdx.x; x < cn * oa * ib; x += blockDim.x) {
    const int c = x / (oa * ib), y = x % (oa * ib), lb = y / oa, ra = y % oa;
    const double* v = in + c * stride + ia * lb;
    double s;
    if (Ma == nullptr) {
      s = v[ra];
    } else {
      const double* row = Ma + ra * ia;
      s = row[0] * v[0];
      for (int l = 1; l < ia; ++l) s = fma(row[l], v[l], s);
    }
    tmp[x] = s;
  }
}

// Dimension-b pass at output (ra, rb) of component c from the staged tmp.
__device__ __forceinline__ double pass_b(const double* tmp, int c, int oa, int ib,
                                         const double* Mb, int ra, int rb) {
  const double* v = tmp + c * oa * ib + ra;
  if (Mb == nullptr) return v[oa * rb];
  const double* row = Mb + rb * ib;
  double s = row[0] * v[0];
  for (int l = 1; l < ib; ++l) s = fma(row[l], v[oa * l], s);
  return s;
}

// Copies a record's two matrices (na x ka and nb x kb elements, either may
// be the identity: nullptr) into LDS at s_mat and points Ma and Mb there.
// The caller synchronises before the first read.
__device__ __forceinline__ void stage_matrices(const double*& Ma, int size_a, const double*& Mb,
                                               int size_b, double* s_mat) {
  double* s_mb = s_mat + (Ma != nullptr ? size_a : 0);
  if (Ma != nullptr) {
    for (int i = threadIdx.x; i < size_a; i += blockDim.x) s_mat[i] = Ma[i];
    Ma = s_mat;
  }
  if (Mb != nullptr) {
    for (int i = threadIdx.x; i < size_b; i += blockDim.x) s_mb[i] = Mb[i];
    Mb = s_mb;
  }
}

// Face -> mortar projection of the records that need one; grid (record,
// component chunk of `chunk`); dynamic LDS: chunk x tmp_max doubles of
// dimension-a sums, then the record's two matrices (mat_max doubles).
template <int NPK>
__global__ void __launch_bounds__(256)
mortar_project_kernel(const MortarRecord* __restrict__ recs, const int32_t* __restrict__ list,
                      const MortarFace* __restrict__ faces, const double* __restrict__ pkf,
                      double* __restrict__ pkm, const double* __restrict__ full,
                      const double* __restrict__ half, int chunk, int tmp_max) {
  extern __shared__ double s_tmp[];
  const MortarRecord r = recs[list[blockIdx.x]];
  const MortarFace fr = faces[r.face];
  const double* Ma = to_mortar(full, half, r.na, r.ka, r.size_a);
  const double* Mb = to_mortar(full, half, r.nb, r.kb, r.size_b);
  stage_matrices(Ma, r.ka * r.na, Mb, r.kb * r.nb, s_tmp + chunk * tmp_max);
  const int nf = r.na * r.nb, km = r.ka * r.kb;
  const int c0 = blockIdx.y * chunk;
  const int cn = min(chunk, NPK - c0);
  const double* in = pkf + static_cast<int64_t>(NPK) * fr.pt_off + c0 * nf;
  double* out = pkm + static_cast<int64_t>(NPK) * r.mortar_off + c0 * km;
  __syncthreads();  // matrices staged
  stage_pass_a(in, nf, cn, r.na, r.nb, r.ka, Ma, s_tmp);
  __syncthreads();
  for (int x = threadIdx.x; x < cn * km; x += blockDim.x) {
    const int c = x / km, k = x % km;
    out[c * km + k] = pass_b(s_tmp, c, r.ka, r.nb, Mb, k % r.ka, k / r.ka);
  }
}

// Package and projection together for the faces in `list`: faces with a
// single mortar that differs from the face and at most 256 points (one per
// thread).  Each package group is formed in LDS and projected onto the
// mortar at once, so the face package never goes to HBM; the arithmetic is
// mortar_package_kernel's followed by mortar_project_kernel's.  Dynamic LDS:
// 13 x (nf_max + tmp_max) + mat_max doubles (the face-to-mortar matrices).
template <bool GH>
__global__ void __launch_bounds__(kDirectPoints)
mortar_direct_kernel(const MixedElement* __restrict__ el, const MortarFace* __restrict__ faces,
                     const int32_t* __restrict__ list, const MortarRecord* __restrict__ recs,
                     const double* __restrict__ u, const double* __restrict__ params,
                     const double* __restrict__ inv_jac, double* __restrict__ pkm,
                     double* __restrict__ mag, const double* __restrict__ full,
                     const double* __restrict__ half, int nf_max, int tmp_max) {
  constexpr int NPK = GH ? 134 : 13;
  extern __shared__ double s_lds[];
  double* s_pk = s_lds;                 // [13][nf] of the current group
  double* s_tmp = s_lds + 13 * nf_max;  // dimension-a sums
  double* s_ma = s_tmp + 13 * tmp_max;  // the two matrices, read by every group
  const MortarFace fr = faces[list[blockIdx.x]];
  const MortarRecord r = recs[fr.rec_begin];
  const MixedElement m = el[fr.e];
  const int nf = fr.nf, ka = r.ka, km = r.ka * r.kb;
  const int f = threadIdx.x;
  const double* Ma = to_mortar(full, half, r.na, r.ka, r.size_a);
  const double* Mb = to_mortar(full, half, r.nb, r.kb, r.size_b);
  stage_matrices(Ma, ka * r.na, Mb, r.kb * r.nb, s_ma);
  PackagePoint<GH> pt{};
  if (f < nf) mag[fr.pt_off + f] = package_setup<GH>(m, fr, f, u, params, inv_jac, pt);
  double* out = pkm + static_cast<int64_t>(NPK) * r.mortar_off;
  for (int gr = 0; gr < kPackageGroups<GH>; ++gr) {
    const int c0 = package_group_begin(GH, gr), cn = package_group_size(GH, gr);
    if (f < nf) {
      if (GH && gr == 10) {
#pragma unroll
        for (int c = 0; c < 4; ++c) s_pk[c * nf + f] = package_speed(pt, c);
      } else {
        double pk[13];
        package_group<GH>(pt, gr, pk);
#pragma unroll
        for (int c = 0; c < 13; ++c) s_pk[c * nf + f] = pk[c];
      }
    }
    __syncthreads();  // the group's package (and, first, the matrices) complete
    stage_pass_a(s_pk, nf, cn, r.na, r.nb, ka, Ma, s_tmp);
    __syncthreads();  // dimension-a sums complete; s_pk free for the next group
    for (int x = threadIdx.x; x < cn * km; x += blockDim.x) {
      const int c = x / km, k = x % km;
      out[(c0 + c) * km + k] = pass_b(s_tmp, c, ka, r.nb, Mb, k % ka, k / ka);
    }
  }
}

// Package block of a record on its mortar: its own projected block, or the
// face block where the mortar equals the face.
template <int NPK>
__device__ __forceinline__ const double* mortar_package(const MortarRecord& r,
                                                        const MortarFace* faces,
                                                        const double* pkf, const double* pkm) {
  return r.mortar_off >= 0 ? pkm + static_cast<int64_t>(NPK) * r.mortar_off
                           : pkf + static_cast<int64_t>(NPK) * faces[r.face].pt_off;
}

template <bool GH>
__global__ void __launch_bounds__(64)
mortar_boundary_terms_kernel(const MortarRecord* __restrict__ recs,
                             const int32_t* __restrict__ list,
                             const MortarFace* __restrict__ faces,
                             const int32_t* __restrict__ gather, const double* __restrict__ pkf,
                             const double* __restrict__ pkm, const double* __restrict__ recv,
                             double* __restrict__ bcm) {
  constexpr int NPK = GH ? 134 : 13;
  constexpr int NC = GH ? 50 : 5;
  const MortarRecord r = recs[list[blockIdx.x]];
  const bool remote = r.partner < 0;
  const double* own = mortar_package<NPK>(r, faces, pkf, pkm);
  // a partner in another partition: its package arrived in the own point order
  const double* ext = remote ? recv + static_cast<int64_t>(NPK) * r.recv_off
                             : mortar_package<NPK>(recs[r.partner], faces, pkf, pkm);
  const int km = r.ka * r.kb;
  double* out = bcm + static_cast<int64_t>(NC) * r.rec_off;
  for (int k = threadIdx.x; k < km; k += blockDim.x) {
    const int s = remote ? k : gather[r.rec_off + k];
    if constexpr (!GH) {
      double in[13], ex[13], bc[5];
#pragma unroll
      for (int c = 0; c < 13; ++c) {
        in[c] = own[c * km + k];
        ex[c] = ext[c * km + s];
      }
      sw_boundary_terms(bc, in, ex);
#pragma unroll
      for (int c = 0; c < 5; ++c) out[c * km + k] = bc[c];
    } else {
      double lam_in[4], lam_ex[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        lam_in[c] = own[(130 + c) * km + k];
        lam_ex[c] = ext[(130 + c) * km + s];
      }
      for (int pr = 0; pr < 10; ++pr) {
        double in[13], ex[13], bc[5];
#pragma unroll
        for (int c = 0; c < 13; ++c) {
          in[c] = own[(13 * pr + c) * km + k];
          ex[c] = ext[(13 * pr + c) * km + s];
        }
        gh_boundary_terms_pair(bc, in, lam_in, ex, lam_ex);
        out[pr * km + k] = bc[0];
        out[(10 + pr) * km + k] = bc[1];
#pragma unroll
        for (int i = 0; i < 3; ++i) out[(20 + 3 * pr + i) * km + k] = bc[2 + i];
      }
    }
  }
}

// Boundary terms of both records of a mortar whose sides are both local, one
// workgroup per mortar (the record in `list` and its partner): each package
// is read once, and each side gets the correction mortar_boundary_terms_kernel
// forms for it (its own package interior, the other exterior), at its own
// point order (own point k is the partner's point gather[k], and the
// partner's gather maps it back).
template <bool GH>
__global__ void __launch_bounds__(64)
mortar_boundary_pairs_kernel(const MortarRecord* __restrict__ recs,
                             const int32_t* __restrict__ list,
                             const MortarFace* __restrict__ faces,
                             const int32_t* __restrict__ gather, const double* __restrict__ pkf,
                             const double* __restrict__ pkm, double* __restrict__ bcm) {
  constexpr int NPK = GH ? 134 : 13;
  constexpr int NC = GH ? 50 : 5;
  const MortarRecord r = recs[list[blockIdx.x]];
  const MortarRecord q = recs[r.partner];
  const double* own = mortar_package<NPK>(r, faces, pkf, pkm);
  const double* ext = mortar_package<NPK>(q, faces, pkf, pkm);
  const int km = r.ka * r.kb;
  double* out = bcm + static_cast<int64_t>(NC) * r.rec_off;
  double* out_q = bcm + static_cast<int64_t>(NC) * q.rec_off;
  for (int k = threadIdx.x; k < km; k += blockDim.x) {
    const int s = gather[r.rec_off + k];
    if constexpr (!GH) {
      double in[13], ex[13], bc[5];
#pragma unroll
      for (int c = 0; c < 13; ++c) {
        in[c] = own[c * km + k];
        ex[c] = ext[c * km + s];
      }
      sw_boundary_terms(bc, in, ex);
#pragma unroll
      for (int c = 0; c < 5; ++c) out[c * km + k] = bc[c];
      sw_boundary_terms(bc, ex, in);
#pragma unroll
      for (int c = 0; c < 5; ++c) out_q[c * km + s] = bc[c];
    } else {
      double lam_in[4], lam_ex[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        lam_in[c] = own[(130 + c) * km + k];
        lam_ex[c] = ext[(130 + c) * km + s];
      }
      for (int pr = 0; pr < 10; ++pr) {
        double in[13], ex[13], bc[5];
#pragma unroll
        for (int c = 0; c < 13; ++c) {
          in[c] = own[(13 * pr + c) * km + k];
          ex[c] = ext[(13 * pr + c) * km + s];
        }
        gh_boundary_terms_pair(bc, in, lam_in, ex, lam_ex);
        out[pr * km + k] = bc[0];
        out[(10 + pr) * km + k] = bc[1];
#pragma unroll
        for (int i = 0; i < 3; ++i) out[(20 + 3 * pr + i) * km + k] = bc[2 + i];
        gh_boundary_terms_pair(bc, ex, lam_ex, in, lam_in);
        out_q[pr * km + s] = bc[0];
        out_q[(10 + pr) * km + s] = bc[1];
#pragma unroll
        for (int i = 0; i < 3; ++i) out_q[(20 + 3 * pr + i) * km + s] = bc[2 + i];
      }
    }
  }
}

// Mortar -> face projection and lift of the faces [face_begin, ...) of one
// dimension, one workgroup per face: its mortars in record order, each
// projected (dimension a staged in LDS for `chunk` components at a time,
// then dimension b per (component, face point) work item) and added into
// dt.  Dynamic LDS: chunk x tmp_max doubles of dimension-a sums, then the
// record's two matrices (mat_max doubles).
template <bool GH>
__global__ void __launch_bounds__(256)
mortar_lift_kernel(const MixedElement* __restrict__ el, const MortarFace* __restrict__ faces,
                   int face_begin, const MortarRecord* __restrict__ recs,
                   const double* __restrict__ mag, const double* __restrict__ bcm,
                   double* __restrict__ dt, const double* __restrict__ full,
                   const double* __restrict__ half, int chunk, int tmp_max) {
  constexpr int NC = GH ? 50 : 5;
  extern __shared__ double s_tmp[];
  const MortarFace fr = faces[face_begin + blockIdx.x];
  const MixedElement m = el[fr.e];
  const int d = fr.dir >> 1, side = fr.dir & 1;
  const int na = d == 0 ? m.n[1] : m.n[0];
  const int nd = d == 0 ? m.n[0] : (d == 1 ? m.n[1] : m.n[2]);
  const int nf = fr.nf;
  double* target0 = dt + NC * m.off;
  for (int q = fr.rec_begin; q < fr.rec_begin + fr.rec_count; ++q) {
    const MortarRecord r = recs[q];
    const double* Pa = from_mortar(full, half, r.na, r.ka, r.size_a);
    const double* Pb = from_mortar(full, half, r.nb, r.kb, r.size_b);
    const int km = r.ka * r.kb;
    const double* src = bcm + static_cast<int64_t>(NC) * r.rec_off;
    if (Pa == nullptr && Pb == nullptr) {
      // the mortar is the face
      for (int x = threadIdx.x; x < NC * nf; x += blockDim.x) {
        const int c = x / nf, f = x % nf;
        const int p = face_point_index(m, d, side, f % na, f / na);
        double& t = target0[c * m.np + p];
        t = t + src[c * km + f] * lift_factor(nd, mag[fr.pt_off + f]);
      }
    } else {
      // the previous mortar's last barrier ended its reads of the matrices
      stage_matrices(Pa, r.na * r.ka, Pb, r.nb * r.kb, s_tmp + chunk * tmp_max);
      __syncthreads();
      for (int c0 = 0; c0 < NC; c0 += chunk) {
        const int cn = min(chunk, NC - c0);
        stage_pass_a(src + c0 * km, km, cn, r.ka, r.kb, r.na, Pa, s_tmp);
        __syncthreads();
        for (int x = threadIdx.x; x < cn * nf; x += blockDim.x) {
          const int c = x / nf, f = x % nf;
          const int fa = f % na, fb = f / na;
          const int p = face_point_index(m, d, side, fa, fb);
          double& t = target0[(c0 + c) * m.np + p];
          t = t + pass_b(s_tmp, c, r.na, r.kb, Pb, fa, fb) * lift_factor(nd, mag[fr.pt_off + f]);
        }
        __syncthreads();  // s_tmp reused; this mortar's dt updates precede the next one's
      }
      continue;
    }
    __syncthreads();  // this mortar's dt updates precede the next one's
  }
}

// Outgoing boundary messages, one workgroup each, into the send buffer's
// group of the receiving rank: the group's directory entry (and, for the
// first message of a group, the message count), the header, then the
// record's mortar package oriented into the receiver's point order (the
// sender orients, ComputeTimeDerivative.hpp:710-721): point s of the own
// mortar goes to the receiver's point gather[s].
template <int NPK>
__global__ void __launch_bounds__(64)
mortar_pack_kernel(const MortarMessage* __restrict__ msgs, const MortarRecord* __restrict__ recs,
                   const MortarFace* __restrict__ faces, const int32_t* __restrict__ gather,
                   const double* __restrict__ pkf, const double* __restrict__ pkm,
                   unsigned char* __restrict__ send, int system, int32_t step_lo,
                   int32_t step_hi) {
  const MortarMessage msg = msgs[blockIdx.x];
  unsigned char* group = send + msg.group_offset[system];
  unsigned char* base = group + msg.offset[system];
  int32_t* hd = reinterpret_cast<int32_t*>(base);
  const int km = msg.km;
  if (threadIdx.x < 16) {
    int32_t v = msg.header[threadIdx.x];
    if (threadIdx.x == 1) v = system;
    if (threadIdx.x == 11) v = NPK;
    if (threadIdx.x == 12) v = NPK * km;
    if (threadIdx.x == 14) v = step_lo;
    if (threadIdx.x == 15) v = step_hi;
    hd[threadIdx.x] = v;
  } else if (threadIdx.x == 16) {
    int64_t* dir = reinterpret_cast<int64_t*>(group);
    dir[1 + msg.index] = msg.offset[system];
    if (msg.index == 0) dir[0] = msg.group_count;
  }
  const MortarRecord r = recs[msg.record];
  const double* own = mortar_package<NPK>(r, faces, pkf, pkm);
  double* data = reinterpret_cast<double*>(base + kMessageHeaderBytes);
  for (int x = threadIdx.x; x < NPK * km; x += blockDim.x) {
    const int c = x / km, k = x % km;
    data[c * km + gather[r.rec_off + k]] = own[x];
  }
}

// The messages of one sending rank's group, one workgroup each (grid = the
// number of messages this rank expects from it): find the receiving slot by
// (receiver element, receiver direction, sender element), check the header
// and copy the package.  status[2 i] = 0 or the failed check, status[2 i + 1]
// = the slot.
template <int NPK>
__global__ void __launch_bounds__(64)
mortar_unpack_kernel(const unsigned char* __restrict__ group, int64_t bytes,
                     const MortarRecvKey* __restrict__ keys, int num_keys,
                     double* __restrict__ recv, int system, int32_t step_lo, int32_t step_hi,
                     int32_t* __restrict__ status) {
  __shared__ int64_t s_off, s_dst;
  __shared__ int32_t s_n, s_code;
  if (threadIdx.x == 0) {
    const int64_t* dir = reinterpret_cast<const int64_t*>(group);
    int code = 0, slot = -1, n = 0;
    int64_t off = 0, dst = 0;
    if (directory_bytes(gridDim.x) > bytes || dir[0] != static_cast<int64_t>(gridDim.x)) {
      code = kMessageWrongCount;
    } else {
      off = dir[1 + blockIdx.x];
      if (off < directory_bytes(gridDim.x) || off % 8 != 0 || off + kMessageHeaderBytes > bytes)
        code = kMessageTruncated;
    }
    const int32_t* hd = reinterpret_cast<const int32_t*>(group + off);
    if (code == 0) {
      if (hd[0] != kMessageMagic) code = kMessageBadMagic;
      else if (hd[1] != system || hd[11] != NPK) code = kMessageWrongSystem;
      else if (hd[14] != step_lo || hd[15] != step_hi) code = kMessageWrongStep;
      else if (hd[12] < 0 || off + kMessageHeaderBytes + 8 * static_cast<int64_t>(hd[12]) > bytes)
        code = kMessageTruncated;
    }
    if (code == 0) {
      int lo = 0, hi = num_keys - 1, found = -1;
      while (lo <= hi) {
        const int mid = (lo + hi) / 2;
        const MortarRecvKey& k = keys[mid];
        const int c = k.recv_e != hd[3] ? (k.recv_e < hd[3] ? -1 : 1)
                      : k.recv_dir != hd[4] ? (k.recv_dir < hd[4] ? -1 : 1)
                      : k.send_e != hd[2] ? (k.send_e < hd[2] ? -1 : 1)
                                          : 0;
        if (c == 0) {
          found = mid;
          break;
        }
        if (c < 0) lo = mid + 1;
        else hi = mid - 1;
      }
      if (found < 0) {
        code = kMessageUnknownMortar;
      } else {
        const MortarRecvKey& k = keys[found];
        if (hd[9] != k.ka || hd[10] != k.kb || hd[12] != NPK * k.ka * k.kb) {
          code = kMessageWrongMesh;
        } else {
          slot = k.slot;
          dst = k.recv_off;
          n = hd[12];
        }
      }
    }
    s_off = off;
    s_dst = dst;
    s_n = n;
    s_code = code;
    status[2 * blockIdx.x] = code;
    status[2 * blockIdx.x + 1] = slot;
  }
  __syncthreads();
  if (s_code != 0) return;
  const double* data = reinterpret_cast<const double*>(group + s_off + kMessageHeaderBytes);
  double* out = recv + NPK * s_dst;
  for (int x = threadIdx.x; x < s_n; x += blockDim.x) out[x] = data[x];
}

// ---- host: the record list ------------------------------------------------------

int map_direction(const int o[3], int dir) {
  const int m = o[dir / 2];
  return (dir % 2 == 1) ? m : (m ^ 1);
}

bool valid_orientation(const int o[3]) {
  bool seen[3] = {false, false, false};
  for (int i = 0; i < 3; ++i) {
    if (o[i] < 0 || o[i] > 5 || seen[o[i] / 2]) return false;
    seen[o[i] / 2] = true;
  }
  return true;
}

// oriented_offset_on_slice (OrientationMapHelpers.cpp:219-241 with
// compute_offset_permutation :42-78): slice point i1 + n1 i2 of this side
// lies at the returned index of the neighbour's slice.
std::vector<int32_t> oriented_offsets(int n1, int n2, int sliced_dim, const int o[3]) {
  const int dims[2] = {sliced_dim == 0 ? 1 : 0, sliced_dim == 2 ? 1 : 2};
  const bool transposed = o[dims[0]] / 2 > o[dims[1]] / 2;
  const bool a1 = o[dims[0]] % 2 == 1, a2 = o[dims[1]] % 2 == 1;
  std::vector<int32_t> out(static_cast<size_t>(n1) * n2);
  for (int i2 = 0; i2 < n2; ++i2)
    for (int i1 = 0; i1 < n1; ++i1) {
      const int j1 = a1 ? i1 : n1 - 1 - i1, j2 = a2 ? i2 : n2 - 1 - i2;
      out[i1 + n1 * i2] = transposed ? j2 + n2 * j1 : j1 + n1 * j2;
    }
  return out;
}

}  // namespace

int build_mortar_plan(int num_elements, const int32_t* extents, int num_records,
                      const int32_t* records, const int32_t* partition, int rank,
                      MortarPlan* plan, std::string* error) {
  *plan = MortarPlan{};
  auto bad = [&](const std::string& msg) {
    *error = msg;
    return SXD_ERR_INVALID_ARGUMENT;
  };
  if (num_records < 1) return bad("need at least one mortar record");
  struct Host {
    int e, dir, en, o[3], size[2], caller, partner = -1, na = 0, nb = 0, ka = 0, kb = 0;
  };
  std::vector<Host> h(num_records);
  std::map<std::array<int, 3>, int> index;
  for (int i = 0; i < num_records; ++i) {
    const int32_t* q = records + 8 * i;
    Host& r = h[i];
    r.e = q[0];
    r.dir = q[1];
    r.en = q[2];
    for (int k = 0; k < 3; ++k) r.o[k] = q[3 + k];
    r.size[0] = q[6];
    r.size[1] = q[7];
    r.caller = i;
    if (r.e < 0 || r.e >= num_elements || r.en < 0 || r.en >= num_elements)
      return bad("record " + std::to_string(i) + ": element index out of range");
    if (r.dir < 0 || r.dir > 5) return bad("record " + std::to_string(i) + ": direction not 0..5");
    if (!valid_orientation(r.o))
      return bad("record " + std::to_string(i) + ": orientation is not a signed permutation");
    for (int k = 0; k < 2; ++k)
      if (r.size[k] < kChildFull || r.size[k] > kChildLowerHalf)
        return bad("record " + std::to_string(i) + ": mortar size not 1..3");
    if (!index.emplace(std::array<int, 3>{r.e, r.dir, r.en}, i).second)
      return bad("record " + std::to_string(i) + ": duplicate (element, direction, neighbour)");
    const int32_t* ext = extents + 3 * r.e;
    const int d = r.dir / 2;
    r.na = d == 0 ? ext[1] : ext[0];
    r.nb = d == 2 ? ext[1] : ext[2];
  }
  for (int i = 0; i < num_records; ++i) {
    Host& r = h[i];
    const int dir_n = map_direction(r.o, r.dir ^ 1);
    const auto it = index.find({r.en, dir_n, r.e});
    if (it == index.end())
      return bad("record " + std::to_string(i) + ": the neighbour has no record back");
    r.partner = it->second;
    const Host& p = h[r.partner];
    for (int k = 0; k < 3; ++k) {
      const int j = r.o[k] / 2;
      if (p.o[j] != 2 * k + r.o[k] % 2)
        return bad("record " + std::to_string(i) + ": orientations of the two sides not inverse");
    }
    const int d = r.dir / 2, dn = dir_n / 2;
    const int dims[2] = {d == 0 ? 1 : 0, d == 2 ? 1 : 2};
    const int pdims[2] = {dn == 0 ? 1 : 0, dn == 2 ? 1 : 2};
    int pn[2];
    for (int k = 0; k < 2; ++k) {
      const int slot = r.o[dims[k]] / 2 == pdims[0] ? 0 : 1;
      pn[k] = slot == 0 ? p.na : p.nb;
      if (r.size[k] != kChildFull && p.size[slot] != kChildFull)
        return bad("record " + std::to_string(i) + ": a mortar dimension halved on both sides");
    }
    r.ka = std::max(r.na, pn[0]);
    r.kb = std::max(r.nb, pn[1]);
  }
  for (int i = 0; i < num_records; ++i) {
    const Host& r = h[i];
    const Host& p = h[r.partner];
    const int d = r.dir / 2;
    const bool transposed = r.o[d == 0 ? 1 : 0] / 2 > r.o[d == 2 ? 1 : 2] / 2;
    if ((transposed ? p.kb : p.ka) != r.ka || (transposed ? p.ka : p.kb) != r.kb)
      return bad("record " + std::to_string(i) + ": the two sides' mortar meshes disagree");
  }
  // the partition: owned elements in global order get local indices
  std::vector<int32_t> local(num_elements, -1);
  int num_owned = 0;
  for (int e = 0; e < num_elements; ++e)
    if (partition == nullptr || partition[e] == rank) local[e] = num_owned++;
  if (num_owned == 0) return bad("rank " + std::to_string(rank) + " owns no element");
  // faces: the owned elements' records grouped by (dimension, element,
  // side), caller order within
  std::vector<int> order;
  for (int i = 0; i < num_records; ++i)
    if (local[h[i].e] >= 0) order.push_back(i);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    const auto kx = std::make_tuple(h[x].dir / 2, h[x].e, h[x].dir % 2);
    const auto ky = std::make_tuple(h[y].dir / 2, h[y].e, h[y].dir % 2);
    return kx < ky;
  });
  const int kept = static_cast<int>(order.size());
  std::vector<int> pos(num_records, -1);
  for (int i = 0; i < kept; ++i) pos[order[i]] = i;
  std::vector<MortarFace> faces;
  std::vector<MortarRecord> recs(kept);
  std::vector<int32_t> gather, proj;
  std::vector<MortarMessage> messages;
  std::vector<MortarRecvKey> keys;
  int64_t face_pts = 0, proj_pts = 0, rec_pts = 0, recv_pts = 0;
  int64_t msg_bytes[2] = {0, 0};
  int dim_begin[4] = {0, 0, 0, 0};
  for (int i = 0; i < kept; ++i) {
    const Host& r = h[order[i]];
    const int le = local[r.e];
    if (faces.empty() || faces.back().e != le || faces.back().dir != r.dir) {
      MortarFace f{};
      f.e = le;
      f.dir = r.dir;
      f.nf = r.na * r.nb;
      f.rec_begin = i;
      f.rec_count = 0;
      f.pt_off = face_pts;
      face_pts += f.nf;
      faces.push_back(f);
      for (int d = r.dir / 2 + 1; d <= 3; ++d) dim_begin[d] = static_cast<int>(faces.size());
    }
    MortarFace& f = faces.back();
    ++f.rec_count;
    MortarRecord& m = recs[i];
    m.face = static_cast<int32_t>(faces.size() - 1);
    m.na = r.na;
    m.nb = r.nb;
    m.ka = r.ka;
    m.kb = r.kb;
    m.size_a = r.size[0];
    m.size_b = r.size[1];
    const int64_t km = static_cast<int64_t>(r.ka) * r.kb;
    const bool identity = r.size[0] == kChildFull && r.size[1] == kChildFull && r.na == r.ka &&
                          r.nb == r.kb;
    m.mortar_off = identity ? -1 : proj_pts;
    if (!identity) {
      proj_pts += km;
      proj.push_back(i);
    }
    m.rec_off = rec_pts;
    rec_pts += km;
    // my mortar point k reads the partner's point offsets[k]: the partner's
    // oriented offsets map its points to mine, mine map mine to its
    const std::vector<int32_t> g = oriented_offsets(r.ka, r.kb, r.dir / 2, r.o);
    gather.insert(gather.end(), g.begin(), g.end());
    const Host& p = h[r.partner];
    if (local[p.e] >= 0) {
      m.partner = pos[r.partner];
      m.recv_off = -1;
      continue;
    }
    // the partner lives in another partition: its package arrives as a
    // message (slot = the order of such records), and this side's package
    // leaves as one, oriented into the partner's point order
    const int slot = static_cast<int>(keys.size());
    m.partner = -1 - slot;
    m.recv_off = recv_pts;
    recv_pts += km;
    MortarRecvKey key{};
    key.recv_e = r.e;
    key.recv_dir = r.dir;
    key.send_e = p.e;
    key.slot = slot;
    key.ka = r.ka;
    key.kb = r.kb;
    key.recv_off = m.recv_off;
    key.send_rank = partition == nullptr ? 0 : partition[p.e];
    keys.push_back(key);
    MortarMessage msg{};
    msg.record = i;
    msg.km = static_cast<int32_t>(km);
    msg.peer = key.send_rank;  // the partner's rank receives this side's package
    int32_t* hd = msg.header;
    hd[0] = kMessageMagic;
    hd[2] = r.e;
    hd[3] = p.e;
    hd[4] = p.dir;
    hd[5] = r.dir;
    for (int k = 0; k < 3; ++k) hd[6 + k] = extents[3 * r.e + k];
    hd[9] = p.ka;
    hd[10] = p.kb;
    messages.push_back(msg);
  }
  // outgoing messages grouped by receiving rank (ascending), record order
  // within a group; each group starts with its directory
  std::stable_sort(messages.begin(), messages.end(),
                   [](const MortarMessage& x, const MortarMessage& y) { return x.peer < y.peer; });
  for (size_t i = 0; i < messages.size();) {
    size_t j = i;
    while (j < messages.size() && messages[j].peer == messages[i].peer) ++j;
    MortarPeer peer{};
    peer.rank = messages[i].peer;
    peer.count = static_cast<int>(j - i);
    for (int sys = 0; sys < 2; ++sys) {
      peer.offset[sys] = msg_bytes[sys];
      int64_t off = directory_bytes(peer.count);
      for (size_t q = i; q < j; ++q) {
        messages[q].group_offset[sys] = peer.offset[sys];
        messages[q].offset[sys] = off;
        off += kMessageHeaderBytes + message_data_bytes(sys == 1, messages[q].km);
      }
      peer.bytes[sys] = off;
      msg_bytes[sys] += off;
    }
    for (size_t q = i; q < j; ++q) {
      messages[q].index = static_cast<int32_t>(q - i);
      messages[q].group_count = peer.count;
    }
    plan->peers.push_back(peer);
    i = j;
  }
  // incoming: the number of messages expected from each sending rank
  for (const MortarRecvKey& k : keys) ++plan->expected[k.send_rank];
  std::sort(keys.begin(), keys.end(), [](const MortarRecvKey& x, const MortarRecvKey& y) {
    return std::make_tuple(x.recv_e, x.recv_dir, x.send_e) <
           std::make_tuple(y.recv_e, y.recv_dir, y.send_e);
  });
  // every face covered exactly once by its mortars
  for (const MortarFace& f : faces) {
    int quarters = 0;
    for (int q = f.rec_begin; q < f.rec_begin + f.rec_count; ++q)
      quarters += (recs[q].size_a == kChildFull ? 2 : 1) * (recs[q].size_b == kChildFull ? 2 : 1);
    if (quarters != 4)
      return bad("face " + std::to_string(f.dir) + " of element " + std::to_string(f.e) +
                 ": its mortars do not cover it exactly once");
  }
  std::vector<bool> has_face(6 * static_cast<size_t>(num_owned), false);
  for (const MortarFace& f : faces) has_face[6 * static_cast<size_t>(f.e) + f.dir] = true;
  for (size_t i = 0; i < has_face.size(); ++i)
    if (!has_face[i])
      return bad("element " + std::to_string(i / 6) + " has no mortar in direction " +
                 std::to_string(i % 6));
  // faces packaged and projected together: one mortar, which differs from
  // the face; the project kernel takes the other projected mortars
  std::vector<int32_t> package_faces, direct, projected;
  int direct_nf = 0, direct_tmp = 1, direct_mat = 0;
  for (size_t i = 0; i < faces.size(); ++i) {
    const MortarFace& f = faces[i];
    const MortarRecord& r = recs[f.rec_begin];
    if (f.rec_count == 1 && r.mortar_off >= 0 && f.nf <= kDirectPoints) {
      direct.push_back(static_cast<int32_t>(i));
      direct_nf = std::max(direct_nf, f.nf);
      direct_tmp = std::max(direct_tmp, r.ka * r.nb);
      direct_mat = std::max(direct_mat, r.ka * r.na + r.kb * r.nb);
    } else {
      package_faces.push_back(static_cast<int32_t>(i));
    }
  }
  for (const int32_t q : proj) {
    const MortarFace& f = faces[recs[q].face];
    if (!(f.rec_count == 1 && f.nf <= kDirectPoints)) projected.push_back(q);
  }
  std::vector<int32_t> pairs, remote_records;
  for (int i = 0; i < kept; ++i) {
    if (recs[i].partner < 0)
      remote_records.push_back(i);
    else if (recs[i].partner > i)
      pairs.push_back(i);
  }
  int kf = 0, km = 0, tmp_max = 1, mat_max = 0;
  for (const MortarRecord& m : recs) {
    kf = std::max(kf, m.na * m.nb);
    km = std::max(km, m.ka * m.kb);
    tmp_max = std::max({tmp_max, m.ka * m.nb, m.na * m.kb});
    mat_max = std::max(mat_max, m.ka * m.na + m.kb * m.nb);
  }
  plan->num_faces = static_cast<int>(faces.size());
  plan->num_records = kept;
  plan->num_remote = static_cast<int>(keys.size());
  plan->remote_points = recv_pts;
  plan->message_bytes[0] = msg_bytes[0];
  plan->message_bytes[1] = msg_bytes[1];
  plan->num_projected = static_cast<int>(projected.size());
  plan->num_package_faces = static_cast<int>(package_faces.size());
  plan->num_direct = static_cast<int>(direct.size());
  plan->num_pairs = static_cast<int>(pairs.size());
  plan->num_remote_records = static_cast<int>(remote_records.size());
  plan->direct_nf = direct_nf;
  plan->direct_tmp = direct_tmp;
  plan->direct_mat = direct_mat;
  plan->face_points = face_pts;
  plan->projected_points = proj_pts;
  plan->record_points = rec_pts;
  for (int d = 0; d < 4; ++d) plan->dim_begin[d] = dim_begin[d];
  plan->kf = kf;
  plan->km = km;
  plan->tmp_max = tmp_max;
  plan->mat_max = mat_max;
  int rc = SXD_OK;
  auto upload = [&](auto** dst, const auto& v) {
    if (rc != SXD_OK || v.empty()) return;
    if (hipMalloc(dst, v.size() * sizeof(v[0])) != hipSuccess) {
      rc = SXD_ERR_OUT_OF_MEMORY;
      return;
    }
    if (hipMemcpy(*dst, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice) != hipSuccess)
      rc = SXD_ERR_DEVICE;
  };
  upload(&plan->d_faces, faces);
  upload(&plan->d_records, recs);
  upload(&plan->d_gather, gather);
  upload(&plan->d_projected, projected);
  upload(&plan->d_package_faces, package_faces);
  upload(&plan->d_direct, direct);
  upload(&plan->d_pairs, pairs);
  upload(&plan->d_remote_records, remote_records);
  upload(&plan->d_messages, messages);
  upload(&plan->d_keys, keys);
  if (rc != SXD_OK) {
    destroy_mortar_plan(plan);
    *error = "device set-up of the mortar tables failed";
  }
  return rc;
}

void destroy_mortar_plan(MortarPlan* plan) {
  if (plan->d_faces != nullptr) (void)hipFree(plan->d_faces);
  if (plan->d_records != nullptr) (void)hipFree(plan->d_records);
  if (plan->d_gather != nullptr) (void)hipFree(plan->d_gather);
  if (plan->d_projected != nullptr) (void)hipFree(plan->d_projected);
  if (plan->d_package_faces != nullptr) (void)hipFree(plan->d_package_faces);
  if (plan->d_direct != nullptr) (void)hipFree(plan->d_direct);
  if (plan->d_pairs != nullptr) (void)hipFree(plan->d_pairs);
  if (plan->d_remote_records != nullptr) (void)hipFree(plan->d_remote_records);
  if (plan->d_messages != nullptr) (void)hipFree(plan->d_messages);
  if (plan->d_keys != nullptr) (void)hipFree(plan->d_keys);
  *plan = MortarPlan{};
}

size_t mortar_work_doubles(bool gh, const MortarPlan& plan) {
  const size_t npk = gh ? 134 : 13, nc = gh ? 50 : 5;
  return npk * static_cast<size_t>(plan.face_points + plan.projected_points +
                                   plan.remote_points) +
         static_cast<size_t>(plan.face_points) + nc * static_cast<size_t>(plan.record_points);
}

namespace {
struct WorkLayout {
  double *pkf, *pkm, *mag, *bcm, *recv;
  WorkLayout(bool gh, const MortarPlan& plan, double* work) {
    const size_t npk = gh ? 134 : 13, nc = gh ? 50 : 5;
    pkf = work;
    pkm = pkf + npk * plan.face_points;
    mag = pkm + npk * plan.projected_points;
    bcm = mag + plan.face_points;
    recv = bcm + nc * plan.record_points;
  }
};
}  // namespace

int launch_mortar_packages(bool gh, const MortarPlan& plan, const MixedElement* el,
                           const double* u, const double* params, const double* inv_jac,
                           double* work, hipStream_t stream) {
  const double* full = device_mortar_matrices();
  const double* half = device_half_mortar_matrices();
  if (full == nullptr || half == nullptr) return SXD_ERR_DEVICE;
  if (work == nullptr) return SXD_ERR_INVALID_ARGUMENT;
  const int npk = gh ? 134 : 13;
  const WorkLayout w(gh, plan, work);
  if (plan.num_package_faces > 0) {
    if (gh)
      hipLaunchKernelGGL(mortar_package_kernel<true>, dim3(plan.num_package_faces), dim3(128), 0,
                         stream, el, plan.d_faces, plan.d_package_faces, u, params, inv_jac,
                         w.pkf, w.mag);
    else
      hipLaunchKernelGGL(mortar_package_kernel<false>, dim3(plan.num_package_faces), dim3(128), 0,
                         stream, el, plan.d_faces, plan.d_package_faces, u, params, inv_jac,
                         w.pkf, w.mag);
    if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  }
  if (plan.num_direct > 0) {
    const size_t lds =
        sizeof(double) * (13 * (plan.direct_nf + plan.direct_tmp) + plan.direct_mat);
    if (gh)
      hipLaunchKernelGGL(mortar_direct_kernel<true>, dim3(plan.num_direct), dim3(kDirectPoints),
                         lds, stream, el, plan.d_faces, plan.d_direct, plan.d_records, u, params,
                         inv_jac, w.pkm, w.mag, full, half, plan.direct_nf, plan.direct_tmp);
    else
      hipLaunchKernelGGL(mortar_direct_kernel<false>, dim3(plan.num_direct), dim3(kDirectPoints),
                         lds, stream, el, plan.d_faces, plan.d_direct, plan.d_records, u, params,
                         inv_jac, w.pkm, w.mag, full, half, plan.direct_nf, plan.direct_tmp);
    if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  }
  if (plan.num_projected > 0) {
    const int chunk = staged_chunk(npk, plan.tmp_max, kProjectCap);
    const dim3 grid(plan.num_projected, (npk + chunk - 1) / chunk);
    const size_t lds = sizeof(double) * (chunk * plan.tmp_max + plan.mat_max);
    if (gh)
      hipLaunchKernelGGL(mortar_project_kernel<134>, grid, dim3(256), lds, stream, plan.d_records,
                         plan.d_projected, plan.d_faces, w.pkf, w.pkm, full, half, chunk,
                         plan.tmp_max);
    else
      hipLaunchKernelGGL(mortar_project_kernel<13>, grid, dim3(256), lds, stream, plan.d_records,
                         plan.d_projected, plan.d_faces, w.pkf, w.pkm, full, half, chunk,
                         plan.tmp_max);
    if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  }
  return SXD_OK;
}

int launch_mortar_boundary_terms(bool gh, const MortarPlan& plan, double* work,
                                 hipStream_t stream) {
  if (work == nullptr) return SXD_ERR_INVALID_ARGUMENT;
  const WorkLayout w(gh, plan, work);
  if (plan.num_pairs > 0) {
    if (gh)
      hipLaunchKernelGGL(mortar_boundary_pairs_kernel<true>, dim3(plan.num_pairs), dim3(64), 0,
                         stream, plan.d_records, plan.d_pairs, plan.d_faces, plan.d_gather, w.pkf,
                         w.pkm, w.bcm);
    else
      hipLaunchKernelGGL(mortar_boundary_pairs_kernel<false>, dim3(plan.num_pairs), dim3(64), 0,
                         stream, plan.d_records, plan.d_pairs, plan.d_faces, plan.d_gather, w.pkf,
                         w.pkm, w.bcm);
    if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  }
  if (plan.num_remote_records > 0) {
    if (gh)
      hipLaunchKernelGGL(mortar_boundary_terms_kernel<true>, dim3(plan.num_remote_records),
                         dim3(64), 0, stream, plan.d_records, plan.d_remote_records, plan.d_faces,
                         plan.d_gather, w.pkf, w.pkm, w.recv, w.bcm);
    else
      hipLaunchKernelGGL(mortar_boundary_terms_kernel<false>, dim3(plan.num_remote_records),
                         dim3(64), 0, stream, plan.d_records, plan.d_remote_records, plan.d_faces,
                         plan.d_gather, w.pkf, w.pkm, w.recv, w.bcm);
    if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  }
  return SXD_OK;
}

int launch_mortar_lift(bool gh, const MortarPlan& plan, const MixedElement* el, double* dt,
                       double* work, hipStream_t stream) {
  const double* full = device_mortar_matrices();
  const double* half = device_half_mortar_matrices();
  if (full == nullptr || half == nullptr) return SXD_ERR_DEVICE;
  if (work == nullptr) return SXD_ERR_INVALID_ARGUMENT;
  const WorkLayout w(gh, plan, work);
  const int chunk = staged_chunk(gh ? 50 : 5, plan.tmp_max, kLiftCap);
  const size_t lds = sizeof(double) * (chunk * plan.tmp_max + plan.mat_max);
  for (int d = 0; d < 3; ++d) {
    const int count = plan.dim_begin[d + 1] - plan.dim_begin[d];
    if (count == 0) continue;
    if (gh)
      hipLaunchKernelGGL(mortar_lift_kernel<true>, dim3(count), dim3(256), lds, stream, el,
                         plan.d_faces, plan.dim_begin[d], plan.d_records, w.mag, w.bcm, dt, full,
                         half, chunk, plan.tmp_max);
    else
      hipLaunchKernelGGL(mortar_lift_kernel<false>, dim3(count), dim3(256), lds, stream, el,
                         plan.d_faces, plan.dim_begin[d], plan.d_records, w.mag, w.bcm, dt, full,
                         half, chunk, plan.tmp_max);
    if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  }
  return SXD_OK;
}

int launch_mortar_corrections(bool gh, const MortarPlan& plan, const MixedElement* el,
                              double* dt, double* work, hipStream_t stream) {
  const int rc = launch_mortar_boundary_terms(gh, plan, work, stream);
  return rc != SXD_OK ? rc : launch_mortar_lift(gh, plan, el, dt, work, stream);
}

int launch_mortar_faces(bool gh, const MortarPlan& plan, const MixedElement* el, const double* u,
                        const double* params, const double* inv_jac, double* dt, double* work,
                        hipStream_t stream) {
  if (plan.num_remote > 0) return SXD_ERR_UNSUPPORTED;
  const int rc = launch_mortar_packages(gh, plan, el, u, params, inv_jac, work, stream);
  return rc != SXD_OK ? rc : launch_mortar_corrections(gh, plan, el, dt, work, stream);
}

int launch_mortar_pack(bool gh, const MortarPlan& plan, double* work, void* send, int64_t step,
                       hipStream_t stream) {
  if (plan.num_remote == 0) return SXD_OK;
  if (work == nullptr || send == nullptr) return SXD_ERR_INVALID_ARGUMENT;
  const WorkLayout w(gh, plan, work);
  const int32_t lo = static_cast<int32_t>(static_cast<uint64_t>(step) & 0xffffffffu);
  const int32_t hi = static_cast<int32_t>(static_cast<uint64_t>(step) >> 32);
  unsigned char* out = static_cast<unsigned char*>(send);
  if (gh)
    hipLaunchKernelGGL(mortar_pack_kernel<134>, dim3(plan.num_remote), dim3(64), 0, stream,
                       plan.d_messages, plan.d_records, plan.d_faces, plan.d_gather, w.pkf, w.pkm,
                       out, 1, lo, hi);
  else
    hipLaunchKernelGGL(mortar_pack_kernel<13>, dim3(plan.num_remote), dim3(64), 0, stream,
                       plan.d_messages, plan.d_records, plan.d_faces, plan.d_gather, w.pkf, w.pkm,
                       out, 0, lo, hi);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

int launch_mortar_unpack(bool gh, const MortarPlan& plan, double* work, const void* recv,
                         int64_t bytes, int count, int64_t step, int32_t* status,
                         hipStream_t stream) {
  if (count == 0) return SXD_OK;
  if (work == nullptr || recv == nullptr || status == nullptr) return SXD_ERR_INVALID_ARGUMENT;
  const WorkLayout w(gh, plan, work);
  const int32_t lo = static_cast<int32_t>(static_cast<uint64_t>(step) & 0xffffffffu);
  const int32_t hi = static_cast<int32_t>(static_cast<uint64_t>(step) >> 32);
  const unsigned char* in = static_cast<const unsigned char*>(recv);
  if (gh)
    hipLaunchKernelGGL(mortar_unpack_kernel<134>, dim3(count), dim3(64), 0, stream, in, bytes,
                       plan.d_keys, plan.num_remote, w.recv, 1, lo, hi, status);
  else
    hipLaunchKernelGGL(mortar_unpack_kernel<13>, dim3(count), dim3(64), 0, stream, in, bytes,
                       plan.d_keys, plan.num_remote, w.recv, 0, lo, hi, status);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

}  // namespace sxd
