// Fused generalized-harmonic DG right-hand side: one workgroup per element,
// one thread per grid point, every term of ComputeTimeDerivative and
// ApplyBoundaryCorrections in a single pass over HBM.
//
// Phases:
//  A  point-wise setup.  Each thread loads its point's 50 fields and runs
//     gh_point_setup(): lapse, shift, inverse metrics, Christoffel symbols and
//     the algebraic part of dt Pi_ab, all in registers.
//  F  face data.  Owners of face points write their side's unit normal,
//     characteristic speeds, gamma2 and lift factor; threads 0 .. 6 N^2 - 1
//     compute the neighbour side of one face point each from the neighbour's
//     g_ab, gamma1, gamma2 and inverse Jacobian (HBM reads of face traces).
//  B  for each symmetric pair (a,b) (10 passes): stage g_ab, Pi_ab, Phi_{i,ab}
//     of the element in LDS; face threads form the lifted UpwindPenalty
//     correction of their face point; every thread forms the 15 derivatives
//     of its point from LDS pencils (sum k = 0..N-1 in order, then the
//     inverse-Jacobian contraction), the five dt components of the pair, adds
//     the corrections of the faces it lies on (xi, eta, zeta order) and
//     writes dt once.
//
// Each output follows the reference's operation sequence (see gh_point.hpp),
// so results equal the point-wise composite path bit for bit.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"
#include "gh_point.hpp"
#include "kernels.hpp"

namespace sxd {
namespace {

template <int N>
struct GhFusedLayout {
  static constexpr int NP = N * N * N;
  static constexpr int NF = N * N;        // points per face
  static constexpr int SLOTS = 6 * NF;    // face points of the element
  static constexpr int kIntStride = 12;   // n[3] nv[3] lam[4] gamma2 lift
  static constexpr int kExtStride = 11;   // n[3] nv[3] lam[4] gamma2
  // LDS of phases F and B (phase A works in registers)
  static constexpr int kFaceInt = 0;
  static constexpr int kFaceExt = kFaceInt + SLOTS * kIntStride;
  static constexpr int kStage = kFaceExt + SLOTS * kExtStride;
  static constexpr int kCorr = kStage + 5 * NP;
  static constexpr int kPhaseBSize = kCorr + SLOTS * 5;
  static constexpr int kDoubles = kPhaseBSize;
  static constexpr int kThreads = ((NP > SLOTS ? NP : SLOTS) + 63) / 64 * 64;
};

// Volume index of face point (a, b) of face `dir` (= 2 d + side).
template <int N>
__device__ __forceinline__ int face_volume_index(int dir, int a, int b) {
  const int d = dir >> 1;
  const int fixed = (dir & 1) ? N - 1 : 0;
  if (d == 0) return fixed + N * (a + N * b);
  if (d == 1) return a + N * (fixed + N * b);
  return a + N * (b + N * fixed);
}

template <int N>
__global__ void __launch_bounds__(GhFusedLayout<N>::kThreads)
gh_rhs_fused_kernel(const int32_t* __restrict__ nbr, const double* __restrict__ u,
                    const double* __restrict__ gammas, const double* __restrict__ inv_jac,
                    double* __restrict__ dt, const double* __restrict__ dtab) {
  using L = GhFusedLayout<N>;
  constexpr int NP = L::NP;
  constexpr int NF = L::NF;
  __shared__ double smem[L::kDoubles];
  const int64_t e = blockIdx.x;
  const int t = threadIdx.x;
  const bool active = t < NP;
  const double* ue = u + e * 50 * NP;
  const int i = t % N, j = (t / N) % N, k = t / (N * N);

  // ---- phase A -------------------------------------------------------------
  GhPointState st;
  double ij[9];
  double gamma1 = 0.0;
  if (active) {
    double up[50];
#pragma unroll
    for (int c = 0; c < 50; ++c) up[c] = ue[c * NP + t];
    const double* ga = gammas + e * 3 * NP + t;
    gamma1 = ga[NP];
    gh_point_setup(st, up, ga[0], gamma1, ga[2 * NP]);
#pragma unroll
    for (int c = 0; c < 9; ++c) ij[c] = inv_jac[(e * 9 + c) * NP + t];
  }

  // ---- phase F -------------------------------------------------------------
  if (active) {
    const int idx[3] = {i, j, k};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (idx[d] != 0 && idx[d] != N - 1) continue;
      const int side = idx[d] == 0 ? 0 : 1;
      const int a = d == 0 ? j : i;
      const int b = d == 2 ? j : k;
      double* fi = smem + L::kFaceInt + ((2 * d + side) * NF + a + N * b) * L::kIntStride;
      GhFaceSide s;
      double mag;
#pragma unroll
      for (int c = 0; c < 3; ++c) s.n[c] = side == 0 ? ij[ij3(d, c)] * -1.0 : ij[ij3(d, c)];
      gh_normalize(s.n, s.nv, &mag, st.inv_sp);
      gh_char_speeds(s.lam, gamma1, st.lapse, st.shift, s.n);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        fi[c] = s.n[c];
        fi[3 + c] = s.nv[c];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) fi[6 + c] = s.lam[c];
      fi[10] = st.gamma2;
      fi[11] = lift_factor(N, mag);
    }
  }
  // neighbour side of face slot t
  int slot_q = 0;        // neighbour volume index of this thread's face slot
  int64_t slot_en = 0;   // neighbour element
  const bool face_thread = t < L::SLOTS;
  if (face_thread) {
    const int dir = t / NF;
    const int f = t % NF;
    const int d = dir >> 1, side = dir & 1;
    slot_en = nbr[6 * e + dir];
    slot_q = face_volume_index<N>(dir ^ 1, f % N, f / N);
    const double* un = u + slot_en * 50 * NP + slot_q;
    double g[10];
#pragma unroll
    for (int c = 0; c < 10; ++c) g[c] = un[c * NP];
    double lapse, shift[3], inv_sp[6], mag;
    gh_lapse_shift(&lapse, shift, inv_sp, g);
    const double* gn = gammas + slot_en * 3 * NP + slot_q;
    GhFaceSide s;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      s.n[c] = inv_jac[(slot_en * 9 + ij3(d, c)) * NP + slot_q];
      if (side == 1) s.n[c] = s.n[c] * -1.0;
    }
    gh_normalize(s.n, s.nv, &mag, inv_sp);
    gh_char_speeds(s.lam, gn[NP], lapse, shift, s.n);
    double* fe = smem + L::kFaceExt + t * L::kExtStride;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      fe[c] = s.n[c];
      fe[3 + c] = s.nv[c];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) fe[6 + c] = s.lam[c];
    fe[10] = gn[2 * NP];
  }

  // differentiation matrix rows of this point
  const double* D = diff_matrix(dtab, N);
  double Di[N], Dj[N], Dk[N];
#pragma unroll
  for (int l = 0; l < N; ++l) {
    Di[l] = active ? D[i * N + l] : 0.0;
    Dj[l] = active ? D[j * N + l] : 0.0;
    Dk[l] = active ? D[k * N + l] : 0.0;
  }

  // ---- phase B -------------------------------------------------------------
  double* stage = smem + L::kStage;
  double* corr = smem + L::kCorr;
  double* dte = dt + e * 50 * NP;
  const int64_t nb_base = slot_en * 50 * NP + slot_q;
#pragma unroll 1
  for (int p = 0; p < 10; ++p) {
    const int comp[5] = {p, 10 + p, 20 + 3 * p, 21 + 3 * p, 22 + 3 * p};
    __syncthreads();  // previous pass done with stage and corr
    if (active) {
#pragma unroll
      for (int c = 0; c < 5; ++c) stage[c * NP + t] = ue[comp[c] * NP + t];
    }
    double ext_v[5];
    if (face_thread) {
#pragma unroll
      for (int c = 0; c < 5; ++c) ext_v[c] = u[nb_base + comp[c] * NP];
    }
    __syncthreads();
    if (face_thread) {
      const int dir = t / NF;
      const int f = t % NF;
      const int pv = face_volume_index<N>(dir, f % N, f / N);
      const double* fi = smem + L::kFaceInt + t * L::kIntStride;
      const double* fe = smem + L::kFaceExt + t * L::kExtStride;
      GhFaceSide si, se;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        si.n[c] = fi[c];
        si.nv[c] = fi[3 + c];
        se.n[c] = fe[c];
        se.nv[c] = fe[3 + c];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        si.lam[c] = fi[6 + c];
        se.lam[c] = fe[6 + c];
      }
      si.gamma2 = fi[10];
      se.gamma2 = fe[10];
      const double fac = fi[11];
      double phi_i[3], phi_e[3], pki[13], pke[13], bc[5];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        phi_i[c] = stage[(2 + c) * NP + pv];
        phi_e[c] = ext_v[2 + c];
      }
      gh_package_pair(pki, stage[pv], stage[NP + pv], phi_i, si);
      gh_package_pair(pke, ext_v[0], ext_v[1], phi_e, se);
      gh_boundary_terms_pair(bc, pki, si.lam, pke, se.lam);
#pragma unroll
      for (int c = 0; c < 5; ++c) corr[t * 5 + c] = bc[c] * fac;
    }
    double out[5];
    if (active) {
      double dlog[15];
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        const double* sc = stage + c * NP;
        double d0 = Di[0] * sc[N * (j + N * k)];
#pragma unroll
        for (int l = 1; l < N; ++l) d0 = fma(Di[l], sc[l + N * (j + N * k)], d0);
        double d1 = Dj[0] * sc[i + N * N * k];
#pragma unroll
        for (int l = 1; l < N; ++l) d1 = fma(Dj[l], sc[i + N * (l + N * k)], d1);
        double d2 = Dk[0] * sc[i + N * j];
#pragma unroll
        for (int l = 1; l < N; ++l) d2 = fma(Dk[l], sc[i + N * (j + N * l)], d2);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          double s = ij[ij3(0, a)] * d0;
          s = fma(ij[ij3(1, a)], d1, s);
          s = fma(ij[ij3(2, a)], d2, s);
          dlog[3 * c + a] = s;
        }
      }
      // dlog[3 c + m] = d_m of component c; c = 0 g, 1 Pi, 2..4 Phi_n
      double phi[3], dphi[9];
#pragma unroll
      for (int n = 0; n < 3; ++n) {
        phi[n] = stage[(2 + n) * NP + t];
#pragma unroll
        for (int m = 0; m < 3; ++m) dphi[m + 3 * n] = dlog[3 * (2 + n) + m];
      }
      gh_point_pair(out, st, p, stage[t], stage[NP + t], phi, dlog, dlog + 3, dphi);
    }
    __syncthreads();  // corrections complete
    if (active) {
      const int idx[3] = {i, j, k};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (idx[d] != 0 && idx[d] != N - 1) continue;
        const int side = idx[d] == 0 ? 0 : 1;
        const int a = d == 0 ? j : i;
        const int b = d == 2 ? j : k;
        const double* cr = corr + ((2 * d + side) * NF + a + N * b) * 5;
#pragma unroll
        for (int c = 0; c < 5; ++c) out[c] = out[c] + cr[c];
      }
#pragma unroll
      for (int c = 0; c < 5; ++c) dte[comp[c] * NP + t] = out[c];
    }
  }
}

template <int N>
int launch_gh_fused_n(int num_elements, const int32_t* nbr, const double* u,
                      const double* gammas, const double* inv_jac, double* dt,
                      const double* dtab, hipStream_t stream) {
  hipLaunchKernelGGL(gh_rhs_fused_kernel<N>, dim3(num_elements),
                     dim3(GhFusedLayout<N>::kThreads), 0, stream, nbr, u, gammas, inv_jac,
                     dt, dtab);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

}  // namespace

bool gh_fused_supported(const int32_t ext[3]) {
  return ext[0] == ext[1] && ext[1] == ext[2] && ext[0] >= 2 && ext[0] <= 8;
}

int launch_gh_rhs_fused(int num_elements, const int32_t ext[3], const int32_t* nbr,
                        const double* u, const double* gammas, const double* inv_jac,
                        double* dt, hipStream_t stream) {
  const double* dtab = device_diff_matrices();
  if (dtab == nullptr) return SXD_ERR_DEVICE;
  if (!gh_fused_supported(ext)) return SXD_ERR_UNSUPPORTED;
  switch (ext[0]) {
    case 2: return launch_gh_fused_n<2>(num_elements, nbr, u, gammas, inv_jac, dt, dtab, stream);
    case 3: return launch_gh_fused_n<3>(num_elements, nbr, u, gammas, inv_jac, dt, dtab, stream);
    case 4: return launch_gh_fused_n<4>(num_elements, nbr, u, gammas, inv_jac, dt, dtab, stream);
    case 5: return launch_gh_fused_n<5>(num_elements, nbr, u, gammas, inv_jac, dt, dtab, stream);
    case 6: return launch_gh_fused_n<6>(num_elements, nbr, u, gammas, inv_jac, dt, dtab, stream);
    case 7: return launch_gh_fused_n<7>(num_elements, nbr, u, gammas, inv_jac, dt, dtab, stream);
    case 8: return launch_gh_fused_n<8>(num_elements, nbr, u, gammas, inv_jac, dt, dtab, stream);
    default: return SXD_ERR_UNSUPPORTED;
  }
}

}  // namespace sxd
