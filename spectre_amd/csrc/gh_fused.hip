// Fused generalized-harmonic DG right-hand side: one workgroup per element,
// every term of ComputeTimeDerivative and ApplyBoundaryCorrections in a single
// pass over HBM.
//
// Register budget.  The point-wise GH algebra keeps ~235 FP64 registers live
// per grid point.  A CDNA4 SIMD has 512 VGPRs per lane.  Two tile shapes:
//  - 256 threads per element: one wave per SIMD, each thread owns P =
//    ceil(Np / 256) grid points (P = 2 at 8^3); with the vector-ALU
//    derivative engine nothing spills, and the two points give each thread
//    two independent instruction streams;
//  - 512 threads (N = 7, 8): one point per thread and two waves per SIMD in
//    256 registers, with part of the phase-A state spilled; the second wave
//    covers the other's LDS and barrier waits.
//
// Phases:
//  A  point-wise setup for each owned point: load the 50 fields, run
//     gh_point_setup() (lapse, shift, inverse metrics, Christoffel symbols and
//     the algebraic part of dt Pi_ab).
//  F  face data.  Owners of face points write their side's unit normal,
//     characteristic speeds, gamma2 and lift factor to LDS; each face point
//     of the element (6 N^2 slots, strided over the threads) gets the
//     neighbour side computed from the neighbour's g_ab, gamma1, gamma2 and
//     inverse Jacobian at the matching point (HBM reads of face traces).
//     GHOST instantiations also take external faces (neighbour -1): their
//     slots read the ghost state at the element's own face point and use the
//     element's gamma1, gamma2 and reversed unit normal (Dirichlet analytic
//     condition, GeneralizedHarmonic/BoundaryConditions/DirichletAnalytic.cpp:
//     70-115, Actions/BoundaryConditionsImpl.hpp:427-586).
//  B  for each symmetric pair (a,b) (10 passes): stage g_ab, Pi_ab, Phi_{i,ab}
//     of the element in LDS (the next pass's values are prefetched while the
//     current pass computes: at 8^3 by direct global-to-LDS copies into the
//     idle stage buffer, otherwise into registers); face slots form the lifted
//     UpwindPenalty correction; every point forms the 15 derivatives of the
//     pair, contracts them with the inverse Jacobian, forms the five dt
//     components of the pair, adds the corrections of the faces it lies on
//     (xi, eta, zeta order) and writes dt once.
//
// Two derivative engines (template parameter MFMA):
//  - vector ALUs: each point sums its three LDS pencils, l = 0 .. N-1 in
//    order.  One barrier per pass: the stage and correction tiles are
//    double-buffered, and a pass's face corrections are added (and its dt
//    written) after the next pass's staging barrier;
//  - matrix cores, described next (5 barriers per pass, logical derivatives
//    go through LDS).
// At 8^3 the vector-ALU engine is faster: the MFMA engine needs the extra
// LDS round trip and barriers, and the register file is already full with
// the point-wise GH state, so few MFMA chains can be in flight.
//
// Derivatives on the matrix cores.  For direction d the logical derivative of
// a component is D (N x N) times the component arranged as N x N^2 (rows:
// the index along d, columns: the N^2 pencils).  v_mfma_f64_16x16x4f64 takes
// the 16 x 16 A operand as block-diag(D, D, ...) over CPT = 16 / N components
// (K = CPT N rows of the B operand are those components' pencil values), so
// one 16 x 16 output tile holds CPT components x N outputs for 16 pencils.
// The instruction accumulates the four K terms of each step one after the
// other with one rounding per term, i.e. it performs exactly the sequential
// fma chain sum_{l=0}^{N-1} D_il u_l of the reference's dgemm order (the zero
// blocks of block-diag(D, ...) add exact zeros), so the MFMA path is bit for
// bit the same as the vector-ALU engine.
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

typedef double mfma_acc_t __attribute__((ext_vector_type(4)));

// TPB: threads per workgroup (256: one wave per SIMD, two points per thread
// at 8^3; 512: two waves per SIMD, one point per thread, 256 registers).
template <int N, bool MFMA, int TPB>
struct GhFusedLayout {
  static constexpr int NP = N * N * N;
  static constexpr int NF = N * N;                   // points per face
  static constexpr int SLOTS = 6 * NF;               // face points of the element
  static constexpr int THREADS = NP < TPB ? (NP + 63) / 64 * 64 : TPB;
  static constexpr int WAVES = THREADS / 64;
  static constexpr int P = (NP + THREADS - 1) / THREADS;          // points per thread
  static constexpr int PF = (SLOTS + THREADS - 1) / THREADS;      // face slots per thread
  // MFMA tiling of the derivative products
  static constexpr int CPT = 16 / N;                 // components per 16-row tile
  static constexpr int ROWS = CPT * N;               // used rows (and K) of a tile
  static constexpr int KSTEPS = (ROWS + 3) / 4;
  static constexpr int TILES = (NF + 15) / 16;       // 16-pencil column tiles
  // LDS (doubles).  Face records and corrections are field-major (field f
  // of slot s at f * SLOTS + s): the lanes of a wave own consecutive slots,
  // so their record accesses fall in distinct banks.
  static constexpr int kIntFields = 12;  // n[3] nv[3] lam[4] gamma2 lift
  static constexpr int kExtFields = 11;  // n[3] nv[3] lam[4] gamma2
  static constexpr int kFaceInt = 0;
  static constexpr int kFaceExt = kFaceInt + SLOTS * kIntFields;
  // Staging by direct global-to-LDS copies (global_load_lds_dwordx4) for the
  // vector-ALU engine at 8^3: the next pass's fields go straight into the
  // idle stage buffer, with no registers in between.  The stage then also
  // holds the pass's parked dt Pi partial sum (component 5).
  static constexpr bool DMA = !MFMA && N == 8;
  static constexpr int STAGED = DMA ? 6 : 5;
  static constexpr int kStage = (kFaceExt + SLOTS * kExtFields + 1) / 2 * 2;  // 16-byte aligned
  static constexpr int kCorr = kStage + (STAGED * NP + 1) / 2 * 2;
  static constexpr int kDeriv = kCorr + 5 * SLOTS;  // MFMA: logical derivatives
  // vector ALUs: second stage and correction buffers (odd passes), so a pass
  // can stage its fields while slower waves still read the previous pass's
  static constexpr int kStage2 = (kDeriv + (MFMA ? 9 * NP : 0) + 1) / 2 * 2;
  static constexpr int kCorr2 = kStage2 + (MFMA ? 0 : (STAGED * NP + 1) / 2 * 2);
  static constexpr int kDoubles = kCorr2 + (MFMA ? 0 : 5 * SLOTS);
  // Staged rows (vector ALUs, N = 8) are swizzled: within row (j, k) the
  // 16-byte pair s of xi values is stored at pair s ^ ((j >> 1) & 3).  The
  // lanes of a wave read the xi rows j = 0..7 of one k (64 bytes apart) two
  // values at a time; with the swizzle the eight rows' pairs fall in eight
  // distinct groups of four banks instead of two.
  static constexpr bool SWIZZLE = !MFMA && N == 8;
};

// Position within its row of xi index i of row j in the staged fields.
template <int N, bool SWIZZLE>
__device__ __forceinline__ int staged_i(int i, int j) {
  if constexpr (SWIZZLE) return i ^ (2 * ((j >> 1) & 3));
  return i;
}

// Staged position of volume point pt = i + N (j + N k).
template <int N, bool SWIZZLE>
__device__ __forceinline__ int staged_point(int pt) {
  const int i = pt % N, j = (pt / N) % N;
  return pt - i + staged_i<N, SWIZZLE>(i, j);
}

// Volume index of face point (a, b) of face `dir` (= 2 d + side).
template <int N>
__device__ __forceinline__ int face_volume_index(int dir, int a, int b) {
  const int d = dir >> 1;
  const int fixed = (dir & 1) ? N - 1 : 0;
  if (d == 0) return fixed + N * (a + N * b);
  if (d == 1) return a + N * (fixed + N * b);
  return a + N * (b + N * fixed);
}

// Volume index of entry l of pencil n (n = a + N b over the two other
// dimensions, lowest fastest) in direction d.
template <int N>
__device__ __forceinline__ int pencil_point(int d, int l, int n) {
  const int a = n % N, b = n / N;
  if (d == 0) return l + N * (a + N * b);
  if (d == 1) return a + N * (l + N * b);
  return a + N * (b + N * l);
}

// Logical derivatives of `ncomp` staged components (stage slots first ..
// first + ncomp - 1) in all three directions, written to
// deriv[(c * 3 + d) * NP + point] (c = 0 .. ncomp - 1).  The jobs
// (direction, component group, pencil tile) are spread over the waves.
template <int N, int NCOMP, int TPB>
__device__ __forceinline__ void mfma_logical_derivatives(const double* stage, double* deriv,
                                                         int first, const double* afrag,
                                                         int wave, int lane) {
  using L = GhFusedLayout<N, true, TPB>;
  constexpr int NP = L::NP;
  constexpr int GROUPS = (NCOMP + L::CPT - 1) / L::CPT;
  constexpr int JOBS = 3 * GROUPS * L::TILES;
  constexpr int PER_WAVE = (JOBS + L::WAVES - 1) / L::WAVES;
  constexpr int BATCH = 1;  // more chains in flight spill registers at N = 8
  const int m_lane = lane % 16;
  const int k_lane = lane / 16;
#pragma unroll
  for (int j0 = 0; j0 < PER_WAVE; j0 += BATCH) {
    double b[BATCH][L::KSTEPS];
    int jd[BATCH], jc0[BATCH], jnc[BATCH], jn[BATCH];
    bool jok[BATCH];
    // operand loads of the whole batch first
#pragma unroll
    for (int x = 0; x < BATCH; ++x) {
      const int job = wave + (j0 + x) * L::WAVES;  // wave-uniform
      jok[x] = (j0 + x < PER_WAVE) && job < JOBS;
      const int jj = jok[x] ? job : 0;
      jd[x] = jj % 3;
      const int gr = (jj / 3) % GROUPS;
      const int tl = jj / (3 * GROUPS);
      jc0[x] = gr * L::CPT;
      jnc[x] = min(L::CPT, NCOMP - jc0[x]);
      jn[x] = 16 * tl + m_lane;
#pragma unroll
      for (int s = 0; s < L::KSTEPS; ++s) {
        const int kk = 4 * s + k_lane;
        const int c2 = kk / N, l = kk % N;
        b[x][s] = (jok[x] && jn[x] < L::NF && kk < jnc[x] * N)
                      ? stage[(first + jc0[x] + c2) * NP + pencil_point<N>(jd[x], l, jn[x])]
                      : 0.0;
      }
    }
    mfma_acc_t acc[BATCH];
#pragma unroll
    for (int x = 0; x < BATCH; ++x) acc[x] = mfma_acc_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < L::KSTEPS; ++s)
#pragma unroll
      for (int x = 0; x < BATCH; ++x)
        acc[x] = __builtin_amdgcn_mfma_f64_16x16x4f64(afrag[s], b[x][s], acc[x], 0, 0, 0);
#pragma unroll
    for (int x = 0; x < BATCH; ++x)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = 4 * r + k_lane;
        const int c2 = m / N, i = m % N;
        if (jok[x] && jn[x] < L::NF && m < jnc[x] * N)
          deriv[((jc0[x] + c2) * 3 + jd[x]) * NP + pencil_point<N>(jd[x], i, jn[x])] =
              acc[x][r];
      }
  }
}

template <int N, bool MFMA, int TPB, bool GHOST>
__global__ void __launch_bounds__((GhFusedLayout<N, MFMA, TPB>::THREADS), 1)
gh_rhs_fused_kernel(const int32_t* __restrict__ nbr, const double* __restrict__ u,
                    const double* __restrict__ gammas, const double* __restrict__ inv_jac,
                    double* __restrict__ dt, const double* __restrict__ dtab,
                    const double* __restrict__ ghost) {
  using L = GhFusedLayout<N, MFMA, TPB>;
  constexpr int NP = L::NP, NF = L::NF, P = L::P, PF = L::PF, T = L::THREADS;
  __shared__ __align__(16) double smem[L::kDoubles];
  const int64_t e = xcd_local_element(blockIdx.x, gridDim.x);
  const int t = threadIdx.x;
  const int wave = t / 64, lane = t % 64;
  const double* ue = u + e * 50 * NP;

  // points owned by this thread: pt[q] = t + q T
  bool act[P];
  int pi_[P], pj[P], pk[P], spt[P];  // spt: staged position (vector-ALU engine)
#pragma unroll
  for (int q = 0; q < P; ++q) {
    const int pt = t + q * T;
    act[q] = pt < NP;
    pi_[q] = pt % N;
    pj[q] = (pt / N) % N;
    pk[q] = act[q] ? pt / (N * N) : 0;
    spt[q] = act[q] ? staged_point<N, L::SWIZZLE>(pt) : 0;
  }

  // ---- phase A -------------------------------------------------------------
  GhPointState st[P];
  double ij[P][9];
  double gamma1[P];
#pragma unroll
  for (int q = 0; q < P; ++q) {
    gamma1[q] = 0.0;
    if (act[q]) {
      const int pt = t + q * T;
      double up[50];
#pragma unroll
      for (int c = 0; c < 50; ++c) up[c] = ue[c * NP + pt];
      const double* ga = gammas + e * 3 * NP + pt;
      gamma1[q] = ga[NP];
      gh_point_setup(st[q], up, ga[0], gamma1[q], ga[2 * NP]);
      // dt Pi_ab through the Christoffel terms: parked in dt (row Pi_ab of
      // this point, which only this thread touches) until pass ab reads it
      // back, so that it holds no registers through the passes
#pragma unroll
      for (int r = 0; r < 10; ++r) dt[(e * 50 + 10 + r) * NP + pt] = st[q].dt_pi_partial[r];
#pragma unroll
      for (int c = 0; c < 9; ++c) ij[q][c] = inv_jac[(e * 9 + c) * NP + pt];
    }
  }

  // ---- phase F -------------------------------------------------------------
#pragma unroll
  for (int q = 0; q < P; ++q) {
    if (!act[q]) continue;
    const int idx[3] = {pi_[q], pj[q], pk[q]};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (idx[d] != 0 && idx[d] != N - 1) continue;
      const int side = idx[d] == 0 ? 0 : 1;
      const int a = d == 0 ? idx[1] : idx[0];
      const int b = d == 2 ? idx[1] : idx[2];
      double* fi = smem + L::kFaceInt + (2 * d + side) * NF + a + N * b;
      GhFaceSide s;
      double mag;
#pragma unroll
      for (int c = 0; c < 3; ++c)
        s.n[c] = side == 0 ? ij[q][ij3(d, c)] * -1.0 : ij[q][ij3(d, c)];
      gh_normalize(s.n, s.nv, &mag, st[q].inv_sp);
      gh_char_speeds(s.lam, gamma1[q], st[q].lapse, st[q].shift, s.n);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        fi[c * L::SLOTS] = s.n[c];
        fi[(3 + c) * L::SLOTS] = s.nv[c];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) fi[(6 + c) * L::SLOTS] = s.lam[c];
      fi[10 * L::SLOTS] = st[q].gamma2;
      fi[11 * L::SLOTS] = lift_factor(N, mag);
    }
  }
  // neighbour side of the face slots s = t + f T
  bool fact[PF];
  int64_t nb_base[PF];  // neighbour element offset + neighbour volume index
  bool nb_ghost[PF];    // GHOST: nb_base indexes the ghost state
#pragma unroll
  for (int f = 0; f < PF; ++f) {
    const int slot = t + f * T;
    fact[f] = slot < L::SLOTS;
    nb_base[f] = 0;
    nb_ghost[f] = false;
    if (!fact[f]) continue;
    const int dir = slot / NF;
    const int fp = slot % NF;
    const int d = dir >> 1, side = dir & 1;
    const int64_t en = nbr[6 * e + dir];
    if constexpr (GHOST) {
      if (en < 0) {
        // the element's own face point and its unit normal, formed as the
        // owner of the point forms it in the loop above
        const int po = face_volume_index<N>(dir, fp % N, fp / N);
        nb_ghost[f] = true;
        nb_base[f] = e * 50 * NP + po;
        double go[10], lo, so[3], iso[6], mag;
#pragma unroll
        for (int c = 0; c < 10; ++c) go[c] = ue[c * NP + po];
        gh_lapse_shift(&lo, so, iso, go);
        double n_in[3], nv_in[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double r = inv_jac[(e * 9 + ij3(d, c)) * NP + po];
          n_in[c] = side == 0 ? r * -1.0 : r;
        }
        gh_normalize(n_in, nv_in, &mag, iso);
        const double* gx = ghost + nb_base[f];
        double g[10], lapse, shift[3], inv_sp[6];
#pragma unroll
        for (int c = 0; c < 10; ++c) g[c] = gx[c * NP];
        gh_lapse_shift(&lapse, shift, inv_sp, g);
        GhFaceSide s;
#pragma unroll
        for (int c = 0; c < 3; ++c) s.n[c] = n_in[c] * -1.0;
        gh_ghost_normal(s.n, s.nv, inv_sp);
        const double* go_ = gammas + e * 3 * NP + po;
        gh_char_speeds(s.lam, go_[NP], lapse, shift, s.n);
        double* fe = smem + L::kFaceExt + slot;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          fe[c * L::SLOTS] = s.n[c];
          fe[(3 + c) * L::SLOTS] = s.nv[c];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) fe[(6 + c) * L::SLOTS] = s.lam[c];
        fe[10 * L::SLOTS] = go_[2 * NP];
        continue;
      }
    }
    const int q = face_volume_index<N>(dir ^ 1, fp % N, fp / N);
    nb_base[f] = en * 50 * NP + q;
    const double* un = u + nb_base[f];
    double g[10];
#pragma unroll
    for (int c = 0; c < 10; ++c) g[c] = un[c * NP];
    double lapse, shift[3], inv_sp[6], mag;
    gh_lapse_shift(&lapse, shift, inv_sp, g);
    const double* gn = gammas + en * 3 * NP + q;
    GhFaceSide s;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      s.n[c] = inv_jac[(en * 9 + ij3(d, c)) * NP + q];
      if (side == 1) s.n[c] = s.n[c] * -1.0;
    }
    gh_normalize(s.n, s.nv, &mag, inv_sp);
    gh_char_speeds(s.lam, gn[NP], lapse, shift, s.n);
    double* fe = smem + L::kFaceExt + slot;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      fe[c * L::SLOTS] = s.n[c];
      fe[(3 + c) * L::SLOTS] = s.nv[c];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) fe[(6 + c) * L::SLOTS] = s.lam[c];
    fe[10 * L::SLOTS] = gn[2 * NP];
  }

  // Derivative operands.  MFMA: the A operand block-diag(D, ..., D), step s
  // holds A[m][4 s + lane / 16] for m = lane % 16.  VALU: the rows of D that
  // each owned point needs in the three directions.
  constexpr int KS = MFMA ? L::KSTEPS : 1;
  constexpr int DR = MFMA ? 1 : N;
  double afrag[KS];
  double Di[P][DR], Dj[P][DR], Dk[P][DR];
  {
    const double* D = diff_matrix(dtab, N);
    if constexpr (MFMA) {
      const int m = lane % 16;
#pragma unroll
      for (int s = 0; s < L::KSTEPS; ++s) {
        const int kk = 4 * s + lane / 16;
        afrag[s] = (m < L::ROWS && kk < L::ROWS && m / N == kk / N) ? D[(m % N) * N + kk % N]
                                                                     : 0.0;
      }
    } else {
#pragma unroll
      for (int q = 0; q < P; ++q) {
        // with N^2 = 64 every wave covers whole k-planes: its k is uniform,
        // so its row of D is read by scalar loads and held in no VGPR
        int kq = pk[q];
        if constexpr (N * N == 64 && T % 64 == 0)
          kq = __builtin_amdgcn_readfirstlane(t / 64) + q * (T / 64);
#pragma unroll
        for (int l = 0; l < N; ++l) {
          Di[q][l] = D[pi_[q] * N + l];
          Dj[q][l] = D[pj[q] * N + l];
          Dk[q][l] = D[kq * N + l];
        }
      }
    }
  }

  // ---- phase B -------------------------------------------------------------
  double* stage = smem + L::kStage;
  double* corr = smem + L::kCorr;
  double* deriv = smem + L::kDeriv;
  double* dte = dt + e * 50 * NP;
  // prefetched fields of the next pass: own points and neighbour face points
  double own_next[P][5], ext_next[PF][5], pin_next[P];
  auto load_pass = [&](int p) {
    const int comp[5] = {p, 10 + p, 20 + 3 * p, 21 + 3 * p, 22 + 3 * p};
    if constexpr (!L::DMA) {
#pragma unroll
      for (int q = 0; q < P; ++q) {
#pragma unroll
        for (int c = 0; c < 5; ++c) own_next[q][c] = act[q] ? ue[comp[c] * NP + t + q * T] : 0.0;
        pin_next[q] = act[q] ? dte[(10 + p) * NP + t + q * T] : 0.0;
      }
    }
#pragma unroll
    for (int f = 0; f < PF; ++f)
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        if constexpr (GHOST) {
          ext_next[f][c] =
              fact[f] ? (nb_ghost[f] ? ghost : u)[nb_base[f] + comp[c] * NP] : 0.0;
        } else {
          ext_next[f][c] = fact[f] ? u[nb_base[f] + comp[c] * NP] : 0.0;
        }
      }
  };
  // Direct copy of pass p's six staged components into the stage buffer at
  // `buf`: 24 jobs of 64 lanes x 16 bytes (component c = job / 4, quarter
  // job % 4), spread over the waves.  Lane L of a job fills 16-byte unit U = 64 (job %
  // 4) + L of the component, i.e. pair U & 3 of staged row U >> 2; it reads
  // the pair the swizzle stores there (pair (U & 3) ^ ((j >> 1) & 3) of the
  // row, j = row % 8).
  auto dma_stage = [&](int p, double* buf) {
    if constexpr (L::DMA) {
#pragma unroll
      for (int m = 0; m < 24 / L::WAVES; ++m) {
        const int job = wave + L::WAVES * m;  // wave-uniform
        const int c = job >> 2, quarter = job & 3;
        const int comp = c == 0 ? p : c == 1 ? 10 + p : c == 5 ? 10 + p : 20 + 3 * p + (c - 2);
        const double* src = (c == 5 ? dte : ue) + comp * NP;
        const int unit = 64 * quarter + lane;
        const int row = unit >> 2;
        const int src_unit = (row << 2) | ((unit & 3) ^ (((row & 7) >> 1) & 3));
        __builtin_amdgcn_global_load_lds(
            (const __attribute__((address_space(1))) void*)(src + 2 * src_unit),
            (__attribute__((address_space(3))) void*)(buf + c * NP + 128 * quarter), 16, 0, 0);
      }
    }
  };
  // inertial derivatives of an owned point from its three logical ones
  auto contract = [&](int q, double d0, double d1, double d2, double* out3) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double s = ij[q][ij3(0, a)] * d0;
      s = fma(ij[q][ij3(1, a)], d1, s);
      s = fma(ij[q][ij3(2, a)], d2, s);
      out3[a] = s;
    }
  };
  // ... from the logical ones of staged component c_local in LDS
  auto contract_lds = [&](int q, int c_local, double* out3) {
    const int pt = t + q * T;
    contract(q, deriv[(c_local * 3 + 0) * NP + pt], deriv[(c_local * 3 + 1) * NP + pt],
             deriv[(c_local * 3 + 2) * NP + pt], out3);
  };
  // lifted boundary corrections of the face slots of the staged pair
  auto face_corrections = [&](const double* ext_v) {
#pragma unroll
    for (int f = 0; f < PF; ++f) {
      if (!fact[f]) continue;
      const int slot = t + f * T;
      const int dir = slot / NF;
      const int fp = slot % NF;
      const int pv = face_volume_index<N>(dir, fp % N, fp / N);
      const double* fi = smem + L::kFaceInt + slot;
      const double* fe = smem + L::kFaceExt + slot;
      GhFaceSide si, se;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        si.n[c] = fi[c * L::SLOTS];
        si.nv[c] = fi[(3 + c) * L::SLOTS];
        se.n[c] = fe[c * L::SLOTS];
        se.nv[c] = fe[(3 + c) * L::SLOTS];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        si.lam[c] = fi[(6 + c) * L::SLOTS];
        se.lam[c] = fe[(6 + c) * L::SLOTS];
      }
      si.gamma2 = fi[10 * L::SLOTS];
      se.gamma2 = fe[10 * L::SLOTS];
      const double fac = fi[11 * L::SLOTS];
      double phi_i[3], phi_e[3], pki[13], pke[13], bc[5];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        phi_i[c] = stage[(2 + c) * NP + pv];
        phi_e[c] = ext_v[5 * f + 2 + c];
      }
      gh_package_pair(pki, stage[pv], stage[NP + pv], phi_i, si);
      gh_package_pair(pke, ext_v[5 * f], ext_v[5 * f + 1], phi_e, se);
      gh_boundary_terms_pair(bc, pki, si.lam, pke, se.lam);
#pragma unroll
      for (int c = 0; c < 5; ++c) corr[c * L::SLOTS + slot] = bc[c] * fac;
    }
  };
  // add the corrections of the faces point q lies on (xi, eta, zeta order;
  // GHOST: external faces first, then internal ones, each in that order)
  int ext_mask = 0;
  if constexpr (GHOST) {
#pragma unroll
    for (int dir = 0; dir < 6; ++dir) ext_mask |= (nbr[6 * e + dir] < 0 ? 1 : 0) << dir;
  }
  auto add_corrections = [&](int q, double* out) {
    const int idx[3] = {pi_[q], pj[q], pk[q]};
#pragma unroll
    for (int step = 0; step < (GHOST ? 6 : 3); ++step) {
      const int d = step % 3;
      if (idx[d] != 0 && idx[d] != N - 1) continue;
      const int side = idx[d] == 0 ? 0 : 1;
      if constexpr (GHOST) {
        if ((((ext_mask >> (2 * d + side)) & 1) != 0) != (step < 3)) continue;
      }
      const int a = d == 0 ? idx[1] : idx[0];
      const int b = d == 2 ? idx[1] : idx[2];
      const double* cr = corr + (2 * d + side) * NF + a + N * b;
#pragma unroll
      for (int c = 0; c < 5; ++c) out[c] = out[c] + cr[c * L::SLOTS];
    }
  };
  if constexpr (MFMA) {
    load_pass(0);
#pragma unroll 1
    for (int p = 0; p < 10; ++p) {
      lds_barrier();  // previous pass done with stage, corr and deriv
      // the pair's own values stay in registers for the point-wise terms
      double own[P][5], ext_v[PF][5], pin[P];
#pragma unroll
      for (int q = 0; q < P; ++q) {
        pin[q] = pin_next[q];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
          own[q][c] = own_next[q][c];
          if (act[q]) stage[c * NP + t + q * T] = own[q][c];
        }
      }
#pragma unroll
      for (int f = 0; f < PF; ++f)
#pragma unroll
        for (int c = 0; c < 5; ++c) ext_v[f][c] = ext_next[f][c];
      if (p + 1 < 10) load_pass(p + 1);
      lds_barrier();  // stage complete
      const int comp[5] = {p, 10 + p, 20 + 3 * p, 21 + 3 * p, 22 + 3 * p};
      double out[P][5];
      mfma_logical_derivatives<N, 2, TPB>(stage, deriv, 0, afrag, wave, lane);  // g, Pi
      face_corrections(&ext_v[0][0]);
      lds_barrier();  // g, Pi derivatives complete
      double dg[P][3], dpi[P][3];
#pragma unroll
      for (int q = 0; q < P; ++q)
        if (act[q]) {
          contract_lds(q, 0, dg[q]);
          contract_lds(q, 1, dpi[q]);
        }
      lds_barrier();  // deriv free again
      mfma_logical_derivatives<N, 3, TPB>(stage, deriv, 2, afrag, wave, lane);  // Phi_n
      lds_barrier();  // Phi derivatives and face corrections complete
#pragma unroll
      for (int q = 0; q < P; ++q) {
        if (!act[q]) continue;
        double dphi[9];  // dphi[m + 3 n] = d_m Phi_n
#pragma unroll
        for (int n = 0; n < 3; ++n) {
          double d3[3];
          contract_lds(q, n, d3);
#pragma unroll
          for (int m = 0; m < 3; ++m) dphi[m + 3 * n] = d3[m];
        }
        gh_point_pair(out[q], st[q], pin[q], own[q][1], own[q] + 2, dg[q], dpi[q], dphi);
      }
#pragma unroll
      for (int q = 0; q < P; ++q) {
        if (!act[q]) continue;
        add_corrections(q, out[q]);
#pragma unroll
        for (int c = 0; c < 5; ++c) dte[comp[c] * NP + t + q * T] = out[q][c];
      }
    }
  } else {
    // pass pp's point-wise outputs plus the corrections of the faces each
    // point lies on (xi, eta, zeta order; GHOST: external faces first), then
    // the dt write
    double prev[P][5];
    auto finish_pass = [&](int pp) {
      const double* cpp = smem + ((pp & 1) ? L::kCorr2 : L::kCorr);
      const int comp[5] = {pp, 10 + pp, 20 + 3 * pp, 21 + 3 * pp, 22 + 3 * pp};
#pragma unroll
      for (int q = 0; q < P; ++q) {
        if (!act[q]) continue;
        const int idx[3] = {pi_[q], pj[q], pk[q]};
#pragma unroll
        for (int step = 0; step < (GHOST ? 6 : 3); ++step) {
          const int d = step % 3;
          if (idx[d] != 0 && idx[d] != N - 1) continue;
          const int side = idx[d] == 0 ? 0 : 1;
          if constexpr (GHOST) {
            if ((((ext_mask >> (2 * d + side)) & 1) != 0) != (step < 3)) continue;
          }
          const int a = d == 0 ? idx[1] : idx[0];
          const int b = d == 2 ? idx[1] : idx[2];
          const double* cr = cpp + (2 * d + side) * NF + a + N * b;
#pragma unroll
          for (int c = 0; c < 5; ++c) prev[q][c] = prev[q][c] + cr[c * L::SLOTS];
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) dte[comp[c] * NP + t + q * T] = prev[q][c];
      }
    };
    if constexpr (L::DMA) {
      // the parked dt Pi partial sums of every point are read by other waves'
      // copies: complete and publish this thread's stores first
      __syncthreads();
      dma_stage(0, smem + L::kStage);
    }
    load_pass(0);
#pragma unroll 1
    for (int p = 0; p < 10; ++p) {
      // Pass p stages into buffer p & 1.  A wave gets here only after the
      // barrier that ends pass p - 1's reads of its stage, so buffer p & 1
      // (last read in pass p - 2) is free.  Pass p's corrections go to
      // correction buffer p & 1; pass p - 1's are read after this pass's
      // staging barrier (which ends their writes), before pass p + 1's
      // barrier lets anyone write that buffer again.  With direct copies,
      // pass p's copy was issued during pass p - 1 (after that pass's
      // barrier, which freed buffer p & 1); each wave waits for its own
      // copies before the barrier that publishes them.
      stage = smem + ((p & 1) ? L::kStage2 : L::kStage);
      corr = smem + ((p & 1) ? L::kCorr2 : L::kCorr);
      if constexpr (!L::DMA) {
#pragma unroll
        for (int q = 0; q < P; ++q)
          if (act[q]) {
#pragma unroll
            for (int c = 0; c < 5; ++c) stage[c * NP + spt[q]] = own_next[q][c];
          }
      }
      double ext_v[PF][5], pin[P];
#pragma unroll
      for (int f = 0; f < PF; ++f)
#pragma unroll
        for (int c = 0; c < 5; ++c) ext_v[f][c] = ext_next[f][c];
      if constexpr (L::DMA) {
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's copies landed
        lds_barrier();
        if (p + 1 < 10) {
          dma_stage(p + 1, smem + (((p + 1) & 1) ? L::kStage2 : L::kStage));
          load_pass(p + 1);
        }
#pragma unroll
        for (int q = 0; q < P; ++q) pin[q] = stage[5 * NP + spt[q]];
      } else {
#pragma unroll
        for (int q = 0; q < P; ++q) pin[q] = pin_next[q];
        if (p + 1 < 10) load_pass(p + 1);
        lds_barrier();
      }
      if (p > 0) finish_pass(p - 1);
      // lifted boundary corrections of the face slots
#pragma unroll
      for (int f = 0; f < PF; ++f) {
        if (!fact[f]) continue;
        const int slot = t + f * T;
        const int dir = slot / NF;
        const int fp = slot % NF;
        const int pv = staged_point<N, L::SWIZZLE>(face_volume_index<N>(dir, fp % N, fp / N));
        double fi[12], fe[11];
#pragma unroll
        for (int x = 0; x < 12; ++x) fi[x] = smem[L::kFaceInt + x * L::SLOTS + slot];
#pragma unroll
        for (int x = 0; x < 11; ++x) fe[x] = smem[L::kFaceExt + x * L::SLOTS + slot];
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
          phi_e[c] = ext_v[f][2 + c];
        }
        gh_package_pair(pki, stage[pv], stage[NP + pv], phi_i, si);
        gh_package_pair(pke, ext_v[f][0], ext_v[f][1], phi_e, se);
        gh_boundary_terms_pair(bc, pki, si.lam, pke, se.lam);
#pragma unroll
        for (int c = 0; c < 5; ++c) corr[c * L::SLOTS + slot] = bc[c] * fac;
      }
      // volume terms of the owned points.  The logical derivatives of all
      // owned points come first: when a thread's two points share their
      // zeta pencil (T a multiple of N^2) it is read from LDS once, and xi
      // pencils (contiguous) are read two doubles at a time for even N.
      constexpr bool kShareZeta = P == 2 && T % (N * N) == 0;
      double dlog[P][15];  // dlog[q][3 c + m] = d_m of staged component c
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        const double* sc = stage + c * NP;
        double zeta[N];
        if constexpr (kShareZeta) {
          const int z0 = staged_i<N, L::SWIZZLE>(pi_[0], pj[0]) + N * pj[0];
#pragma unroll
          for (int l = 0; l < N; ++l) zeta[l] = sc[z0 + N * N * l];
        }
#pragma unroll
        for (int q = 0; q < P; ++q) {
          if (!act[q]) continue;
          const int i = pi_[q], j = pj[q], k = pk[q];
          const double* row = sc + N * (j + N * k);
          double xr[N];
          if constexpr (N % 2 == 0) {
#pragma unroll
            for (int l = 0; l < N; l += 2) {
              const double2 v =
                  *reinterpret_cast<const double2*>(row + staged_i<N, L::SWIZZLE>(l, j));
              xr[l] = v.x;
              xr[l + 1] = v.y;
            }
          } else {
#pragma unroll
            for (int l = 0; l < N; ++l) xr[l] = row[l];
          }
          double d0 = Di[q][0] * xr[0];
#pragma unroll
          for (int l = 1; l < N; ++l) d0 = fma(Di[q][l], xr[l], d0);
          double d1 = Dj[q][0] * sc[staged_i<N, L::SWIZZLE>(i, 0) + N * N * k];
#pragma unroll
          for (int l = 1; l < N; ++l)
            d1 = fma(Dj[q][l], sc[staged_i<N, L::SWIZZLE>(i, l) + N * (l + N * k)], d1);
          double d2;
          if constexpr (kShareZeta) {
            d2 = Dk[q][0] * zeta[0];
#pragma unroll
            for (int l = 1; l < N; ++l) d2 = fma(Dk[q][l], zeta[l], d2);
          } else {
            const int z0 = staged_i<N, L::SWIZZLE>(i, j) + N * j;
            d2 = Dk[q][0] * sc[z0];
#pragma unroll
            for (int l = 1; l < N; ++l) d2 = fma(Dk[q][l], sc[z0 + N * N * l], d2);
          }
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            double s = ij[q][ij3(0, a)] * d0;
            s = fma(ij[q][ij3(1, a)], d1, s);
            s = fma(ij[q][ij3(2, a)], d2, s);
            dlog[q][3 * c + a] = s;
          }
        }
      }
      double out[P][5];
#pragma unroll
      for (int q = 0; q < P; ++q) {
        if (!act[q]) continue;
        const int pt = spt[q];
        // c = 0 g, 1 Pi, 2..4 Phi_n
        double phi[3], dphi[9];
#pragma unroll
        for (int n = 0; n < 3; ++n) {
          phi[n] = stage[(2 + n) * NP + pt];
#pragma unroll
          for (int m = 0; m < 3; ++m) dphi[m + 3 * n] = dlog[q][3 * (2 + n) + m];
        }
        gh_point_pair(out[q], st[q], pin[q], stage[NP + pt], phi, dlog[q], dlog[q] + 3, dphi);
      }
      // this pass's corrections are added after the next pass's staging
      // barrier, which also ends their writes
#pragma unroll
      for (int q = 0; q < P; ++q)
#pragma unroll
        for (int c = 0; c < 5; ++c) prev[q][c] = out[q][c];
    }
    lds_barrier();  // the last pass's corrections complete
    finish_pass(9);
  }
}

template <int N, bool MFMA, int TPB, bool GHOST>
int launch_gh_fused_n(int num_elements, const int32_t* nbr, const double* u,
                      const double* gammas, const double* inv_jac, double* dt,
                      const double* dtab, const double* ghost, hipStream_t stream) {
  hipLaunchKernelGGL((gh_rhs_fused_kernel<N, MFMA, TPB, GHOST>), dim3(num_elements),
                     dim3(GhFusedLayout<N, MFMA, TPB>::THREADS), 0, stream, nbr, u, gammas,
                     inv_jac, dt, dtab, ghost);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

// External faces (ghost != null) are instantiated for the vector-ALU engine.
template <int N, int TPB>
int launch_gh_fused_engine(int variant, int num_elements, const int32_t* nbr, const double* u,
                           const double* gammas, const double* inv_jac, double* dt,
                           const double* dtab, const double* ghost, hipStream_t stream) {
  if (ghost != nullptr) {
    if (variant == kGhFusedMfma) return SXD_ERR_UNSUPPORTED;
    return launch_gh_fused_n<N, false, TPB, true>(num_elements, nbr, u, gammas, inv_jac, dt,
                                                  dtab, ghost, stream);
  }
  if (variant == kGhFusedMfma)
    return launch_gh_fused_n<N, true, TPB, false>(num_elements, nbr, u, gammas, inv_jac, dt,
                                                  dtab, nullptr, stream);
  return launch_gh_fused_n<N, false, TPB, false>(num_elements, nbr, u, gammas, inv_jac, dt,
                                                 dtab, nullptr, stream);
}

// Below 7^3 points every tile has at most 256 threads; 512 only differs from
// 256 at N = 7, 8.
template <int N>
int launch_gh_fused_variant(int variant, int threads, int num_elements, const int32_t* nbr,
                            const double* u, const double* gammas, const double* inv_jac,
                            double* dt, const double* dtab, const double* ghost,
                            hipStream_t stream) {
  if constexpr (N >= 7) {
    if (threads == 512)
      return launch_gh_fused_engine<N, 512>(variant, num_elements, nbr, u, gammas, inv_jac, dt,
                                            dtab, ghost, stream);
  }
  return launch_gh_fused_engine<N, 256>(variant, num_elements, nbr, u, gammas, inv_jac, dt, dtab,
                                        ghost, stream);
}

}  // namespace

bool gh_fused_supported(const int32_t ext[3]) {
  return ext[0] == ext[1] && ext[1] == ext[2] && ext[0] >= 2 && ext[0] <= 8;
}

int launch_gh_rhs_fused(int num_elements, const int32_t ext[3], const int32_t* nbr,
                        const double* u, const double* gammas, const double* inv_jac,
                        double* dt, int variant, int threads, const double* ghost,
                        hipStream_t stream) {
  const double* dtab = device_diff_matrices();
  if (dtab == nullptr) return SXD_ERR_DEVICE;
  if (!gh_fused_supported(ext)) return SXD_ERR_UNSUPPORTED;
  if (variant != kGhFusedValu && variant != kGhFusedMfma) return SXD_ERR_INVALID_ARGUMENT;
  switch (ext[0]) {
#define SXD_GH_FUSED_CASE(n) \
  case n:                    \
    return launch_gh_fused_variant<n>(variant, threads, num_elements, nbr, u, gammas, inv_jac, dt, \
                                      dtab, ghost, stream);
    SXD_GH_FUSED_CASE(2)
    SXD_GH_FUSED_CASE(3)
    SXD_GH_FUSED_CASE(4)
    SXD_GH_FUSED_CASE(5)
    SXD_GH_FUSED_CASE(6)
    SXD_GH_FUSED_CASE(7)
    SXD_GH_FUSED_CASE(8)
#undef SXD_GH_FUSED_CASE
    default: return SXD_ERR_UNSUPPORTED;
  }
}

}  // namespace sxd
