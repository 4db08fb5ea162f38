// Shared device-side definitions for the CDNA4 kernels.
//
// Floating point: everything is FP64.  Kernels are compiled with
// -ffp-contract=off; every accumulation statement s += a * b of the
// reference is written as an explicit fma(a, b, s), everything else is
// evaluated unfused in the reference's order.  This makes the kernels
// reproduce the reference's arithmetic step for step.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "spectral.hpp"

namespace sxd {

// Device copy of the packed LGL differentiation matrices, n = 2..20
// (row-major, see spectral.hpp); kernels take its address as an argument.
__device__ __forceinline__ const double* diff_matrix(const double* table, int n) {
  return table + packed_d_offset(n);
}

// Largest LDS copy of the three D matrices of an element (extents <= 20).
constexpr int kStagedDiffDoubles = 3 * 20 * 20;

// Copies the D matrices of extents n[0..2] from the device table into LDS at
// dst (n0^2 + n1^2 + n2^2 doubles, at most kStagedDiffDoubles) and points
// D[0..2] at the copies.  The caller synchronises the workgroup before use.
__device__ __forceinline__ void stage_diff_matrices(const double* table, const int n[3],
                                                    double* dst, const double* D[3]) {
  int off = 0;
  for (int d = 0; d < 3; ++d) {
    const double* src = diff_matrix(table, n[d]);
    for (int x = threadIdx.x; x < n[d] * n[d]; x += blockDim.x) dst[off + x] = src[x];
    D[d] = dst + off;
    off += n[d] * n[d];
  }
}

// Offset of the (nf, nm) block pair in the packed mortar table (the same
// value as packed_mortar_offset, in O(nf) steps).
__host__ __device__ inline int mortar_table_offset(int nf, int nm) {
  constexpr int kSumB = kMaxPoints * (kMaxPoints + 1) / 2;  // sum_{b=1}^{20} b
  int off = 0;
  for (int a = kMinPoints; a < nf; ++a) off += 2 * a * (kSumB - a * (a - 1) / 2);
  return off + 2 * nf * ((nm - 1) * nm / 2 - (nf - 1) * nf / 2);
}

// Tensor storage order (DataStructures/Tensor/Structure.hpp:98-195).
// Symmetric 4-d pair (a <= b): (0,0)(0,1)(0,2)(0,3)(1,1)(1,2)(1,3)(2,2)(2,3)(3,3)
__host__ __device__ constexpr int sym4(int a, int b) {
  return a <= b ? (a == 0 ? b : a == 1 ? 3 + b : a == 2 ? 5 + b : 9)
                : (b == 0 ? a : b == 1 ? 3 + a : b == 2 ? 5 + a : 9);
}
// Symmetric 3-d pair: (0,0)(0,1)(0,2)(1,1)(1,2)(2,2)
__host__ __device__ constexpr int sym3(int a, int b) {
  return a <= b ? (a == 0 ? b : a == 1 ? 2 + b : 5) : (b == 0 ? a : b == 1 ? 2 + a : 5);
}
// First (row) index of each symmetric 4-d pair in storage order.
__host__ __device__ constexpr int pair_a(int p) {
  return p < 4 ? 0 : p < 7 ? 1 : p < 9 ? 2 : 3;
}
__host__ __device__ constexpr int pair_b(int p) {
  return p < 4 ? p : p < 7 ? p - 3 : p < 9 ? p - 5 : 3;
}
// InverseJacobian(logical j, inertial i) block.
__host__ __device__ constexpr int ij3(int j, int i) { return j + 3 * i; }

// Workgroup barrier that orders LDS accesses only.  __syncthreads() also
// releases global memory, which makes every wave wait for its outstanding
// global loads (s_waitcnt vmcnt(0)) at the barrier; with a local-address-space
// fence, prefetches of the next pass stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Element handled by workgroup b of a one-workgroup-per-element launch of E
// workgroups.  The dispatcher hands workgroups to the 8 XCDs of the chip in
// turn (b % 8), and each XCD has its own L2; this bijection gives XCD x one
// contiguous run of about E / 8 elements, so that an element and the
// neighbours whose face traces it reads (numbered next to it) share an L2.
constexpr int kNumXcds = 8;
__device__ __forceinline__ int64_t xcd_local_element(int64_t b, int64_t num_elements) {
  const int64_t x = b % kNumXcds, slot = b / kNumXcds;
  const int64_t per = num_elements / kNumXcds, rem = num_elements % kNumXcds;
  return x * per + (x < rem ? x : rem) + slot;
}

// Heaviside step of UpwindPenalty (Blaze/StepFunction.hpp:46-48).
__device__ __forceinline__ double step_fn(double x) { return x < 0.0 ? 0.0 : 1.0; }

// Lift factor of dg::lift_flux (LiftFlux.hpp:56-60): -0.5 N (N-1) |n|.
__device__ __forceinline__ double lift_factor(int n_perp, double mag) {
  return (-0.5 * static_cast<double>(n_perp * (n_perp - 1))) * mag;
}

// Device address of the packed D tables on the current device, uploading
// them on first use (one small allocation per device for the life of the
// process).  Returns nullptr on failure.
const double* device_diff_matrices();

}  // namespace sxd
