// HBM probe kernels of known access pattern, for calibrating the rocprofv3
// byte counters (FETCH_SIZE, WRITE_SIZE) and for a measured copy bandwidth
// beside the 8 TB/s data-sheet figure.  Not part of the DG path: built into
// its own library (libsxd_hbm_probe.so) and used by the benchmark only.
//
// Every kernel walks its arrays once, with 16-byte loads/stores that are
// contiguous across a wavefront (1 KiB per wavefront instruction), in a
// grid-stride loop: the algorithmic bytes are exactly
//   read:  n doubles read, one double written per workgroup
//   write: n doubles written, nothing read
//   copy:  n doubles read and n written
// and two reads with the access patterns of the DG kernels' loads:
//   read8:    n doubles read with 8-byte loads (contiguous across a
//             wavefront, 512 bytes per instruction)
//   gather64: n / 8 doubles read, one per 64-byte line (stride 8 doubles, as
//             the xi-face traces of an 8^3 element)
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kThreads = 256;

__global__ void __launch_bounds__(kThreads)
probe_read_kernel(const double2* __restrict__ src, int64_t n2, double* __restrict__ partial) {
  __shared__ double s[kThreads];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kThreads) + threadIdx.x; i < n2;
       i += static_cast<int64_t>(gridDim.x) * kThreads) {
    const double2 v = src[i];
    acc += v.x + v.y;
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w /= 2) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(kThreads)
probe_write_kernel(double2* __restrict__ dst, int64_t n2, double value) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kThreads) + threadIdx.x; i < n2;
       i += static_cast<int64_t>(gridDim.x) * kThreads)
    dst[i] = make_double2(value, value);
}

__global__ void __launch_bounds__(kThreads)
probe_copy_kernel(const double2* __restrict__ src, double2* __restrict__ dst, int64_t n2) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kThreads) + threadIdx.x; i < n2;
       i += static_cast<int64_t>(gridDim.x) * kThreads)
    dst[i] = src[i];
}

__global__ void __launch_bounds__(kThreads)
probe_read8_kernel(const double* __restrict__ src, int64_t n, int64_t stride,
                   double* __restrict__ partial) {
  __shared__ double s[kThreads];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kThreads) + threadIdx.x; i * stride < n;
       i += static_cast<int64_t>(gridDim.x) * kThreads)
    acc += src[i * stride];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w /= 2) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

}  // namespace

extern "C" {

// Number of workgroups the probes launch (and doubles `partial` must hold).
int sxd_probe_workgroups(void) { return 256 * 8; }

// mode 0: read n doubles of src (partial gets sxd_probe_workgroups() sums);
// mode 1: write n doubles of dst; mode 2: copy src to dst; mode 3: read n
// doubles of src with 8-byte loads; mode 4: read every 8th double of the
// first n of src (8-byte loads, 64 bytes apart).  n must be even
// and the arrays 16-byte aligned device memory.  Returns 0, or -1 on a bad
// argument or launch failure.
int sxd_probe_run(int mode, const double* src, double* dst, double* partial, int64_t n,
                  void* stream) {
  if (n < 0 || (n & 1) != 0) return -1;
  if ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15) return -1;
  const int64_t n2 = n / 2;
  const dim3 grid(sxd_probe_workgroups()), block(kThreads);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (mode == 0) {
    if (src == nullptr || partial == nullptr) return -1;
    hipLaunchKernelGGL(probe_read_kernel, grid, block, 0, s,
                       reinterpret_cast<const double2*>(src), n2, partial);
  } else if (mode == 1) {
    if (dst == nullptr) return -1;
    hipLaunchKernelGGL(probe_write_kernel, grid, block, 0, s, reinterpret_cast<double2*>(dst),
                       n2, 1.0);
  } else if (mode == 2) {
    if (src == nullptr || dst == nullptr) return -1;
    hipLaunchKernelGGL(probe_copy_kernel, grid, block, 0, s,
                       reinterpret_cast<const double2*>(src), reinterpret_cast<double2*>(dst),
                       n2);
  } else if (mode == 3 || mode == 4) {
    if (src == nullptr || partial == nullptr) return -1;
    hipLaunchKernelGGL(probe_read8_kernel, grid, block, 0, s, src, n, mode == 3 ? 1 : 8,
                       partial);
  } else {
    return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"
