// LGL spectral tables (host) and their packed device layout.
#pragma once

#include <vector>

namespace sxd {

constexpr int kMinPoints = 2;
constexpr int kMaxPoints = 20;

// Offset of the n-point matrix in the packed table: sum_{m=2}^{n-1} m^2.
constexpr int packed_d_offset(int n) {
  int off = 0;
  for (int m = kMinPoints; m < n; ++m) off += m * m;
  return off;
}
constexpr int kPackedDSize = packed_d_offset(kMaxPoints + 1);

void lgl_points_and_weights(int n, double* x, double* w);
void lgl_differentiation_matrix(int n, double* d);  // row-major
// All D matrices for n = 2..20, row-major, at packed_d_offset(n).
const std::vector<double>& packed_differentiation_matrices();

}  // namespace sxd
