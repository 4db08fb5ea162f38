"""spectre_amd: SpECTRE's DG volume time-derivative path on MI355X (gfx950).

Native HIP kernels behind a C ABI (include/spectre_amd.h, libspectre_amd.so);
this package is the Python host layer over them.  Modules:

* spectral   - LGL points, weights, differentiation matrices
* dg         - device entry points on torch tensors (partial derivatives,
               ScalarWave / GH time derivatives, UpwindPenalty, lift, RK3,
               whole-domain RHS)
* workloads  - periodic bricks and analytic initial data (PlaneWave,
               Kerr-Schild) for tests and benchmarks
"""
from ._lib import SpectreAmdError, lib  # noqa: F401

__version__ = "0.1.0"
