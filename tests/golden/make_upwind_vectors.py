"""Write UpwindPenalty input/output vectors from the reference's own Python
test oracles.

The reference checks its C++ UpwindPenalty (ScalarWave and generalized
harmonic) against pure-numpy implementations used by its pypp tests:
  tests/Unit/Evolution/Systems/ScalarWave/BoundaryCorrections/UpwindPenalty.py
  tests/Unit/Evolution/Systems/GeneralizedHarmonic/BoundaryCorrections/UpwindPenalty.py
with random inputs in [-1, 1) (tests/Unit/Helpers/Evolution/DiscontinuousGalerkin/
BoundaryCorrections.hpp:836-914).  No closed form pins those random-input
values, so this script (run only where the reference tree exists) imports
those two modules, evaluates dg_package_data on both sides and
dg_boundary_terms, and stores inputs and outputs as JSON.  Nothing of the
reference is copied: the fixture holds numbers only.

Usage: python tests/golden/make_upwind_vectors.py [reference_root]
"""
import importlib.util
import json
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent


def load(path, name):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def sym4(a, b):
    a, b = min(a, b), max(a, b)
    return (0, 4, 7, 9)[a] + b - a


def sw_vectors(mod, rng, n):
    cases = []
    for _ in range(n):
        side = {}
        for s in ("int", "ext"):
            psi, pi = rng.uniform(-1, 1), rng.uniform(-1, 1)
            phi = rng.uniform(-1, 1, 3)
            g2 = rng.uniform(-1, 1)
            nrm = rng.uniform(-1, 1, 3)
            nrm /= np.linalg.norm(nrm)
            side[s] = dict(psi=psi, pi=pi, phi=phi.tolist(), gamma2=g2, normal=nrm.tolist())
        pk = {s: mod.dg_package_data(v["psi"], v["pi"], np.array(v["phi"]), v["gamma2"],
                                     np.array(v["normal"]), None, None)
              for s, v in side.items()}
        bc = mod.dg_boundary_terms(*pk["int"], *pk["ext"], False)
        cases.append(dict(inputs=side, bc_psi=float(bc[0]), bc_pi=float(bc[1]),
                          bc_phi=np.asarray(bc[2]).tolist()))
    return cases


def gh_vectors(mod, rng, n):
    cases = []
    for _ in range(n):
        side = {}
        for s in ("int", "ext"):
            g = rng.uniform(-1, 1, (4, 4))
            g = 0.5 * (g + g.T)
            pi = rng.uniform(-1, 1, (4, 4))
            pi = 0.5 * (pi + pi.T)
            phi = rng.uniform(-1, 1, (3, 4, 4))
            phi = 0.5 * (phi + phi.transpose(0, 2, 1))
            gamma1, gamma2 = rng.uniform(-1, 1), rng.uniform(-1, 1)
            lapse = rng.uniform(0.5, 1.5)
            shift = rng.uniform(-1, 1, 3)
            # near-flat inverse spatial metric, unit normal with respect to it
            a = rng.uniform(-0.1, 0.1, (3, 3))
            inv = np.eye(3) + 0.5 * (a + a.T)
            ncov = rng.uniform(-1, 1, 3)
            nvec = inv @ ncov
            mag = np.sqrt(nvec @ ncov)
            ncov, nvec = ncov / mag, nvec / mag
            side[s] = dict(g=g.tolist(), pi=pi.tolist(), phi=phi.tolist(), gamma1=gamma1,
                           gamma2=gamma2, lapse=lapse, shift=shift.tolist(),
                           normal_covector=ncov.tolist(), normal_vector=nvec.tolist())
        pk = {s: mod.dg_package_data(np.array(v["g"]), np.array(v["pi"]), np.array(v["phi"]),
                                     v["gamma1"], v["gamma2"], v["lapse"], np.array(v["shift"]),
                                     np.array(v["normal_covector"]), np.array(v["normal_vector"]),
                                     None, None)
              for s, v in side.items()}
        bc = mod.dg_boundary_terms(*pk["int"], *pk["ext"], False)
        cases.append(dict(inputs=side, bc_g=np.asarray(bc[0]).tolist(),
                          bc_pi=np.asarray(bc[1]).tolist(), bc_phi=np.asarray(bc[2]).tolist()))
    return cases


def main():
    ref = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    base = ref / "tests/Unit/Evolution/Systems"
    sw = load(base / "ScalarWave/BoundaryCorrections/UpwindPenalty.py", "sw_upwind_ref")
    gh = load(base / "GeneralizedHarmonic/BoundaryCorrections/UpwindPenalty.py", "gh_upwind_ref")
    rng = np.random.default_rng(20241129)
    out = {"source": ["tests/Unit/Evolution/Systems/ScalarWave/BoundaryCorrections/UpwindPenalty.py",
                      "tests/Unit/Evolution/Systems/GeneralizedHarmonic/BoundaryCorrections/UpwindPenalty.py"],
           "note": "static mesh (no mesh velocity); inputs in [-1, 1)",
           "scalar_wave": sw_vectors(sw, rng, 40),
           "generalized_harmonic": gh_vectors(gh, rng, 40)}
    (HERE / "upwind_penalty_vectors.json").write_text(json.dumps(out, indent=0))
    print("wrote", len(out["scalar_wave"]), "SW and", len(out["generalized_harmonic"]), "GH cases")


if __name__ == "__main__":
    main()
