"""Write input/output vectors of the GH constraint-preserving Bjorhus boundary
condition (ConstraintPreserving and ConstraintPreservingPhysical types, static
mesh) from the reference's own Python test oracle.

The reference checks its C++ ConstraintPreservingBjorhus against the numpy
functions of
  tests/Unit/Evolution/Systems/GeneralizedHarmonic/BoundaryConditions/Bjorhus.py
(dt_spacetime_metric_static_mesh, dt_pi_ConstraintPreserving_static_mesh,
dt_phi_ConstraintPreserving_static_mesh and the ..._ConstraintPreservingPhysical_
static_mesh pair), which evaluate the corrections at one point.  No closed form pins those values, so this script (run only where
the reference tree exists) imports that module with its sibling test modules,
evaluates it on seeded, physically consistent random points (a spacetime
metric near Minkowski with its lapse, shift, inverse and normal vector; a
unit normal covector; random fields, derivatives, time derivatives, gauge
source and constraint; coordinates at radius 3..10; shifts along +-n so that
both signs of every characteristic speed occur, and every fourth point has no
incoming speed at all) and stores inputs and outputs as JSON.  Nothing of the
reference is copied: the fixture holds numbers only.

Usage: python tests/golden/make_bjorhus_vectors.py [reference_root]
"""
import json
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent


def sym(rng, shape, scale):
    x = rng.uniform(-scale, scale, shape)
    return 0.5 * (x + np.swapaxes(x, -1, -2))


def point(rng, k):
    g = np.diag([-1.0, 1.0, 1.0, 1.0]) + sym(rng, (4, 4), 0.15)
    inv_g = np.linalg.inv(g)
    lapse = 1.0 / np.sqrt(-inv_g[0, 0])
    shift = inv_g[0, 1:] * lapse ** 2
    spatial_inv = inv_g[1:, 1:] + np.outer(shift, shift) / lapse ** 2
    n = rng.normal(size=3)
    n /= np.sqrt(n @ spatial_inv @ n)
    # steer the shift along +-n so that the speeds change sign between cases
    n_up = spatial_inv @ n
    push = (0.6, -0.6, 0.0, -1.6)[k % 4] * lapse  # the last: every speed positive
    shift = shift + push * n_up
    # rebuild g consistently from (lapse, shift, spatial metric)
    gamma = np.linalg.inv(spatial_inv)
    beta_dn = gamma @ shift
    g = np.zeros((4, 4))
    g[0, 0] = -lapse ** 2 + shift @ beta_dn
    g[0, 1:] = g[1:, 0] = beta_dn
    g[1:, 1:] = gamma
    inv_g = np.linalg.inv(g)
    t_up = np.concatenate([[1.0 / lapse], -shift / lapse])
    r = rng.uniform(3.0, 10.0)
    x = rng.normal(size=3)
    x *= r / np.linalg.norm(x)
    return dict(
        normal_covector=n, normal_vector=n_up, spacetime_metric=g,
        pi=sym(rng, (4, 4), 0.5), phi=sym(rng, (3, 4, 4), 0.5), inertial_coords=x,
        gamma1=rng.uniform(-1.0, 0.5), gamma2=rng.uniform(0.1, 2.0), lapse=lapse, shift=shift,
        inverse_spacetime_metric=inv_g, spacetime_unit_normal_vector=t_up,
        three_index_constraint=sym(rng, (3, 4, 4), 0.2), gauge_source=rng.uniform(-0.3, 0.3, 4),
        spacetime_deriv_gauge_source=rng.uniform(-0.3, 0.3, (4, 4)),
        dt_spacetime_metric=sym(rng, (4, 4), 1.0), dt_pi=sym(rng, (4, 4), 1.0),
        dt_phi=sym(rng, (3, 4, 4), 1.0), d_spacetime_metric=sym(rng, (3, 4, 4), 0.5),
        d_pi=sym(rng, (3, 4, 4), 0.5), d_phi=sym(rng, (3, 3, 4, 4), 0.5))


ORDER = ["normal_covector", "normal_vector", "spacetime_metric", "pi", "phi", "inertial_coords",
         "gamma1", "gamma2", "lapse", "shift", "inverse_spacetime_metric",
         "spacetime_unit_normal_vector", "three_index_constraint", "gauge_source",
         "spacetime_deriv_gauge_source", "dt_spacetime_metric", "dt_pi", "dt_phi",
         "d_spacetime_metric", "d_pi", "d_phi"]


def main():
    root = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    sys.path.insert(0, str(root / "tests" / "Unit"))
    import Evolution.Systems.GeneralizedHarmonic.BoundaryConditions.Bjorhus as bj

    rng = np.random.default_rng(2024)
    cases = []
    for k in range(48):
        p = point(rng, k)
        args = [p[name] for name in ORDER]
        dt_g = bj.dt_spacetime_metric_static_mesh(*args)
        dt_pi = bj.dt_pi_ConstraintPreserving_static_mesh(*args)
        dt_phi = bj.dt_phi_ConstraintPreserving_static_mesh(*args)
        dt_pi_phys = bj.dt_pi_ConstraintPreservingPhysical_static_mesh(*args)
        dt_phi_phys = bj.dt_phi_ConstraintPreservingPhysical_static_mesh(*args)
        cases.append({"inputs": {k2: np.asarray(v).tolist() for k2, v in p.items()},
                      "dt_spacetime_metric": np.asarray(dt_g).tolist(),
                      "dt_pi": np.asarray(dt_pi).tolist(),
                      "dt_phi": np.asarray(dt_phi).tolist(),
                      "dt_pi_physical": np.asarray(dt_pi_phys).tolist(),
                      "dt_phi_physical": np.asarray(dt_phi_phys).tolist()})
    out = HERE / "bjorhus_vectors.json"
    out.write_text(json.dumps({"source": "tests/Unit/Evolution/Systems/GeneralizedHarmonic/"
                                         "BoundaryConditions/Bjorhus.py (ConstraintPreserving "
                                         "and ConstraintPreservingPhysical, static mesh)", "cases": cases}))
    print(f"wrote {len(cases)} cases to {out}")


if __name__ == "__main__":
    main()
