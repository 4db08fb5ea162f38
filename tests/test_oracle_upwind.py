"""Pins of the oracle's UpwindPenalty (ScalarWave and GH) against vectors
written from the reference's own Python test oracles
(tests/golden/make_upwind_vectors.py; the reference's C++ tests compare against
the same functions with Catch approx, 1e-14 relative, scale 1), and the
invariant that continuous data across a face gives no correction.
"""
import json
import pathlib

import numpy as np

import oracle_lib as ol
from spectre_amd import workloads as wl

VEC = json.loads((pathlib.Path(__file__).parent / "golden" / "upwind_penalty_vectors.json").read_text())


def approx_equal(a, b, eps=1e-14):
    a, b = np.asarray(a), np.asarray(b)
    return np.all(np.abs(a - b) <= eps * (1.0 + np.maximum(np.abs(a), np.abs(b))))


def test_scalar_wave_upwind_matches_reference_vectors():
    cases = VEC["scalar_wave"]
    n = len(cases)
    arr = {s: {"u": np.zeros((5, n)), "g2": np.zeros(n), "n": np.zeros((3, n))}
           for s in ("int", "ext")}
    for k, c in enumerate(cases):
        for s in ("int", "ext"):
            v = c["inputs"][s]
            arr[s]["u"][:, k] = [v["psi"], v["pi"], *v["phi"]]
            arr[s]["g2"][k] = v["gamma2"]
            arr[s]["n"][:, k] = v["normal"]
    bc = ol.sw_upwind(arr["int"]["u"], arr["int"]["g2"], arr["int"]["n"],
                      arr["ext"]["u"], arr["ext"]["g2"], arr["ext"]["n"])
    for k, c in enumerate(cases):
        assert approx_equal(bc[0, k], c["bc_psi"]), k
        assert approx_equal(bc[1, k], c["bc_pi"]), k
        assert approx_equal(bc[2:, k], c["bc_phi"]), k


def test_gh_upwind_matches_reference_vectors():
    cases = VEC["generalized_harmonic"]
    n = len(cases)
    arr = {s: {"u": np.zeros((50, n)), "aux": np.zeros((6, n)), "nrm": np.zeros((6, n))}
           for s in ("int", "ext")}
    for k, c in enumerate(cases):
        for s in ("int", "ext"):
            v = c["inputs"][s]
            g, pi, phi = np.array(v["g"]), np.array(v["pi"]), np.array(v["phi"])
            for (a, b) in wl.PAIRS:
                p = wl.sym4(a, b)
                arr[s]["u"][p, k] = g[a, b]
                arr[s]["u"][10 + p, k] = pi[a, b]
                for i in range(3):
                    arr[s]["u"][20 + i + 3 * p, k] = phi[i, a, b]
            arr[s]["aux"][:, k] = [v["gamma1"], v["gamma2"], v["lapse"], *v["shift"]]
            arr[s]["nrm"][:, k] = [*v["normal_covector"], *v["normal_vector"]]
    bc = ol.gh_upwind_direct(arr["int"]["u"], arr["int"]["aux"], arr["int"]["nrm"],
                             arr["ext"]["u"], arr["ext"]["aux"], arr["ext"]["nrm"])
    for k, c in enumerate(cases):
        bg, bpi, bphi = np.array(c["bc_g"]), np.array(c["bc_pi"]), np.array(c["bc_phi"])
        for (a, b) in wl.PAIRS:
            p = wl.sym4(a, b)
            assert approx_equal(bc[p, k], bg[a, b]), (k, a, b)
            assert approx_equal(bc[10 + p, k], bpi[a, b]), (k, a, b)
            for i in range(3):
                assert approx_equal(bc[20 + i + 3 * p, k], bphi[i, a, b]), (k, i, a, b)


def test_upwind_zero_for_continuous_states():
    rng = np.random.default_rng(3)
    n = 50
    u = rng.uniform(-1, 1, (5, n))
    g2 = rng.uniform(0, 1, n)
    nrm = rng.normal(size=(3, n))
    nrm /= np.linalg.norm(nrm, axis=0)
    bc = ol.sw_upwind(u, g2, nrm, u, g2, -nrm)
    assert np.max(np.abs(bc)) < 1e-15
