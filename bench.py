#!/usr/bin/env python3
"""Benchmark of the DG right-hand side on one MI355X.

A step is one full DG time derivative (volume terms + UpwindPenalty boundary
corrections + LGL lift) of every element of the workload, with all inputs
resident in HBM.  Default workload: generalized harmonic, 512 elements
(8x8x8 periodic brick, x in [3,11], y, z in [-4,4]), 8^3 LGL points each,
50 evolved components, Kerr-Schild (M = 1, chi = 0.5) plus 1e-6 relative
noise, gamma0 = gamma2 = 1, gamma1 = -1, harmonic gauge.  FP64 throughout.

Prints one JSON line: metric, value (grid points per second), roofline of the
RHS kernel against HBM bandwidth (algorithmic bytes / measured time) and a CPU
baseline (the oracle, a C++ restatement of the reference path, timed on the
host cores over a bounded sample).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload gh|sw]
"""
import argparse
import json
import os
import sys
import time

METRIC = "DG RHS grid-pts/s FP64 on 1 GPU, % roofline; mixed-p batch gain; adjoint/forward cost"
HBM_PEAK_GBS = 8000.0
LLC_BYTES = 256 << 20  # MI355X Infinity Cache
FP64_PEAK_TFLOPS = 78.6

# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
# (each scaled by what the counter reads for probe kernels of known access
# pattern: FETCH_SIZE x 2.0, WRITE_SIZE x 1.0), for the workload's default
# kernel; the summaries are profiles/gh_rhs_fused_kernel_gh_kerr_schild_512el_8cubed.md,
# profiles/hbm_probe_counter_calibration.md and
# profiles/sw_rhs_fused_kernel_sw_plane_wave_64el_5cubed.md.
MEASURED_TRAFFIC = {("gh", "fused"): 450124288.0, ("scalar_wave", "fused"): 2243413.0}


def algorithmic_bytes_per_point(nc, n_aux, n):
    """Compulsory HBM traffic per grid point: u (nc), inverse Jacobian (9) and
    damping parameters (n_aux) read once, dt (nc) written once, plus the
    neighbour's face trace of u (nc per face point, 6 N^2 face points per N^3
    volume points)."""
    return 8.0 * (nc + 9 + n_aux + nc) + 8.0 * nc * (6.0 * n * n / n ** 3)


def mixed_algorithmic_bytes(brick, nc, n_aux):
    """Compulsory HBM bytes of one mixed-p RHS: every element's fields,
    inverse Jacobian and damping parameters read once, dt written once, plus
    the neighbour face traces each face reads (nc values per neighbour face
    point)."""
    import numpy as np
    total = 8.0 * (2 * nc + 9 + n_aux) * brick.total_points
    for e in range(brick.num_elements):
        for d in range(3):
            for side in range(2):
                en = brick.neighbors[e, 2 * d + side]
                ext = brick.extents[en]
                total += 8.0 * nc * int(np.prod(ext)) // int(ext[d])
    return total


def time_calls(fn, steps, warmup):
    """Mean ms per call with HIP events on the launch stream."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn()
    stop.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def hbm_probe(mib, reps=5):
    """Probe kernels of known access pattern (spectre_amd/csrc/hbm_probe.hip):
    a streaming read, a streaming write and a copy of `mib` MiB each, a read
    with 8-byte loads and a gather of one double per 64-byte line, `reps`
    launches each on the current stream, timed with HIP events.  Profiled with
    rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE), the counters' KiB per
    dispatch against the known bytes give the factor that converts the
    counters into bytes; the copy rate is a measured HBM bandwidth."""
    import ctypes
    import torch
    root = os.path.dirname(os.path.abspath(__file__))
    probe = ctypes.CDLL(os.path.join(root, "spectre_amd", "libsxd_hbm_probe.so"))
    probe.sxd_probe_run.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    partial = torch.empty(probe.sxd_probe_workgroups(), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    out = {"probe_doubles": n}
    for mode, name, read, written in ((0, "read", n * 8, partial.numel() * 8),
                                      (1, "write", 0, n * 8), (2, "copy", n * 8, n * 8),
                                      (3, "read8", n * 8, partial.numel() * 8),
                                      (4, "gather64", n, partial.numel() * 8)):
        call = lambda: probe.sxd_probe_run(mode, ctypes.c_void_p(src.data_ptr()),
                                           ctypes.c_void_p(dst.data_ptr()),
                                           ctypes.c_void_p(partial.data_ptr()), n,
                                           ctypes.c_void_p(stream.cuda_stream))
        if call() != 0:
            raise RuntimeError(f"probe {name} failed")
        ms = time_calls(lambda: call(), reps, 1)
        out[name] = {"bytes_read": read, "bytes_written": written, "ms": round(ms, 5),
                     "gb_per_s": round((read + written) / (ms * 1e-3) / 1e9, 1)}
    torch.cuda.synchronize()
    return out


def measure_sw_graph(calls=1000, reps=5):
    """ScalarWave plane wave on 64 elements of 5^3 points (8000 points): one
    launch is far too small to fill the GPU, so the RHS is also timed inside a
    HIP graph of `calls` RHS evaluations (the library's launches captured by
    torch.cuda.graph on the capture stream).  The per-call difference between
    call-by-call enqueueing and graph replay is the launch overhead."""
    import numpy as np
    import torch
    from spectre_amd import dg, workloads as wl

    brick, u, g2, inv = wl.scalar_wave_plane_wave_3d(seed=0)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    d_u, d_g, d_inv = dev(u), dev(g2), dev(inv)
    d_dt = torch.empty_like(d_u)
    domain = dg.Domain.from_brick(brick)
    domain.set_kernel("scalar_wave", "fused")
    rhs = lambda: domain.scalar_wave_rhs(d_u, d_g, d_inv, out=d_dt)
    direct_ms = time_calls(rhs, calls, 50)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        for _ in range(calls):
            rhs()
    graph.replay()
    torch.cuda.synchronize()
    graph_ms = time_calls(graph.replay, reps, 1) / calls
    points = brick.num_elements * brick.num_points
    return {"workload": "scalar_wave_plane_wave_64el_5^3", "grid_points": points,
            "graph_calls": calls, "rhs_us_graph": round(graph_ms * 1e3, 3),
            "rhs_us_call_by_call": round(direct_ms * 1e3, 3),
            "launch_overhead_us": round((direct_ms - graph_ms) * 1e3, 3),
            "grid_points_per_s_graph": round(points / (graph_ms * 1e-3), 1)}


def measure_mixed(system, steps, warmup, rk3_steps=5):
    """Mixed-p workload (1000 elements of a 10x10x10 brick, extents drawn from
    {4..12} per element and dimension): the two batched organisations over
    the whole domain (the mortar kernels, and the mixed-p face kernels)
    against one launch per group of equal extents, with the autotuner's
    threads per workgroup, and one RK3 step replayed from a HIP graph (the
    autotuned organisation) against the same step enqueued call by call."""
    import numpy as np
    import torch
    from spectre_amd import dg, workloads as wl

    if system == "scalar_wave":
        brick, u, params, inv = wl.scalar_wave_plane_wave_mixed(seed=0)
        nc, n_aux = 5, 1
        rhs_name = "scalar_wave_rhs"
    else:
        brick, u, params, inv = wl.gh_kerr_schild_mixed(seed=0)
        nc, n_aux = 50, 3
        rhs_name = "gh_rhs"
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    d_u, d_p, d_inv = dev(u), dev(params), dev(inv)
    d_dt = torch.empty_like(d_u)
    domain = dg.MixedDomain.from_brick(brick)
    rhs = getattr(domain, rhs_name)
    chosen, _ = domain.autotune(system, d_u, d_p, d_inv, d_dt, repetitions=5)
    threads = domain.threads(system)
    ms = {}
    for kernel in ("mortar", "mixed_batched", "mixed_per_extents"):
        domain.set_kernel(system, kernel)
        ms[kernel] = time_calls(lambda: rhs(d_u, d_p, d_inv, out=d_dt), steps, warmup)
    domain.set_kernel(system, chosen)
    # one RK3 step: graph replay against call-by-call enqueueing
    dt_step = 1e-3
    d_u2 = d_u.clone()
    stepper = dg.Rk3Stepper(domain, system, d_u2, d_p, d_inv)
    graph_ms = time_calls(lambda: stepper.step(dt_step, 1), rk3_steps, 1)
    stepper.close()
    stage, l = torch.empty_like(d_u), torch.empty_like(d_u)

    def eager_step():
        rhs(d_u2, d_p, d_inv, out=l)
        dg.rk3_substep(0, dt_step, d_u2, d_u2, l, out=stage)
        rhs(stage, d_p, d_inv, out=l)
        dg.rk3_substep(1, dt_step, d_u2, stage, l, out=stage)
        rhs(stage, d_p, d_inv, out=l)
        dg.rk3_substep(2, dt_step, d_u2, stage, l, out=d_u2)

    eager_ms = time_calls(eager_step, rk3_steps, 1)
    points = brick.total_points
    alg = mixed_algorithmic_bytes(brick, nc, n_aux)
    b = min(ms["mortar"], ms["mixed_batched"])
    return {
        "workload": f"{system}_mixed_p_1000el_extents_4..12",
        "grid_points": points,
        "extents_groups": len({tuple(e) for e in brick.extents.tolist()}),
        "autotuned": {"kernel": chosen, "threads_per_workgroup": threads},
        "batched_ms": round(b, 6),
        "mortar_ms": round(ms["mortar"], 6),
        "mixed_batched_ms": round(ms["mixed_batched"], 6),
        "per_extents_ms": round(ms["mixed_per_extents"], 6),
        "batch_gain": round(ms["mixed_per_extents"] / b, 3),
        "grid_points_per_s": round(points / (b * 1e-3), 1),
        "hbm_frac": round(alg / (b * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "algorithmic_bytes": alg,
        "rk3_step_graph_ms": round(graph_ms, 6),
        "rk3_step_call_by_call_ms": round(eager_ms, 6),
    }


def measure_gauss(steps=50, warmup=5):
    """ScalarWave RHS on Legendre-Gauss points (512 elements of 8^3, plane
    wave on [-1,1]^3, composite kernels: Gauss derivatives, point-wise terms,
    three face launches lifting onto whole pencils) against the LGL composite
    and fused RHS of the same brick size."""
    import numpy as np
    import torch
    from spectre_amd import dg, workloads as wl

    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    brick, u, g2, inv, det_inv, face_inv = wl.scalar_wave_plane_wave_gauss(
        seed=0, elements=(8, 8, 8), extents=(8, 8, 8))
    d_u, d_g, d_inv, d_det, d_face = dev(u), dev(g2), dev(inv), dev(det_inv), dev(face_inv)
    d_dt = torch.empty_like(d_u)
    domain = dg.Domain.from_brick(brick)
    ms = {"gauss_composite": time_calls(
        lambda: domain.scalar_wave_rhs_gauss(d_u, d_g, d_inv, d_det, d_face, out=d_dt),
        steps, warmup)}
    for kernel in ("composite", "fused"):
        domain.set_kernel("scalar_wave", kernel)
        ms["lgl_" + kernel] = time_calls(lambda: domain.scalar_wave_rhs(d_u, d_g, d_inv, out=d_dt),
                                         steps, warmup)
    points = brick.num_elements * brick.num_points
    return {"workload": "scalar_wave_plane_wave_gauss_512el_8^3", "grid_points": points,
            "ms": {k: round(v, 6) for k, v in ms.items()},
            "gauss_grid_points_per_s": round(points / (ms["gauss_composite"] * 1e-3), 1)}


def measure_mortars(steps=30, warmup=3):
    """The non-conforming mortar path: a periodic brick of 8^3 cells, every
    16th cell split in all three dimensions and every 16th (offset 5) in one,
    elements rotated at random, extents 6..8 (768 elements); ScalarWave
    (plane wave) and GH (Kerr-Schild).  And the cost of the general path on
    a conforming, aligned 512 x 8^3 brick: the mortar kernels against the
    mixed-p batched and the fused kernels of the same brick."""
    import numpy as np
    import torch
    from spectre_amd import dg, workloads as wl

    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    refine = {c: (0, 1, 2) for c in range(0, 512, 16)}
    refine.update({c: (1,) for c in range(5, 512, 16)})
    out = {}
    brick, u, g2, inv = wl.scalar_wave_mortar_workload(seed=0, base=(8, 8, 8), refine=refine,
                                                       lowest=6, highest=8)
    domain = dg.MortarDomain.from_brick(brick)
    d_u, d_g, d_i = dev(u), dev(g2), dev(inv)
    d_dt = torch.empty_like(d_u)
    ms = time_calls(lambda: domain.scalar_wave_rhs(d_u, d_g, d_i, out=d_dt), steps, warmup)
    out["sw_refined_rotated"] = {"elements": brick.num_elements, "mortars": len(brick.records),
                                 "grid_points": brick.total_points, "ms": round(ms, 6),
                                 "grid_points_per_s": round(brick.total_points / (ms * 1e-3), 1)}
    brick, u, ga, inv = wl.gh_mortar_workload(seed=0, base=(8, 8, 8), refine=refine, lowest=6,
                                              highest=8, noise=1.0e-6)
    domain = dg.MortarDomain.from_brick(brick)
    d_u, d_g, d_i = dev(u), dev(ga), dev(inv)
    d_dt = torch.empty_like(d_u)
    ms = time_calls(lambda: domain.gh_rhs(d_u, d_g, d_i, out=d_dt), steps, warmup)
    out["gh_refined_rotated"] = {"elements": brick.num_elements, "mortars": len(brick.records),
                                 "grid_points": brick.total_points, "ms": round(ms, 6),
                                 "grid_points_per_s": round(brick.total_points / (ms * 1e-3), 1)}
    # conforming aligned 512 x 8^3: mortar kernels vs mixed batched vs fused
    for system in ("scalar_wave", "gh"):
        if system == "scalar_wave":
            mb, u, par, inv = wl.scalar_wave_plane_wave_mixed(seed=0, elements=(8, 8, 8),
                                                             lowest=8, highest=8)
        else:
            mb, u, par, inv = wl.gh_kerr_schild_mixed(seed=0, elements=(8, 8, 8), lowest=8,
                                                     highest=8)
        nb = np.asarray(mb.neighbors)
        rec = np.array([[e, d, nb[e, d], 1, 3, 5, 1, 1] for e in range(nb.shape[0])
                        for d in range(6)], dtype=np.int32)
        d_u, d_p, d_i = dev(u), dev(par), dev(inv)
        d_dt = torch.empty_like(d_u)
        rhs = (lambda dm: (lambda: dm.scalar_wave_rhs(d_u, d_p, d_i, out=d_dt))) \
            if system == "scalar_wave" else (lambda dm: (lambda: dm.gh_rhs(d_u, d_p, d_i, out=d_dt)))
        mo = dg.MortarDomain(mb.extents, rec)
        mx = dg.MixedDomain.from_brick(mb)
        mx.set_kernel(system, "mixed_batched")
        res = {"mortar": time_calls(rhs(mo), steps, warmup),
               "mixed_batched": time_calls(rhs(mx), steps, warmup)}
        if system == "scalar_wave":
            ub, uu, ug, ui = wl.scalar_wave_plane_wave_3d(seed=0, elements=(8, 8, 8),
                                                          extents=(8, 8, 8))
        else:
            ub, uu, ug, ui = wl.gh_kerr_schild_brick(seed=0, elements=(8, 8, 8),
                                                     extents=(8, 8, 8))
        ud = dg.Domain.from_brick(ub)
        ud.set_kernel(system, "fused")
        e_u, e_g, e_i = dev(uu), dev(ug), dev(ui)
        e_dt = torch.empty_like(e_u)
        fn = (lambda: ud.scalar_wave_rhs(e_u, e_g, e_i, out=e_dt)) if system == "scalar_wave" \
            else (lambda: ud.gh_rhs(e_u, e_g, e_i, out=e_dt))
        res["fused"] = time_calls(fn, steps, warmup)
        out[system + "_conforming_512el_8^3_ms"] = {k: round(v, 6) for k, v in res.items()}
    return out


def measure_boundaries(steps=50, warmup=5):
    """Cost of the external boundary conditions: the RHS of a non-periodic
    brick whose outer faces are closed by a condition, against the periodic
    RHS of the same brick and data (one launch per RHS for the fused kernels).
    GH: 512 elements of 8^3 (Kerr-Schild, Dirichlet analytic ghost), fused
    vector-ALU kernel, 256 threads.  ScalarWave: 512 elements of 8^3 (plane
    wave; Dirichlet, Sommerfeld spherical radiation, constraint-preserving
    Sommerfeld), fused kernel; and the mixed-p brick of 1000 elements with
    extents 4..12 (Dirichlet, constraint-preserving Sommerfeld), batched."""
    import numpy as np
    import torch
    from spectre_amd import dg, workloads as wl

    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    out = {}
    brick, u, gammas, inv, ghost = wl.gh_kerr_schild_dirichlet(seed=0, elements=(8, 8, 8),
                                                               extents=(8, 8, 8))
    periodic = wl.Brick(brick.lower, brick.upper, brick.elements, brick.extents)
    d_u, d_g, d_inv, d_gh = dev(u), dev(gammas), dev(inv), dev(ghost)
    d_dt = torch.empty_like(d_u)
    ms = {}
    for name, b in (("periodic", periodic), ("dirichlet", brick)):
        domain = dg.Domain.from_brick(b)
        domain.set_kernel("gh", "fused")
        domain.set_threads("gh", 256)
        if name == "periodic":
            fn = lambda: domain.gh_rhs(d_u, d_g, d_inv, out=d_dt)
        else:
            fn = lambda: domain.gh_rhs_dirichlet(d_u, d_g, d_inv, d_gh, out=d_dt)
        ms[name] = round(time_calls(fn, steps, warmup), 6)
    out["gh_512el_8^3_fused_ms"] = ms
    del d_u, d_g, d_inv, d_gh, d_dt

    brick, u, g2, inv, ghost = wl.scalar_wave_plane_wave_dirichlet(
        seed=0, elements=(8, 8, 8), extents=(8, 8, 8), t=0.1, noise=1e-2,
        lower=(1.0, -4.0, -4.0), upper=(9.0, 4.0, 4.0))
    periodic = wl.Brick(brick.lower, brick.upper, brick.elements, brick.extents)
    d_u, d_g, d_inv, d_gh = dev(u), dev(g2), dev(inv), dev(ghost)
    d_dt = torch.empty_like(d_u)
    ms = {}
    for name in ("periodic", "dirichlet", "spherical_radiation", "cpsr"):
        domain = dg.Domain.from_brick(periodic if name == "periodic" else brick)
        domain.set_kernel("scalar_wave", "fused")
        fn = {"periodic": lambda: domain.scalar_wave_rhs(d_u, d_g, d_inv, out=d_dt),
              "dirichlet": lambda: domain.scalar_wave_rhs_dirichlet(d_u, d_g, d_inv, d_gh,
                                                                    out=d_dt),
              "spherical_radiation": lambda: domain.scalar_wave_rhs_spherical_radiation(
                  d_u, d_g, d_inv, "sommerfeld", out=d_dt),
              "cpsr": lambda: domain.scalar_wave_rhs_constraint_preserving_spherical_radiation(
                  d_u, d_g, d_inv, "sommerfeld", out=d_dt)}[name]
        ms[name] = round(time_calls(fn, steps, warmup), 6)
    out["scalar_wave_512el_8^3_fused_ms"] = ms

    brick, u, g2, inv, ghost = wl.scalar_wave_plane_wave_mixed_dirichlet(
        seed=0, elements=(10, 10, 10), lowest=4, highest=12, t=0.1, noise=1e-2)
    periodic = wl.MixedBrick(brick.lower, brick.upper, brick.elements, brick.extents)
    d_u, d_g, d_inv, d_gh = dev(u), dev(g2), dev(inv), dev(ghost)
    d_dt = torch.empty_like(d_u)
    ms = {}
    for name in ("periodic", "dirichlet", "cpsr"):
        domain = dg.MixedDomain.from_brick(periodic if name == "periodic" else brick)
        domain.set_kernel("scalar_wave", "mixed_batched")
        fn = {"periodic": lambda: domain.scalar_wave_rhs(d_u, d_g, d_inv, out=d_dt),
              "dirichlet": lambda: domain.scalar_wave_rhs_dirichlet(d_u, d_g, d_inv, d_gh,
                                                                    out=d_dt),
              "cpsr": lambda: domain.scalar_wave_rhs_constraint_preserving_spherical_radiation(
                  d_u, d_g, d_inv, "sommerfeld", out=d_dt)}[name]
        ms[name] = round(time_calls(fn, steps, warmup), 6)
    out["scalar_wave_mixed_p_1000el_extents_4..12_batched_ms"] = ms
    return out


def measure_adjoint(steps=40, reps=3):
    """Adjoint ScalarWave workload (512 elements x 6^3, 40 RK3 steps at CFL
    0.25, J = sum M psi^2): time of the forward evolution against the
    backward (hand-written adjoint kernels), with and without recomputing
    each 8-step segment from its checkpoint."""
    import numpy as np
    import torch
    from spectre_amd import dg, evolution as ev, workloads as wl

    brick, u, g2, inv, dt = wl.scalar_wave_adjoint_workload()
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    domain = dg.Domain.from_brick(brick)
    d_g2, d_inv = dev(g2), dev(inv)
    mass = ev.mass_weights(brick.extents, d_inv, brick.num_elements)

    def run(recompute):
        u0 = dev(u).requires_grad_(True)
        torch.cuda.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t2 = torch.cuda.Event(enable_timing=True)
        t0.record()
        uf = ev.ScalarWaveEvolution.apply(u0, domain, d_g2, d_inv, dt, steps, 8, recompute)
        t1.record()
        J = ev.psi_energy(uf, mass)
        J.backward()
        t2.record()
        torch.cuda.synchronize()
        return t0.elapsed_time(t1), t1.elapsed_time(t2)

    run(False)
    fwd, bwd = zip(*[run(False) for _ in range(reps)])
    _, bwd_rc = zip(*[run(True) for _ in range(reps)])
    f = min(fwd)
    return {"workload": "scalar_wave_adjoint_512el_6^3_40_rk3_steps",
            "grid_points": brick.num_elements * brick.num_points, "steps": steps,
            "dt": dt, "forward_ms": round(f, 4), "backward_ms": round(min(bwd), 4),
            "backward_with_recompute_ms": round(min(bwd_rc), 4),
            "adjoint_forward_cost": round(min(bwd) / f, 3),
            "adjoint_forward_cost_with_recompute": round(min(bwd_rc) / f, 3)}


def measure_gh_adjoint(steps=20, reps=2, mixed=False):
    """Adjoint GH workload (300 elements, 20 RK3 steps, J = L2 norm of the
    gauge constraint): forward evolution against the backward (hand-written
    adjoint kernels), with a checkpoint every 8 and every steps / 2 steps
    (earlier segments recomputed) and with every stage input kept by the
    forward.
    Uniform: 6^3 points per element; mixed: extents drawn from {6..10} per
    element and dimension."""
    import numpy as np
    import torch
    from spectre_amd import dg, evolution as ev, workloads as wl

    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    if mixed:
        brick, u, gammas, inv, dt = wl.gh_adjoint_workload_mixed()
        domain = dg.MixedDomain.from_brick(brick)
        d_g, d_inv = dev(gammas), dev(inv)
        index = ev.ragged_point_index(brick.extents, 50).to(d_inv.device)
        mass = ev.ragged_mass_weights(brick.extents, d_inv)
        functional = lambda uf: ev.ragged_gauge_constraint_norm(uf, index, mass)
        points, name = brick.total_points, "gh_adjoint_300el_mixed_6..10_20_rk3_steps_gauge_constraint"
    else:
        brick, u, gammas, inv, dt = wl.gh_adjoint_workload()
        domain = dg.Domain.from_brick(brick)
        d_g, d_inv = dev(gammas), dev(inv)
        mass = ev.mass_weights(brick.extents, d_inv, brick.num_elements)
        functional = lambda uf: ev.gauge_constraint_norm(uf, mass)
        points, name = brick.num_elements * brick.num_points, \
            "gh_adjoint_300el_6^3_20_rk3_steps_gauge_constraint"

    def run(k):
        u0 = dev(u).requires_grad_(True)
        torch.cuda.synchronize()
        t0, t1, t2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        t0.record()
        uf = ev.GhEvolution.apply(u0, domain, d_g, d_inv, dt, steps, k)
        t1.record()
        J = functional(uf)
        J.backward()
        t2.record()
        torch.cuda.synchronize()
        return t0.elapsed_time(t1), t1.elapsed_time(t2)

    # checkpoint every 8 steps and every steps / 2 steps (earlier segments
    # recomputed in the backward; the forward keeps the last segment's stage
    # inputs), and one segment (the forward keeps every stage input: 3 states
    # per step, a few GB at these sizes, no recomputation)
    half = max(steps // 2, 1)
    timings = {}
    for k in sorted({8, half, steps}):
        run(k)
        fwd, bwd = zip(*[run(k) for _ in range(reps)])
        timings[k] = (min(fwd), min(bwd))
    u_d, gb = dev(u), dev(u)
    out = torch.empty_like(u_d)
    rhs_ms = time_calls(lambda: domain.gh_rhs(u_d, d_g, d_inv, out=out), 5, 1)
    adj_ms = time_calls(lambda: domain.gh_rhs_adjoint(u_d, d_g, d_inv, gb, out=out), 3, 1)
    f8, b8 = timings[8]
    fh, bh = timings[half]
    fa, ba = timings[steps]
    return {"workload": name, "grid_points": int(points), "steps": steps, "dt": dt,
            "forward_ms": round(f8, 3), "backward_with_recompute_ms": round(b8, 3),
            "adjoint_forward_cost": round(b8 / f8, 3),
            f"forward_ms_checkpoint_{half}": round(fh, 3),
            f"backward_ms_checkpoint_{half}": round(bh, 3),
            f"adjoint_forward_cost_checkpoint_{half}": round(bh / fh, 3),
            "forward_ms_all_stages_kept": round(fa, 3),
            "backward_ms_all_stages_kept": round(ba, 3),
            "adjoint_forward_cost_all_stages_kept": round(ba / fa, 3),
            "rhs_ms": round(rhs_ms, 4), "adjoint_rhs_ms": round(adj_ms, 4)}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", choices=["gh", "sw"], default="gh")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="wall-clock budget of the CPU baseline sample (0 = skip)")
    ap.add_argument("--kernel", default="tune",
                    choices=["tune", "auto", "fused", "fused_mfma", "composite"],
                    help="kernel organisation of the RHS; 'tune' runs the library's "
                         "autotuner on the workload first and keeps its choice")
    ap.add_argument("--threads", type=int, default=0,
                    help="threads per workgroup with an explicit --kernel (0 = library default)")
    ap.add_argument("--rotate", type=int, default=0,
                    help="number of input/output buffer sets the timed steps cycle through "
                         "(0 = enough to exceed 3x the 256 MiB Infinity Cache, at most 8)")
    ap.add_argument("--extensions", type=int, default=1,
                    help="also measure the mixed-p batched form (1) or not (0)")
    ap.add_argument("--only-extension", default="",
                    choices=["", "sw_graph", "mixed_sw", "mixed_gh", "adjoint", "gh_adjoint",
                             "gh_adjoint_mixed", "boundaries", "gauss", "mortars"],
                    help="run only this extension measurement (for profiling) and print "
                         "its JSON")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (a plain copy "
                         "kernel of known size, to calibrate HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not "
              "supported (across GPUs the path shards by elements with one face-trace "
              "exchange per RHS; that is not built here)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.copy_calibration:
        print(json.dumps(hbm_probe(args.copy_calibration)))
        return

    root = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, root)
    from spectre_amd import dg, workloads as wl

    if args.only_extension:
        only = {"sw_graph": measure_sw_graph,
                "mixed_sw": lambda: measure_mixed("scalar_wave", 20, 3),
                "mixed_gh": lambda: measure_mixed("gh", 20, 3),
                "adjoint": measure_adjoint,
                "gh_adjoint": measure_gh_adjoint,
                "gh_adjoint_mixed": lambda: measure_gh_adjoint(mixed=True),
                "boundaries": measure_boundaries,
                "gauss": measure_gauss,
                "mortars": measure_mortars}
        print(json.dumps(only[args.only_extension]()))
        return

    if args.workload == "gh":
        brick, u, aux, inv = wl.gh_kerr_schild_brick(seed=0)
        nc, n_aux = 50, 3
        name = "gh_kerr_schild_brick_512el_8^3"
    else:
        brick, u, aux, inv = wl.scalar_wave_plane_wave_3d(seed=0)
        nc, n_aux = 5, 1
        name = "scalar_wave_plane_wave_64el_5^3"
    n = brick.extents[0]
    points = brick.num_elements * brick.num_points

    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    d_u, d_aux, d_inv = dev(u), dev(aux), dev(inv)
    d_dt = torch.empty_like(d_u)
    domain = dg.Domain.from_brick(brick)
    rhs = domain.gh_rhs if args.workload == "gh" else domain.scalar_wave_rhs
    system = "gh" if args.workload == "gh" else "scalar_wave"
    if args.kernel == "tune":
        domain.autotune(system, d_u, d_aux, d_inv, d_dt, repetitions=10)
    else:
        domain.set_kernel(system, args.kernel)
        if args.threads:
            domain.set_threads(system, args.threads)
    kernel_used = domain.kernel(system)

    # The timed steps cycle through `rotate` copies of the inputs and outputs
    # whose total size is at least three times the 256 MiB last-level
    # (Infinity) cache, so every step reads its inputs from HBM rather than
    # from what the previous step left in that cache.
    footprint = sum(t.numel() * 8 for t in (d_u, d_aux, d_inv, d_dt))
    rotate = args.rotate if args.rotate > 0 else \
        max(1, min(8, -(-3 * LLC_BYTES // footprint)))
    sets = [(d_u, d_aux, d_inv, d_dt)] + [
        (d_u.clone(), d_aux.clone(), d_inv.clone(), torch.empty_like(d_dt))
        for _ in range(rotate - 1)]
    for k in range(args.warmup):
        a, p, j, o = sets[k % rotate]
        rhs(a, p, j, out=o)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for k in range(args.steps):
        a, p, j, o = sets[k % rotate]
        rhs(a, p, j, out=o)
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = start.elapsed_time(stop) / args.steps
    value = points / (ms_per_step * 1e-3)
    # the same launches on one set of buffers (inputs partly served by the
    # last-level cache), for comparison
    ms_same_buffers = time_calls(lambda: rhs(d_u, d_aux, d_inv, out=d_dt), args.steps, 2)
    del sets

    bpp = algorithmic_bytes_per_point(nc, n_aux, n)
    bytes_per_launch = bpp * points
    achieved_gbs = bytes_per_launch / (ms_per_step * 1e-3) / 1e9
    # exact algorithmic flops per point (oracle run with a counting scalar type)
    flops_pp = None
    golden = os.path.join(root, "tests", "golden", "flop_counts.json")
    for entry in json.load(open(golden)):
        if entry["system"] == args.workload and entry["extents"] == list(brick.extents):
            flops_pp = entry["flops_per_point"]
    tflops = flops_pp * points / (ms_per_step * 1e-3) / 1e12 if flops_pp else None
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved_gbs, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved_gbs / HBM_PEAK_GBS, 4),
        "traffic": MEASURED_TRAFFIC.get((system, kernel_used)),
        "kernel": f"{system} RHS, {kernel_used} kernel organisation",
        "algorithmic_bytes_per_point": bpp,
        "algorithmic_bytes_per_launch": bytes_per_launch,
        "algorithmic_flops_per_point": flops_pp,
        "achieved_fp64_tflops": round(tflops, 3) if tflops else None,
        "fp64_peak_tflops": FP64_PEAK_TFLOPS,
        "fp64_frac": round(tflops / FP64_PEAK_TFLOPS, 4) if tflops else None,
    }

    cpu = None
    if args.cpu_seconds > 0:
        sys.path.insert(0, os.path.join(root, "tests"))
        import oracle_lib as ol
        # GH: the volume terms arranged as the reference computes them on a
        # CPU (per-dimension transposes and matrix products, materialised
        # temporaries); ScalarWave: the oracle
        fnt = ol.gh_rhs_refstyle_threads if args.workload == "gh" else ol.sw_rhs_threads
        out = np.empty_like(u)

        def rate(threads, seconds):
            reps = 0
            t = time.perf_counter()
            while True:
                fnt(threads, brick.extents, brick.neighbors, u, aux, inv, out=out)
                reps += 1
                if time.perf_counter() - t >= seconds:
                    break
            return reps, reps * points / (time.perf_counter() - t)

        single_reps, single = rate(1, min(5.0, args.cpu_seconds / 4))
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        reps, multi = rate(threads, args.cpu_seconds)
        arrangement = ("GH volume terms arranged as the reference computes them (per-dimension "
                       "transposes and matrix products, materialised temporaries), faces point "
                       "by point" if args.workload == "gh" else "the oracle")
        cpu = {"value": round(multi, 1), "unit": "grid-points/s", "cores": threads,
               "kind": "port",
               "sample": f"{reps} full RHS evaluations of the same workload on {threads} "
                         f"host threads (elements split over threads; {arrangement}); "
                         f"1-thread rate {single:.4g} pts/s from {single_reps} evaluations",
               "single_core_value": round(single, 1)}

    extensions = None
    if args.extensions:
        extensions = {"sw_64el_graph": measure_sw_graph(),
                      "mixed_p": [measure_mixed(s, 20, 3) for s in ("scalar_wave", "gh")],
                      "adjoint": measure_adjoint(),
                      "gh_adjoint": measure_gh_adjoint(),
                      "gh_adjoint_mixed": measure_gh_adjoint(mixed=True),
                      "boundaries": measure_boundaries(),
                      "gauss": measure_gauss(),
                      "mortars": measure_mortars()}

    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "grid-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": name, "elements": brick.num_elements,
                   "extents": list(brick.extents), "grid_points": points,
                   "components": nc, "parallelism": "single GPU",
                   "kernel": kernel_used, "threads_per_workgroup": domain.threads(system),
                   "kernel_choice": args.kernel},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extensions": extensions,
        "wall_s_timed_region": round(wall, 4),
        "buffer_sets_rotated": rotate,
        "ms_per_step_same_buffers": round(ms_same_buffers, 6),
    }
    print(json.dumps(result))


if __name__ == "__main__":
    main()
