"""Time evolution of a DG system with Rk3HesthavenSsp and its gradient.

`ScalarWaveEvolution` is a torch.autograd.Function: forward runs `steps`
steps of the ScalarWave DG RHS (the hand-written HIP kernels behind the C ABI)
and returns the final state; backward applies the hand-written adjoint
kernels (transposed RHS and transposed RK3 substeps) step by step in reverse.
The forward keeps a checkpoint of the state every `checkpoint_every` steps;
the backward walks the segments in reverse and, if `recompute` is set,
recomputes each segment's step inputs from its checkpoint before reversing it
(the schedule a system with a state-dependent linearisation needs; the
ScalarWave RHS is linear, so its step adjoint does not read those states and
`recompute=False` skips that work).

Reference: Time/TimeSteppers/Rk3HesthavenSsp.cpp:55-81 (the forward step);
the reference has no adjoint.

Every device call of these Functions goes through the pybind11 module
spectre_amd._Pybindings (the right-hand sides, their adjoints and the RK3
substeps on the domain's handle), which calls the C ABI.
"""
import numpy as np
import torch

from . import _Pybindings as _pyb
from . import dg


def _rhs(domain, system):
    """RHS callable fn(u, params, inv_jac, out) of `system` on `domain`."""
    call = _pyb.scalar_wave_rhs if system == "scalar_wave" else _pyb.gh_rhs
    return lambda u, params, inv_jac, out: call(domain.handle, u, params, inv_jac, out)


def _sw_rhs_adjoint(domain):
    return lambda gbar, gamma2, inv_jac, out: _pyb.scalar_wave_rhs_adjoint(
        domain.handle, gbar, gamma2, inv_jac, out)


def rk3_adjoint_substep(substep, dt, lam, a, mu, lam_out=None):
    """Transpose of one RK3 substep (sxd_rk3_adjoint_substep)."""
    _pyb.rk3_adjoint_substep(int(substep), float(dt), lam, a, mu, lam_out)
    return mu, lam_out


def rk3_step(domain, rhs, u, params, inv_jac, dt, stage, l):
    """One Rk3HesthavenSsp step of u in place (stage, l: work arrays)."""
    rhs(u, params, inv_jac, out=l)
    _pyb.rk3_substep(0, dt, u, u, l, stage)
    rhs(stage, params, inv_jac, out=l)
    _pyb.rk3_substep(1, dt, u, stage, l, stage)
    rhs(stage, params, inv_jac, out=l)
    _pyb.rk3_substep(2, dt, u, stage, l, u)


def rk3_step_adjoint(adjoint_rhs, lam3, params, inv_jac, dt, a, lam, mu):
    """Adjoint of one RK3 step of a linear RHS: mu = S^T lam3 (a, lam: work)."""
    adjoint_rhs(lam3, params, inv_jac, out=a)
    rk3_adjoint_substep(2, dt, lam3, a, mu, lam)        # lam = lam2, mu = lam3 / 3
    adjoint_rhs(lam, params, inv_jac, out=a)
    rk3_adjoint_substep(1, dt, lam, a, mu, lam)         # lam = lam1, mu += 3/4 lam2
    adjoint_rhs(lam, params, inv_jac, out=a)
    rk3_adjoint_substep(0, dt, lam, a, mu, None)        # mu += lam1 + dt A^T lam1
    return mu


class ScalarWaveEvolution(torch.autograd.Function):
    """u_final = S^steps u0 for the ScalarWave DG RHS on `domain`."""

    @staticmethod
    def forward(ctx, u0, domain, gamma2, inv_jac, dt, steps, checkpoint_every=8,
                recompute=False):
        domain.check_fields("scalar_wave", u0, gamma2, inv_jac)
        u = u0.detach().clone()
        stage, l = torch.empty_like(u), torch.empty_like(u)
        checkpoints = []
        rhs = _rhs(domain, "scalar_wave")
        for n in range(steps):
            if n % checkpoint_every == 0:
                checkpoints.append(u.clone())
            rk3_step(domain, rhs, u, gamma2, inv_jac, dt, stage, l)
        ctx.domain, ctx.dt, ctx.steps = domain, dt, steps
        ctx.checkpoint_every, ctx.recompute = checkpoint_every, recompute
        ctx.checkpoints = checkpoints
        ctx.save_for_backward(gamma2, inv_jac)
        return u

    @staticmethod
    def backward(ctx, grad_out):
        gamma2, inv_jac = ctx.saved_tensors
        domain, dt, k = ctx.domain, ctx.dt, ctx.checkpoint_every
        lam3 = grad_out.detach().contiguous().clone()
        a, lam, mu = (torch.empty_like(lam3) for _ in range(3))
        stage, l = torch.empty_like(lam3), torch.empty_like(lam3)
        for seg in reversed(range(len(ctx.checkpoints))):
            first = seg * k
            last = min(ctx.steps, first + k)
            if ctx.recompute:
                # inputs of the segment's steps, from its checkpoint
                u = ctx.checkpoints[seg].clone()
                states = []
                for _ in range(first, last):
                    states.append(u.clone())
                    rk3_step(domain, _rhs(domain, "scalar_wave"), u, gamma2, inv_jac, dt,
                             stage, l)
            for _ in range(last - first):
                rk3_step_adjoint(_sw_rhs_adjoint(domain), lam3, gamma2, inv_jac, dt, a, lam, mu)
                lam3, mu = mu, lam3
        ctx.checkpoints = None
        return lam3, None, None, None, None, None, None, None


def rk3_step_states(rhs, u, params, inv_jac, dt, l, complete=True):
    """One RK3 step of u (in place) returning its stage inputs (u, u1, u2).
    complete=False stops after the stage inputs (u is left unchanged): the
    step's result is not needed."""
    u0 = u.clone() if complete else u
    rhs(u, params, inv_jac, out=l)
    u1 = _pyb.rk3_substep(0, dt, u, u, l, torch.empty_like(u))
    rhs(u1, params, inv_jac, out=l)
    u2 = _pyb.rk3_substep(1, dt, u, u1, l, torch.empty_like(u))
    if complete:
        rhs(u2, params, inv_jac, out=l)
        _pyb.rk3_substep(2, dt, u, u2, l, u)
    return u0, u1, u2


class GhEvolution(torch.autograd.Function):
    """u_final after `steps` RK3 steps of the GH DG RHS on `domain`.

    The GH RHS is nonlinear, so each step's adjoint needs the step's stage
    inputs u, u1, u2: the forward keeps the state every `checkpoint_every`
    steps, and the backward recomputes the stage inputs of one segment at a
    time from its checkpoint, then applies the transposed substeps with the
    adjoint RHS linearised at those stages (sxd_gh_rhs_adjoint).  The last
    segment's stage inputs are kept by the forward itself (the backward
    starts there), so only the earlier segments are recomputed, each on a
    second stream while the adjoint of the segment after it runs."""

    @staticmethod
    def forward(ctx, u0, domain, gammas, inv_jac, dt, steps, checkpoint_every=8, overlap=None):
        domain.check_fields("gh", u0, gammas, inv_jac)
        u = u0.detach().clone()
        l = torch.empty_like(u)
        checkpoints, last_stages = [], []
        last_first = (max(steps - 1, 0) // checkpoint_every) * checkpoint_every
        rhs = _rhs(domain, "gh")
        for n in range(steps):
            if n % checkpoint_every == 0:
                checkpoints.append(u.clone())
            if n >= last_first:
                last_stages.append(rk3_step_states(rhs, u, gammas, inv_jac, dt, l))
            else:
                rk3_step(domain, rhs, u, gammas, inv_jac, dt, torch.empty_like(u), l)
        ctx.domain, ctx.dt, ctx.steps, ctx.k = domain, dt, steps, checkpoint_every
        # overlap the recompute with the adjoint (see backward): on by default
        # for mixed-p domains; the uniform fused RHS kernel fills whole CUs
        # and only competes with the adjoint kernels
        ctx.overlap = isinstance(domain, dg.MixedDomain) if overlap is None else bool(overlap)
        ctx.checkpoints, ctx.last_stages = checkpoints, last_stages
        ctx.save_for_backward(gammas, inv_jac)
        return u

    @staticmethod
    def backward(ctx, grad_out):
        gammas, inv_jac = ctx.saved_tensors
        domain, dt, k = ctx.domain, ctx.dt, ctx.k
        lam3 = grad_out.detach().contiguous().clone()
        a, lam, mu = (torch.empty_like(lam3) for _ in range(3))
        nseg = len(ctx.checkpoints)
        main = torch.cuda.current_stream(lam3.device)
        side, twin = None, domain
        if nseg > 1 and ctx.overlap:
            # The stage inputs of segment seg - 1 do not depend on segment
            # seg's adjoint: they are recomputed on a second stream (default
            # priority) by a twin domain with work arrays of its own, while
            # this stream runs the adjoint kernels.
            side = torch.cuda.Stream(device=lam3.device)
            side.wait_stream(main)
            twin = getattr(domain, "_recompute_twin", None)
            if twin is None:
                twin = domain._recompute_twin = domain.twin()

        def recompute(seg):
            first = seg * k
            last = min(ctx.steps, first + k)
            with torch.cuda.stream(side if side is not None else main):
                u = ctx.checkpoints[seg].clone()
                l2 = torch.empty_like(u)
                # the segment's last step is not completed: its result is
                # the next checkpoint
                stages = [rk3_step_states(_rhs(twin, "gh"), u, gammas, inv_jac, dt, l2,
                                          complete=n < last - 1)
                          for n in range(first, last)]
                done = None
                if side is not None:
                    done = torch.cuda.Event()
                    done.record(side)
            return stages, done

        pending = recompute(nseg - 2) if nseg > 1 and side is not None else None
        for seg in reversed(range(nseg)):
            if seg == nseg - 1:
                stages = ctx.last_stages
            else:
                stages, done = pending if pending is not None else recompute(seg)
                if done is not None:
                    main.wait_event(done)
                    for st in stages:
                        for t in st:
                            t.record_stream(main)
                pending = recompute(seg - 1) if seg > 0 and side is not None else None
            h = domain.handle
            for s0, s1, s2 in reversed(stages):
                _pyb.gh_rhs_adjoint(h, s2, gammas, inv_jac, lam3, a)
                rk3_adjoint_substep(2, dt, lam3, a, mu, lam)
                _pyb.gh_rhs_adjoint(h, s1, gammas, inv_jac, lam, a)
                rk3_adjoint_substep(1, dt, lam, a, mu, lam)
                _pyb.gh_rhs_adjoint(h, s0, gammas, inv_jac, lam, a)
                rk3_adjoint_substep(0, dt, lam, a, mu, None)
                lam3, mu = mu, lam3
        ctx.checkpoints = ctx.last_stages = None
        return lam3, None, None, None, None, None, None, None


def gauge_constraint(u):
    """Gauge constraint C_a of the first-order GH system with the harmonic gauge
    source (H_a = 0): the trace of the Christoffel symbols written with Pi,
    Phi and the normals, as gh::gauge_constraint / gh::trace_christoffel
    (Evolution/Systems/GeneralizedHarmonic/Constraints.cpp:983-998,
    PointwiseFunctions/GeneralRelativity/GeneralizedHarmonic/Christoffel.cpp:64-119).
    u [E][50][Np] -> C [E][4][Np] (differentiable torch operations: products
    and sums over the small tensor indices, the inverse spatial metric by
    cofactors)."""
    pairs = [(a, b) for a in range(4) for b in range(a, 4)]
    idx = {}
    for p, (a, b) in enumerate(pairs):
        idx[(a, b)] = idx[(b, a)] = p
    perm = torch.tensor([idx[(a, b)] for a in range(4) for b in range(4)], device=u.device)
    x = u.permute(0, 2, 1)                                          # [E][Np][50]
    # index_select rather than advanced indexing: its backward is a scatter-add
    # (each storage slot receives at most two terms, so the sum is exact in any
    # order) instead of torch's sort-based index_put backward
    gather = lambda first: torch.index_select(x, -1, first + perm).reshape(
        x.shape[0], x.shape[1], 4, 4)
    g = gather(0)                                                   # [E][Np][4][4]
    pi = gather(10)
    phi = torch.stack([torch.index_select(x, -1, 20 + i + 3 * perm).reshape(g.shape)
                       for i in range(3)], 2)
    s = g[..., 1:, 1:]                                              # spatial metric
    c00 = s[..., 1, 1] * s[..., 2, 2] - s[..., 1, 2] * s[..., 1, 2]
    c01 = s[..., 1, 2] * s[..., 0, 2] - s[..., 0, 1] * s[..., 2, 2]
    c02 = s[..., 0, 1] * s[..., 1, 2] - s[..., 1, 1] * s[..., 0, 2]
    c11 = s[..., 2, 2] * s[..., 0, 0] - s[..., 0, 2] * s[..., 0, 2]
    c12 = s[..., 0, 2] * s[..., 0, 1] - s[..., 0, 0] * s[..., 1, 2]
    c22 = s[..., 0, 0] * s[..., 1, 1] - s[..., 0, 1] * s[..., 0, 1]
    det = s[..., 0, 0] * c00 + s[..., 0, 1] * c01 + s[..., 0, 2] * c02
    inv_sp = torch.stack([torch.stack([c00, c01, c02], -1), torch.stack([c01, c11, c12], -1),
                          torch.stack([c02, c12, c22], -1)], -2) / det[..., None, None]
    g0i = g[..., 1:, 0]                                             # [E][Np][3]
    shift = (inv_sp * g0i[..., None, :]).sum(-1)
    lapse = torch.sqrt(-g[..., 0, 0] + (shift * g0i).sum(-1))
    n_up = torch.cat([(1.0 / lapse)[..., None], -shift / lapse[..., None]], -1)   # [E][Np][4]
    m = -1.0 / lapse ** 2
    top = torch.cat([m[..., None], (-shift * m[..., None])], -1)    # [E][Np][4]
    lower = inv_sp + shift[..., :, None] * shift[..., None, :] * m[..., None, None]
    lower = torch.cat([(-shift * m[..., None])[..., None], lower], -1)  # [E][Np][3][4]
    inv_g = torch.cat([top[..., None, :], lower], -2)               # [E][Np][4][4]
    phi_ig = (phi * inv_g[..., None, :, :]).sum((-2, -1))           # [E][Np][3]
    k = -0.5 * ((pi * inv_g).sum((-2, -1)) + (n_up[..., 1:] * phi_ig).sum(-1))
    c = []
    for a in range(4):
        term = (inv_sp * phi[..., :, 1:, a]).sum((-2, -1)) + (n_up * pi[..., :, a]).sum(-1)
        if a == 0:
            term = term - lapse * k
        else:
            term = term - 0.5 * phi_ig[..., a - 1]
        c.append(term)
    return torch.stack(c, 1)                                        # [E][4][Np]


def gauge_constraint_norm(u, mass):
    """J = ( sum over points of M sum_a C_a^2 )^(1/2)."""
    c = gauge_constraint(u)
    return torch.sqrt(torch.sum(mass[:, None, :] * c * c))


def mass_weights(domain_extents, inv_jac, num_elements):
    """Diagonal LGL mass matrix per point, w_i w_j w_k det(J) [E][Np] (the
    Jacobian determinant as 1 / det of the inverse Jacobian)."""
    from . import spectral
    n0, n1, n2 = domain_extents
    w = [spectral.lgl_weights(n) for n in (n0, n1, n2)]
    wk, wj, wi = np.meshgrid(w[2], w[1], w[0], indexing="ij")
    wq = torch.from_numpy((wi * wj * wk).ravel()).to(inv_jac.device)
    m = inv_jac.reshape(num_elements, 3, 3, -1)   # [E][i][j][p], block j + 3 i
    det_inv = torch.linalg.det(m.permute(0, 3, 1, 2))  # [E][p]
    return wq[None, :] / det_inv


def ragged_point_index(extents, nc):
    """Gather index of a ragged field (element e's nc blocks of Np_e points at
    nc * offset_e) as [nc][total points], points element by element:
    u_flat[index] is the field point-major, for point-wise functionals."""
    ext = np.asarray(extents, dtype=np.int64).reshape(-1, 3)
    points = np.prod(ext, axis=1)
    offsets = np.concatenate([[0], np.cumsum(points)])
    index = np.empty((nc, int(offsets[-1])), dtype=np.int64)
    comps = np.arange(nc)[:, None]
    for e, n in enumerate(points):
        index[:, offsets[e]:offsets[e + 1]] = nc * offsets[e] + comps * n + np.arange(n)[None, :]
    return torch.from_numpy(index)


def ragged_mass_weights(extents, inv_jac):
    """mass_weights of a mixed-p domain: [total points] in the ragged point
    order (inv_jac: flat ragged field of 9 components)."""
    from . import spectral
    ext = np.asarray(extents, dtype=np.int64).reshape(-1, 3)
    out, off = [], 0
    for n0, n1, n2 in ext:
        n = int(n0 * n1 * n2)
        w = [spectral.lgl_weights(int(k)) for k in (n0, n1, n2)]
        wk, wj, wi = np.meshgrid(w[2], w[1], w[0], indexing="ij")
        m = inv_jac[9 * off:9 * (off + n)].reshape(3, 3, n)
        det_inv = torch.linalg.det(m.permute(2, 0, 1))
        out.append(torch.from_numpy((wi * wj * wk).ravel()).to(inv_jac.device) / det_inv)
        off += n
    return torch.cat(out)


def ragged_gauge_constraint_norm(u, index, mass):
    """gauge_constraint_norm of a ragged GH state u (flat), with the gather
    index of ragged_point_index(extents, 50) and ragged_mass_weights."""
    pts = torch.index_select(u, 0, index.reshape(-1)).reshape(index.shape)
    return gauge_constraint_norm(pts[None], mass[None])


def psi_energy(u, mass):
    """J = sum over points of M psi^2 (psi = component 0)."""
    return torch.sum(mass * u[:, 0, :] ** 2)
