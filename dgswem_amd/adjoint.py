"""Gauge-misfit loss and its gradient through the library's discrete adjoint, as a
torch.autograd.Function.

    loss = GaugeMisfit.apply(manning, dp, problem)
    loss.backward()          # manning.grad, dp.grad

`GaugeMisfit.forward` runs the forward model on the GPU (dg_adj_gauges) and evaluates
J = weight * sum_k sum_g (eta_g(U^k) - d_g^k)^2; `backward` runs dg_adj_gradient: a checkpointed
forward sweep and the hand-written adjoint kernels of the stage RHS and SSP-RK stages.  Both start
from the problem's initial state, which is restored before every evaluation.  Tensors may live on
any device; the library takes host copies.  torch is only the autograd front end here.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import gpu
from .fortio import Case


@dataclass
class MisfitProblem:
    """A run (its initial state), gauges, the observed gauge series and the loss weight."""
    solver: gpu.GpuSolver
    nsteps: int
    data: np.ndarray                       # [nsteps, ng]
    weight: float = 1.0
    interval: int = 0                      # checkpoint spacing (0: about sqrt(nsteps))
    state0: tuple = field(default=None)    # initial (ze, qx, qy, wdflg)
    it0: int = 0

    def reset(self) -> None:
        ze, qx, qy, wd = self.state0
        self.solver.set_state(ze, qx, qy, wd)
        self.solver.it = self.it0

    def loss(self, manning: np.ndarray, dp: np.ndarray) -> float:
        s = self.solver
        s.set_manning(manning); s.set_depth(dp)
        self.reset()
        eta = s.gauges(self.nsteps)
        self.reset()
        d = eta - self.data
        return float(self.weight * np.sum(d * d))

    def gradient(self, manning: np.ndarray, dp: np.ndarray):
        s = self.solver
        s.set_manning(manning); s.set_depth(dp)
        self.reset()
        return s.gradient(self.data, self.weight, self.interval)


def seeded_gauges(case: Case, ng: int, seed: int):
    """ng gauge points: seeded elements and seeded points inside them (reference coordinates)."""
    rng = np.random.default_rng(seed)
    elem = rng.choice(case.ne, size=ng, replace=False) + 1
    a = rng.uniform(0.1, 0.8, ng); b = rng.uniform(0.1, 0.9, ng) * (1.0 - a)
    r = -1.0 + 2.0 * a; s = -1.0 + 2.0 * b
    return elem.astype(np.int32), r, s


def hump_state(case: Case, amplitude: float = 0.5, width: float = 0.25):
    """Initial state at rest with a Gaussian elevation hump centred in the basin (cell means only),
    so that the whole domain moves from the first step."""
    x = case.x[case.nm - 1].mean(axis=1); y = case.y[case.nm - 1].mean(axis=1)
    cx, cy = 0.5 * (case.x.min() + case.x.max()), 0.5 * (case.y.min() + case.y.max())
    L = width * max(np.ptp(case.x), np.ptp(case.y))
    ze = np.zeros((case.ne, case.dofh))
    ze[:, 0] = amplitude * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / L ** 2)
    z = np.zeros_like(ze)
    return ze, z, z.copy(), np.ones(case.ne, np.int32)


def twin_problem(case: Case, nsteps: int, ng: int = 50, seed: int = 0, perturb: float = 0.2,
                 weight: float | None = None, interval: int = 0, state0=None) -> MisfitProblem:
    """Gauge data from a twin run whose Manning field is perturbed by a seeded smooth relative
    factor (1 + perturb * field), then a problem set up at the unperturbed parameters.  The runs
    start from `state0` (default: a Gaussian hump at rest, see hump_state; "cold": the solver's
    cold start, at rest with the initial dry check of src/prep_DG.F:1214-1300)."""
    s = gpu.GpuSolver(case)
    elem, r, sr = seeded_gauges(case, ng, seed)
    s.set_gauges(elem, r, sr)
    if isinstance(state0, str):
        if state0 != "cold":
            raise ValueError(f"unknown initial state {state0!r}")
        state0 = s.state()
    elif state0 is None:
        state0 = hump_state(case)
    s.set_state(*state0)
    rng = np.random.default_rng(seed + 1)
    kx, ky, ph = rng.uniform(1.0, 3.0), rng.uniform(1.0, 3.0), rng.uniform(0.0, 2 * np.pi)
    x = (case.x - case.x.min()) / np.ptp(case.x); y = (case.y - case.y.min()) / np.ptp(case.y)
    f = np.sin(2 * np.pi * kx * x + ph) * np.cos(2 * np.pi * ky * y)
    s.set_manning(case.manning * (1.0 + perturb * f))
    data = s.gauges(nsteps)
    s.set_manning(case.manning)
    w = weight if weight is not None else 1.0 / (nsteps * ng)
    p = MisfitProblem(solver=s, nsteps=nsteps, data=data, weight=w, interval=interval,
                      state0=state0, it0=0)
    p.reset()
    return p


def rest_state(case: Case):
    """All-wet state at rest (zero elevation and flow).  Where the bottom rises above H0 this is
    not a valid wet/dry state; twin_problem's "cold" start is."""
    z = np.zeros((case.ne, case.dofh))
    return z, z.copy(), z.copy(), np.ones(case.ne, np.int32)


def ensemble_problem(cases, nsteps: int, ng: int = 20, seed: int = 0, perturb: float = 0.2,
                     state0: str = "rest"):
    """Batched twin problems: each member gets its own gauges and its own perturbed-Manning twin
    data (as in twin_problem, seeds seed+i), all with the same loss weight 1/(nsteps*ng), from
    rest (the cold start: elements whose bottom rises above H0 start dry) or from a hump.  Returns
    the ensemble's MisfitProblem (gauges and data concatenated in member order) and the members'
    own problems; the gradient of the summed loss restricted to a member's nodes is that member's
    gradient."""
    w = 1.0 / (nsteps * ng)
    probs = []
    for i, c in enumerate(cases):
        st = "cold" if state0 == "rest" else hump_state(c)
        probs.append(twin_problem(c, nsteps=nsteps, ng=ng, seed=seed + i, perturb=perturb, weight=w,
                                  state0=st))
    ens = gpu.Ensemble([p.solver for p in probs])
    elems, rs, ss = [], [], []
    for i, c in enumerate(cases):
        el, r, s_ = seeded_gauges(c, ng, seed + i)
        elems.append(el + ens.eoff[i]); rs.append(r); ss.append(s_)
    ens.set_gauges(np.concatenate(elems).astype(np.int32), np.concatenate(rs), np.concatenate(ss))
    data = np.concatenate([p.data for p in probs], axis=1)
    st0 = tuple(np.concatenate([p.state0[j] for p in probs]) for j in range(4))
    eprob = MisfitProblem(solver=ens, nsteps=nsteps, data=data, weight=w, state0=st0, it0=0)
    eprob.reset()
    return eprob, probs


def _make_function():
    import torch

    class GaugeMisfit(torch.autograd.Function):
        """Forward: when a gradient will be asked for, one dg_adj_gradient call runs the
        checkpointed forward sweep and the reverse sweep together and returns the loss with the
        gradient, which backward then only scales; otherwise the plain forward model."""

        @staticmethod
        def forward(ctx, manning, dp, problem: MisfitProblem, want_grad: bool):
            mn = manning.detach().to("cpu", torch.float64).numpy()
            d = dp.detach().to("cpu", torch.float64).numpy()
            ctx.devices = (manning.device, dp.device)
            if want_grad:
                J, gn, gdp = problem.gradient(mn, d)
                ctx.grads = (gn, gdp)
            else:
                J = problem.loss(mn, d)
                ctx.grads = None
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_out):
            gn, gdp = ctx.grads
            g = float(grad_out)
            return (torch.from_numpy(g * gn).to(ctx.devices[0]),
                    torch.from_numpy(g * gdp).to(ctx.devices[1]), None, None)

    return GaugeMisfit


_FN = None


def gauge_misfit(manning, dp, problem: MisfitProblem):
    """torch front end: differentiable loss of torch float64 tensors `manning` and `dp` (per node)."""
    import torch
    global _FN
    if _FN is None:
        _FN = _make_function()
    want = torch.is_grad_enabled() and (manning.requires_grad or dp.requires_grad)
    return _FN.apply(manning, dp, problem, want)
