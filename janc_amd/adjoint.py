"""Differentiable steps: torch.autograd wrappers around jh_advance / jh_step_vjp.

``step(sim, U, aux, dt)`` is one fixed-dt step as a ``torch.autograd.Function``
(forward = the FP64 step kernels keeping the step's residuals, the RK stage
states and their LF maxima, ``jh_step_keep``; backward = the hand-written
adjoint kernels from those residuals, ``jh_step_vjp_res``, which recompute
nothing).  ``rollout(sim, U, aux, dt, nsteps, checkpoint_every)`` runs many
steps keeping only every ``checkpoint_every``-th state; its backward replays
each segment from its checkpoint keeping every step's residuals (one
segment's worth at a time) and sweeps the step VJPs in reverse.

These replace JAX's reverse mode over the reference's ``advance_one_step``
(set_simulator.py:117-122), the custom VJP of ``get_T_nasa7``
(thermo_model.py:155-190) included.
"""
import torch


def _vjp(sim, U, aux, dt, Ub, ab, out, res):
    return sim.step_vjp(U, aux, dt, Ub.contiguous(), ab.contiguous(), known_output=out,
                        residuals=res)


class _Step(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U, aux, sim, dt):
        U, aux = U.detach().contiguous(), aux.detach().contiguous()
        Uo, ao, res = sim.step_keep(U, aux, dt)
        ctx.save_for_backward(U, aux, Uo, ao, res)
        ctx.sim, ctx.dt = sim, dt
        return Uo, ao

    @staticmethod
    def backward(ctx, Ub, ab):
        U, aux, Uo, ao, res = ctx.saved_tensors
        Ub = torch.zeros_like(U) if Ub is None else Ub
        ab = torch.zeros_like(aux) if ab is None else ab
        gU, ga = _vjp(ctx.sim, U, aux, ctx.dt, Ub, ab, (Uo, ao), res)
        return gU, ga, None, None


def step(sim, U, aux, dt):
    """One differentiable fixed-dt step (FP64)."""
    return _Step.apply(U, aux, sim, float(dt))


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U, aux, sim, dt, nsteps, every):
        U = U.detach().contiguous().clone()
        aux = aux.detach().contiguous().clone()
        ckpt = []
        for s in range(0, nsteps, every):          # one fused call per segment
            ckpt.append((U.clone(), aux.clone()))
            sim.advance(U, aux, min(every, nsteps - s), dt=dt)
        ctx.sim, ctx.dt, ctx.nsteps, ctx.every = sim, dt, nsteps, every
        ctx.ckpt = ckpt
        ctx.final = (U.clone(), aux.clone())
        return U, aux

    @staticmethod
    def backward(ctx, Ub, ab):
        sim, dt, nsteps, every = ctx.sim, ctx.dt, ctx.nsteps, ctx.every
        Ub = Ub.contiguous().clone()
        ab = (torch.zeros_like(ctx.ckpt[0][1]) if ab is None else ab).contiguous().clone()
        for seg in reversed(range(len(ctx.ckpt))):
            U, aux = ctx.ckpt[seg]
            first = seg * every
            last = min(first + every, nsteps)
            steps = []
            for _ in range(first, last):           # replay the segment, keeping residuals
                Un, an, res = sim.step_keep(U, aux, dt)
                steps.append((U, aux, Un, an, res))
                U, aux = Un, an
            for Us, As, Un, an, res in reversed(steps):
                Ub, ab = _vjp(sim, Us, As, dt, Ub, ab, (Un, an), res)
            steps = None
        ctx.ckpt = ctx.final = None
        return Ub, ab, None, None, None, None


def rollout(sim, U, aux, dt, nsteps, checkpoint_every=10):
    """nsteps differentiable fixed-dt steps with checkpoints every
    ``checkpoint_every`` steps (memory: one state per checkpoint plus one
    segment during the backward sweep)."""
    return _Rollout.apply(U, aux, sim, float(dt), int(nsteps), int(checkpoint_every))


# ---------------------------------------------------------------- objectives
# Differentiable (torch) state construction and losses for gradients with
# respect to initial temperature and composition fields.

def _poly(coef, T):
    """Horner evaluation, coef (ns, m) highest power first, T (nx, ny)."""
    acc = torch.zeros((coef.shape[0],) + tuple(T.shape), dtype=T.dtype, device=T.device)
    for j in range(coef.shape[1]):
        acc = acc * T + coef[:, j].reshape(-1, 1, 1)
    return acc


def _consts(mech, ref):
    import numpy as np
    f = lambda a: torch.as_tensor(np.asarray(a), dtype=ref.dtype, device=ref.device)  # noqa
    return f(1.0 / mech.species_M), f(mech.cp), f(mech.h)


def mixture(mech, T, Y9):
    """(R, cp, h, gamma) of the mixture (NASA-7, per unit mass), as in the
    reference's get_thermo_nasa7 (model/thermo_model.py:94-132)."""
    invM, cp_c, h_c = _consts(mech, T)
    R = (invM.reshape(-1, 1, 1) * Y9).sum(0)
    lo = T < mech.Tcr
    cp = torch.where(lo, (_poly(cp_c[0], T) * Y9).sum(0), (_poly(cp_c[1], T) * Y9).sum(0))
    h = torch.where(lo, (_poly(h_c[0], T) * Y9).sum(0), (_poly(h_c[1], T) * Y9).sum(0))
    return R, cp, h, cp / (cp - R)


def initial_state(mech, T, p, Y, u=None, v=None):
    """(U, aux) from T, p (nx, ny) and the first ns-1 mass fractions Y
    (ns-1, nx, ny), differentiably: rho = p/(R T), rho*E = rho (h - R T) +
    kinetic energy, aux = (gamma, T)."""
    Y9 = torch.cat([Y, 1.0 - Y.sum(0, keepdim=True)], 0)
    R, cp, h, gamma = mixture(mech, T, Y9)
    rho = p / (R * T)
    u = torch.zeros_like(T) if u is None else u
    v = torch.zeros_like(T) if v is None else v
    rhoE = rho * h - p + 0.5 * rho * (u * u + v * v)
    U = torch.cat([rho[None], (rho * u)[None], (rho * v)[None], rhoE[None], rho[None] * Y], 0)
    return U, torch.stack([gamma, T])


def formation_enthalpy(mech, ref, T_ref=298.15):
    """h_k(T_ref) per unit mass of every species (NASA-7)."""
    _, _, h_c = _consts(mech, ref)
    r = 0 if T_ref < mech.Tcr else 1
    T = torch.full((1, 1), T_ref, dtype=ref.dtype, device=ref.device)
    return _poly(h_c[r], T).reshape(-1)


def heat_release(mech, U0, UN, dx, dy):
    """Chemical heat released between states U0 and UN: the decrease of
    sum_k h_k(298.15 K) rho Y_k integrated over the domain (N2 by closure)."""
    hf = formation_enthalpy(mech, U0)

    def chem(U):
        rY = U[4:]
        rN2 = U[0] - rY.sum(0)
        return (hf[:-1].reshape(-1, 1, 1) * rY).sum(0) + hf[-1] * rN2

    return (chem(U0) - chem(UN)).sum() * dx * dy


def pressure(mech, U, aux):
    """p = rho R T (aux_func.py:39-51)."""
    invM, _, _ = _consts(mech, U)
    Y = U[4:] / U[0]
    Y9 = torch.cat([Y, 1.0 - Y.sum(0, keepdim=True)], 0)
    R = (invM.reshape(-1, 1, 1) * Y9).sum(0)
    return U[0] * R * aux[1]
