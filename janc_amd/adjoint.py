"""Differentiable steps: torch.autograd wrappers around jh_advance / jh_step_vjp.

``step(sim, U, aux, dt)`` is one fixed-dt step as a ``torch.autograd.Function``
(forward = the FP64 step kernels, backward = the hand-written adjoint
kernels).  ``rollout(sim, U, aux, dt, nsteps, checkpoint_every)`` runs many
steps keeping only every ``checkpoint_every``-th state; its backward replays
each segment from its checkpoint (the RK stages of every step are recomputed
inside ``jh_step_vjp``) and sweeps the step VJPs in reverse.

These replace JAX's reverse mode over the reference's ``advance_one_step``
(set_simulator.py:117-122), the custom VJP of ``get_T_nasa7``
(thermo_model.py:155-190) included.
"""
import torch


def _vjp(sim, U, aux, dt, Ub, ab):
    return sim.step_vjp(U, aux, dt, Ub.contiguous(), ab.contiguous())


class _Step(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U, aux, sim, dt):
        Uo, ao = sim.advance_one_step(U.detach(), aux.detach(), sim.dx, sim.dy, dt)
        ctx.save_for_backward(U.detach(), aux.detach())
        ctx.sim, ctx.dt = sim, dt
        return Uo, ao

    @staticmethod
    def backward(ctx, Ub, ab):
        U, aux = ctx.saved_tensors
        Ub = torch.zeros_like(U) if Ub is None else Ub
        ab = torch.zeros_like(aux) if ab is None else ab
        gU, ga = _vjp(ctx.sim, U, aux, ctx.dt, Ub, ab)
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
        for s in range(nsteps):
            if s % every == 0:
                ckpt.append((U.clone(), aux.clone()))
            U, aux = sim.advance_one_step(U, aux, sim.dx, sim.dy, dt)
        ctx.sim, ctx.dt, ctx.nsteps, ctx.every = sim, dt, nsteps, every
        ctx.ckpt = ckpt
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
            states = [(U, aux)]
            for _ in range(first, last - 1):       # replay the segment
                U, aux = sim.advance_one_step(U, aux, sim.dx, sim.dy, dt)
                states.append((U, aux))
            for Us, As in reversed(states):
                Ub, ab = _vjp(sim, Us, As, dt, Ub, ab)
        ctx.ckpt = None
        return Ub, ab, None, None, None, None


def rollout(sim, U, aux, dt, nsteps, checkpoint_every=10):
    """nsteps differentiable fixed-dt steps with checkpoints every
    ``checkpoint_every`` steps (memory: one state per checkpoint plus one
    segment during the backward sweep)."""
    return _Rollout.apply(U, aux, sim, float(dt), int(nsteps), int(checkpoint_every))
