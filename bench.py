#!/usr/bin/env python3
"""Throughput of the JANC 2-D reacting-Euler step on one MI355X.

One step = the reference's advance_func (simulation/set_simulator.py:444-447):
CFL dt, three SSP-RK3 stages of WENO5 + global Lax-Friedrichs flux
divergence with Halley temperature updates, point-implicit 9-species /
19-reaction H2-air chemistry and the final temperature solve.

Headline (``value``): the FP64 plain step on the 4000 x 400 (1.6M-cell)
periodic-in-x detonation annulus, inputs resident in HBM, K steps timed
between two device synchronisations.  Also reported: the mixed-precision and
FP64 steps on the 8000 x 500 (4M-cell) tube, the per-kernel HIP-event times
with the roofline of the dominant kernel, and the FP64 CPU oracle timed on a
200 x 80 crop of the headline workload on this host.

Prints ONE JSON line.  Single GPU only.
"""
import argparse
import json
import os
import sys
import time

METRIC = ("cell-updates/s per RK3 step w/ chemistry: mixed vs FP64 at stated err; "
          "adjoint/fwd")
HBM_PEAK_GBS = 8000.0
FP64_VECTOR_PEAK_TFLOPS = 78.6

# Algorithmic HBM words per cell per launch (FP64: 8 B/word, mixed: 4 B/word).
#   stage 1: read U (12) + T (1), write U1 (12) + aux (2)         = 27
#   stage 2/3: additionally read U^n (12)                          = 39
#   chemistry: read U (12) + T (1), write rho*Y (8) + aux (2)      = 23
WORDS = {"stage": (27 + 39 + 39) / 3.0, "chem": 23.0, "reduce": 13.0, "bc": 0.0}

# Measured HBM bytes per cell per stage-kernel launch of the FP64 headline
# step (rocprofv3 FETCH_SIZE + WRITE_SIZE passes, each scaled by the same
# counter's reading for a 256 MiB device-to-device copy; summary under
# profiles/).  Reads exceed the algorithmic 25 words: each region cell's
# state is read for its primitives and again per variable.
STAGE_TRAFFIC_B_PER_CELL_FP64 = 459.8

# Measured FP64 FLOP per cell per launch of the FP64 headline kernels
# (rocprofv3 SQ_INSTS_VALU_FLOPS_FP64 x 64 lanes / cells, FMA = 2; summary
# under profiles/): the bench turns them into live FP64-vector fractions.
FP64_FLOP_PER_CELL = {"stage": 4627.0, "chem": 3239.0}


# Per-kernel HIP events bracket the launches of every KERNEL_TIMING_EVERY-th
# step of the timed region: events around every launch add ~5 % to a step.
KERNEL_TIMING_EVERY = 5


def _run(sim, U, aux, steps, warmup, torch):
    """Warm up, then time `steps` fused steps; returns (seconds, per-kernel)."""
    if warmup:
        sim.advance(U, aux, warmup)
    torch.cuda.synchronize()
    sim.timing(U, enable=KERNEL_TIMING_EVERY)
    sim.timing(U)                       # drop anything queued before
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, _, dts = sim.advance(U, aux, steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kt = sim.timing(U)
    sim.timing(U, enable=False)
    flags = sim.health(U)
    return t1 - t0, kt, dts, flags


def _roofline(kt, cells, word_bytes):
    out = {}
    for kind, (ms, n) in kt.items():
        if n == 0 or WORDS.get(kind, 0.0) == 0.0:
            continue
        avg_s = ms / n / 1e3
        bytes_per_launch = WORDS[kind] * word_bytes * cells
        ach = bytes_per_launch / avg_s / 1e9
        out[kind] = {"avg_ms": ms / n, "launches": n, "algorithmic_bytes": bytes_per_launch,
                     "achieved_GBs": ach, "frac_hbm": ach / HBM_PEAK_GBS}
    return out


CPU_CROP = (400, 400)

# flux_config variants (set_flux_solver, solver_2D/flux.py:19-40); the
# headline is the first
SCHEMES = {
    "lf_weno": {"solver_type": "flux_splitting", "interface_reconstruction": "WENO5_JS",
                "split_method": "LF"},
    "hllc_weno": {"solver_type": "godunov", "interface_reconstruction": "WENO5_JS",
                  "riemann_solver": "HLLC"},
    "hllc_muscl": {"solver_type": "godunov", "interface_reconstruction": "MUSCL",
                   "riemann_solver": "HLLC"},
    "lf_muscl": {"solver_type": "flux_splitting", "interface_reconstruction": "MUSCL",
                 "split_method": "LF"},
    "sw_weno": {"solver_type": "flux_splitting", "interface_reconstruction": "WENO5_JS",
                "split_method": "SW"},
}


def _schemes_leg(U0, aux0, geom, theta, steps, warmup, tune, torch):
    """The FP64 step of the headline workload with each other flux scheme."""
    from janc_amd import cases
    from janc_amd.solver import Simulator
    cells = U0.shape[1] * U0.shape[2]
    out = {}
    for name, fc in SCHEMES.items():
        if name == "lf_weno":
            continue
        sim = Simulator(cases.simulation_config(geom, cfl=0.3, precision="fp64", theta=theta,
                                                flux=fc))
        U = torch.as_tensor(U0).cuda()
        aux = torch.as_tensor(aux0).cuda()
        tuned = None
        if tune and fc["solver_type"] == "flux_splitting":
            tuned = sim.autotune(U, aux, theta=theta)
        el, kt, dts, flags = _run(sim, U, aux, steps, warmup, torch)
        out[name] = {"flux_config": fc, "value": cells * steps / el,
                     "ms_per_step": el / steps * 1e3, "steps": steps,
                     "kernels": {k: {"ms_total": v[0], "launches": v[1]} for k, v in kt.items()},
                     "roofline": _roofline(kt, cells, 8), "health_flags": flags,
                     "autotune": tuned}
        del U, aux
        torch.cuda.empty_cache()
    return out


def _cpu_baseline(U, aux, geom, theta, steps):
    """FP64 oracle (numpy, one core) on a 400 x 400 crop of the headline state."""
    from oracle.step import Oracle
    from janc_amd.mechanism import DEFAULT_MECH
    crop = (slice(None), slice(0, CPU_CROP[0]), slice(0, CPU_CROP[1]))
    Uc, ac = U[crop].copy(), aux[crop].copy()
    bcfg = dict(geom["boundary_config"])
    inlets = [k for k, v in bcfg.items() if v == "injection_inlet"]
    for k in inlets:
        bcfg[k] = "slip_wall"
    o = Oracle(DEFAULT_MECH, bcfg, geom["dx"], geom["dy"])
    if inlets:
        from oracle.inlet import attach_inlet
        for k in inlets:
            attach_inlet(o, k.replace("_boundary", ""), theta["P0_inj"], theta["T0_inj"],
                         theta["Yinj"])
    t = 0.0
    o.advance_func(Uc, ac, t, 0.3, theta)          # warm caches
    t0 = time.perf_counter()
    for _ in range(steps):
        Uc, ac, t, _ = o.advance_func(Uc, ac, t, 0.3, theta)
    el = time.perf_counter() - t0
    cells = Uc.shape[1] * Uc.shape[2]
    return {"value": cells * steps / el, "unit": "cell-updates/s", "cores": 1,
            "kind": "port",
            "sample": "%d CFL steps of the FP64 numpy oracle on the %dx%d corner crop "
                      "(seed columns included) of the 4000x400 headline state, same BC kinds; "
                      "%.1f s" % (steps, CPU_CROP[0], CPU_CROP[1], el)}


def _snapshot_leg(U0, aux0, geom, theta, steps, every, torch):
    """HDF5 snapshots of the headline state (H5Saver layout, gzip level 4):
    `steps` FP64 steps with an asynchronous snapshot every `every` steps vs
    the same steps without, and one synchronous save timed alone."""
    import tempfile
    from janc_amd import cases
    from janc_amd.h5 import H5Saver
    from janc_amd.solver import Simulator
    sim = Simulator(cases.simulation_config(geom, cfl=0.3, precision="fp64", theta=theta))
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for mode in ("none", "async"):
            U = torch.as_tensor(U0).cuda()
            aux = torch.as_tensor(aux0).cuda()
            sim.advance(U, aux, 2)
            saver = H5Saver(os.path.join(d, mode + ".h5"), "2D") if mode == "async" else None
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for k in range(0, steps, every):
                sim.advance(U, aux, every)
                if saver is not None:
                    saver.save_async(k // every, 0.0, k + every, u=U)
            torch.cuda.synchronize()
            t_loop = time.perf_counter() - t0
            if saver is not None:
                saver.wait()
            t_all = time.perf_counter() - t0
            out[mode] = {"loop_s": t_loop, "until_written_s": t_all}
            if saver is not None:
                t0 = time.perf_counter()
                saver.save(999, 0.0, 0, u=U)
                out["sync_save_s"] = time.perf_counter() - t0
                saver.close()
                size = os.path.getsize(os.path.join(d, "async.h5"))
                out["file_bytes"] = size
                out["raw_bytes_per_snapshot"] = U.numel() * 8
                out["snapshots"] = steps // every + 1
            del U, aux
    # stepping slowed by the overlapped saves, and the whole including the
    # last snapshot's tail (written after the last step)
    out["loop_overhead_frac"] = out["async"]["loop_s"] / out["none"]["loop_s"] - 1.0
    out["overhead_frac"] = out["async"]["until_written_s"] / out["none"]["loop_s"] - 1.0
    out["workload"] = ("%d FP64 steps of the 4000x400 annulus, u (12x4000x400 float64) saved "
                       "every %d steps with H5Saver.save_async (pinned D2H on a side stream, "
                       "gzip on 16 host threads)" % (steps, every))
    return out


def _fields(U, aux, inv_M):
    """rho, p, T, Y_k (float64 torch) of a state."""
    import torch
    U = U.double()
    aux = aux.double()
    rho = U[0]
    Y = U[4:] / rho
    Y9 = torch.cat([Y, (1.0 - Y.sum(0))[None]])
    R = (inv_M[:, None, None] * Y9).sum(0)
    return {"rho": rho, "p": rho * R * aux[1], "T": aux[1], "Y": Y}


def _accuracy(mech, steps, every, torch):
    """Mixed vs FP64 step on the 8000 x 500 tube at one fixed dt (the initial
    CFL-0.3 dt): relative L2 and Linf of rho, p, T, Y_k every `every` steps."""
    from janc_amd import cases
    from janc_amd.solver import Simulator
    U3, aux3, g3 = cases.mixed_tube(mech)
    probe = Simulator(cases.simulation_config(g3, cfl=0.3))
    dt = float(probe.CFL(U3, aux3, 0.3))
    del probe
    inv_M = torch.as_tensor(1.0 / mech.species_M).cuda()
    sims, states = {}, {}
    for prec, dtype in (("mixed", torch.float32), ("fp64", torch.float64)):
        sims[prec] = Simulator(cases.simulation_config(g3, dt=dt, precision=prec))
        states[prec] = (torch.as_tensor(U3, dtype=dtype).cuda(),
                        torch.as_tensor(aux3, dtype=dtype).cuda())
    report = []
    done = 0
    while done < steps:
        n = min(every, steps - done)
        for prec in sims:
            sims[prec].advance(*states[prec], n)
        done += n
        fm = _fields(*states["mixed"], inv_M)
        fd = _fields(*states["fp64"], inv_M)
        row = {"step": done}
        for k in ("rho", "p", "T"):
            d = (fm[k] - fd[k])
            row[k] = {"L2": float(d.norm() / fd[k].norm()),
                      "Linf": float(d.abs().max() / fd[k].abs().max())}
        ys = {}
        for i, name in enumerate(mech.species[:-1]):
            d = fm["Y"][i] - fd["Y"][i]
            ys[name] = {"L2": float(d.norm() / max(float(fd["Y"][i].norm()), 1e-300)),
                        "Linf": float(d.abs().max() / max(float(fd["Y"][i].abs().max()), 1e-300)),
                        "max": float(fd["Y"][i].abs().max())}
        row["Y"] = ys
        report.append(row)
    flags = {p: sims[p].health(states[p][0]) for p in sims}
    return {"workload": "8000x500 tube (dx 1.25e-5 m), fixed dt %.4g s, mixed vs FP64 GPU "
                        "step (the FP64 step is parity-checked against the oracle)" % dt,
            "dt": dt, "steps": steps, "health_flags": flags, "every": report}


def _batch_leg(mech, steps, tune, torch):
    """64 uneven mixed-precision tubes advanced together (batched launches
    captured in one HIP graph over `steps` steps) vs the same cases advanced
    one after another, each in one fused library call."""
    from janc_amd import cases
    from janc_amd.batch import BatchSimulator
    from janc_amd.solver import Simulator
    cs = cases.batch_cases(mech)
    cfgs = [cases.simulation_config(g, cfl=0.3, precision="mixed") for _, _, g in cs]
    f32 = torch.float32
    U0 = [torch.as_tensor(U, dtype=f32).cuda() for U, _, _ in cs]
    A0 = [torch.as_tensor(a, dtype=f32).cuda() for _, a, _ in cs]
    cells = sum(U.shape[1] * U.shape[2] for U in U0)
    Us = [U.clone() for U in U0]
    As = [a.clone() for a in A0]
    b = BatchSimulator(cfgs, Us, As)
    tuned = b.autotune(steps=steps) if tune else None
    for U, a, u0, a0 in zip(Us, As, U0, A0):
        U.copy_(u0)
        a.copy_(a0)
    b.advance(steps, graph=True)            # capture + first replay
    for U, a, u0, a0 in zip(Us, As, U0, A0):
        U.copy_(u0)
        a.copy_(a0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.advance(steps, graph=True)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    flags = b.health()
    # sequential: one Simulator per case, same tile variant, warm contexts
    sims = [Simulator(c) for c in cfgs]
    Uq = [U.clone() for U in U0]
    Aq = [a.clone() for a in A0]
    for sim, U, a in zip(sims, Uq, Aq):
        sim.set_tile(U, b.get_tile())
        Uw, aw = U.clone(), a.clone()
        sim.advance(Uw, aw, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for sim, U, a in zip(sims, Uq, Aq):
        sim.advance(U, a, steps)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    same = all(torch.equal(x, y) for x, y in zip(Us, Uq))
    return {"workload": "64 tubes, grids drawn from {128x32, 256x64, 512x128, 1024x256}, "
                        "phi U[0.6,1.4], p0 U[0.5,3] atm, driver T U[2000,3200] K, mixed "
                        "precision, CFL 0.3, %d steps in one HIP graph" % steps,
            "cases": len(cs), "cells": cells,
            "value": cells * steps / t_batch, "unit": "cell-updates/s",
            "ms_per_step": t_batch / steps * 1e3,
            "sequential_value": cells * steps / t_seq,
            "speedup_vs_sequential": t_seq / t_batch,
            "bit_identical_to_sequential": bool(same),
            "tile_variant": b.get_tile(), "autotune": tuned, "health_flags": flags}


def _adjoint_leg(mech, steps, every, ndir, torch):
    """Config 4: J = heat released + 1/2 |p_final/p_ref - p_target/p_ref|^2 /
    cells after `steps` FP64 steps of a 1024 x 256 tube, gradients with
    respect to the initial T and Y_k fields through the checkpointed adjoint
    rollout; timed against the forward rollout, and checked against central
    differences (eps 1e-6) on `ndir` seeded random directions."""
    import numpy as np
    from janc_amd import adjoint, cases
    from janc_amd.solver import Simulator
    T, p, Y, u, v, g = cases.adjoint_case(mech)
    dev = "cuda"
    Tt, pt, Yt = (torch.tensor(x, device=dev) for x in (T, p, Y))
    ut, vt = torch.tensor(u, device=dev), torch.tensor(v, device=dev)
    U0, a0 = adjoint.initial_state(mech, Tt, pt, Yt, ut, vt)
    probe = Simulator(cases.simulation_config(g, cfl=0.3, precision="fp64"))
    dt = 0.8 * float(probe.CFL(U0, a0))
    sim = Simulator(cases.simulation_config(g, dt=dt, precision="fp64"))
    cells = T.size
    dx, dy = g["dx"], g["dy"]
    UN, aN = sim.advance(U0.clone(), a0.clone(), steps)[:2]
    rng = np.random.default_rng(4)
    s_field = torch.tensor(cases.smooth_field(rng, *T.shape), device=dev)
    p_ref = 101325.0
    p_target = adjoint.pressure(mech, UN, aN) * (1.0 + 0.01 * s_field)
    q_scale = float(adjoint.heat_release(mech, U0, UN, dx, dy).abs()) + 1e-300

    def J(U0, UN, aN):
        q = adjoint.heat_release(mech, U0, UN, dx, dy) / q_scale
        r = (adjoint.pressure(mech, UN, aN) - p_target) / p_ref
        return q + 0.5 * (r * r).sum() / cells

    def J_fwd(Tx, Yx):
        U0, a0 = adjoint.initial_state(mech, Tx, pt, Yx, ut, vt)
        UN, aN = sim.advance(U0.clone(), a0.clone(), steps)[:2]
        return float(J(U0, UN, aN))

    # forward rollout alone (fused library call)
    Uf, af = U0.clone(), a0.clone()
    sim.advance(Uf, af, 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.advance(Uf, af, steps)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    # forward + adjoint
    Tg = Tt.clone().requires_grad_(True)
    Yg = Yt.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    U0g, a0g = adjoint.initial_state(mech, Tg, pt, Yg, ut, vt)
    UNg, aNg = adjoint.rollout(sim, U0g, a0g, dt, steps, checkpoint_every=every)
    torch.cuda.synchronize()
    t_roll = time.perf_counter() - t0
    J(U0g, UNg, aNg).backward()
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t0
    # one step: forward vs VJP
    W = torch.randn(U0.shape, dtype=torch.float64, device=dev,
                    generator=torch.Generator(device=dev).manual_seed(3))
    Wa = torch.zeros_like(a0)
    Uw, aw = U0.clone(), a0.clone()
    sim.step_vjp(U0, a0, dt, W, Wa)
    torch.cuda.synchronize()
    reps = 10
    t0 = time.perf_counter()
    sim.advance(Uw, aw, reps)
    torch.cuda.synchronize()
    t_fwd1 = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.step_vjp(U0, a0, dt, W, Wa)
    torch.cuda.synchronize()
    t_vjp1 = (time.perf_counter() - t0) / reps
    # the VJP as adjoint.step's backward runs it: the step's output is known
    U1s, a1s = sim.advance_one_step(U0, a0, dx, dy, dt)
    sim.step_vjp(U0, a0, dt, W, Wa, known_output=(U1s, a1s))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.step_vjp(U0, a0, dt, W, Wa, known_output=(U1s, a1s))
    torch.cuda.synchronize()
    t_vjpk = (time.perf_counter() - t0) / reps
    # central differences on random directions: two-point at the
    # configuration's eps = 1e-6 and at 1e-7, and four-point at 1e-6.  The
    # two-point quotient's own truncation error (chemistry makes J strongly
    # nonlinear in T) falls as eps^2 while the rounding of J grows as 1/eps;
    # the four-point quotient's truncation error falls as eps^4
    errs = {"1e-06": [], "1e-07": [], "1e-06_4pt": []}
    for d in range(ndir):
        dT = Tt * torch.tensor(rng.standard_normal(T.shape), device=dev)
        dY = Yt * torch.tensor(rng.standard_normal(Y.shape), device=dev)
        ad = float((Tg.grad * dT).sum() + (Yg.grad * dY).sum())
        Jp = {}
        for k in (-2, -1, 1, 2):
            Jp[k] = J_fwd(Tt + k * 1e-6 * dT, Yt + k * 1e-6 * dY)
        fds = {"1e-06": (Jp[1] - Jp[-1]) / 2e-6,
               "1e-07": (J_fwd(Tt + 1e-7 * dT, Yt + 1e-7 * dY) -
                         J_fwd(Tt - 1e-7 * dT, Yt - 1e-7 * dY)) / 2e-7,
               "1e-06_4pt": (8.0 * (Jp[1] - Jp[-1]) - (Jp[2] - Jp[-2])) / 12e-6}
        for key, fd in fds.items():
            errs[key].append(abs(fd - ad) / max(abs(fd), abs(ad), 1e-300))
    return {"workload": "1024x256 tube (dx 1.25e-5 m, 51-column 30 atm/3000 K driver, 0.5%% T "
                        "and 3 m/s velocity noise), FP64, fixed dt %.3e s, %d steps, checkpoints "
                        "every %d; J = heat release + pressure mismatch; grad wrt initial T, Y_k"
                        % (dt, steps, every),
            "cells": cells,
            "step_fwd_ms": t_fwd1 * 1e3, "step_vjp_ms": t_vjpk * 1e3,
            "step_vjp_recompute_all_ms": t_vjp1 * 1e3,
            "adjoint_over_fwd_step": t_vjpk / t_fwd1,
            "rollout_fwd_ms": t_fwd * 1e3,
            "rollout_checkpointed_fwd_ms": t_roll * 1e3,
            "rollout_fwd_plus_adjoint_ms": t_total * 1e3,
            "adjoint_over_fwd_rollout": (t_total - t_roll) / t_fwd,
            "fd_directions": ndir,
            "fd_max_rel_err": {e: max(v) for e, v in errs.items()} if ndir else None,
            "fd_median_rel_err": ({e: float(np.median(v)) for e, v in errs.items()}
                                  if ndir else None)}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--mixed-steps", type=int, default=100)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--accuracy-steps", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the tile autotuner")
    ap.add_argument("--tile", type=int, default=None,
                    help="force the headline stage-kernel variant (skips its autotuning)")
    ap.add_argument("--batch-steps", type=int, default=50)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=100)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--fd-directions", type=int, default=32)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="one 256 MiB device-to-device copy before the timed region "
                         "(reference dispatch for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the FP64 headline workload (for profiler runs)")
    ap.add_argument("--flux", choices=list(SCHEMES), default="lf_weno",
                    help="flux scheme of the headline workload (default: the reference's "
                         "LF flux splitting with WENO5_JS)")
    ap.add_argument("--scheme-steps", type=int, default=50)
    ap.add_argument("--snapshot-steps", type=int, default=600)
    ap.add_argument("--no-snapshots", action="store_true")
    ap.add_argument("--no-schemes", action="store_true")
    args = ap.parse_args()
    if args.headline_only:
        args.no_mixed = args.no_cpu = args.no_batch = args.no_adjoint = True
        args.no_schemes = args.no_snapshots = True
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the single-block step runs on one GPU; --gpus N > 1 is not "
              "supported (the x-slab decomposition is not built)", file=sys.stderr)
        sys.exit(2)

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import numpy as np
    import torch
    from janc_amd import cases
    from janc_amd.mechanism import Mechanism
    from janc_amd.solver import Simulator

    mech = Mechanism()
    # ---- headline: FP64 on the 4000 x 400 annulus ----
    U0, aux0, geom, theta = cases.rdc_annulus(mech)
    nx, ny = U0.shape[1], U0.shape[2]
    sim = Simulator(cases.simulation_config(geom, cfl=0.3, precision="fp64", theta=theta,
                                            flux=SCHEMES[args.flux]))
    U = torch.as_tensor(U0).cuda()
    aux = torch.as_tensor(aux0).cuda()
    if args.flux != "lf_weno":
        tuned = None
    elif args.tile is not None:
        sim.set_tile(U, args.tile, theta)
        tuned = {"variant": args.tile, "forced": True}
    else:
        tuned = sim.autotune(U, aux, theta=theta) if not args.no_tune else None
    if args.calibrate_copy:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    el, kt, dts, flags = _run(sim, U, aux, args.steps, args.warmup, torch)
    cells = nx * ny
    value = cells * args.steps / el
    roof = _roofline(kt, cells, 8)
    dom = max((k for k in roof), key=lambda k: kt[k][0])
    r = roof[dom]
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "FP64 plain step, 4000x400 periodic-in-x detonation annulus "
                               "(ZND-like seed 200 columns 30 atm/3000 K, 0.5% T noise), "
                               "9sp-19r H2-air, CFL 0.3; bottom injection inlet (0.2 MPa / "
                               "300 K manifold), top pressure outlet 1 atm",
                   "flux_config": SCHEMES[args.flux],
                   "nx": nx, "ny": ny, "cells": cells, "precision": "fp64",
                   "dt_first": float(dts[0]), "dt_last": float(dts[-1]),
                   "health_flags": flags},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": r["achieved_GBs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": r["frac_hbm"],
                     "traffic": (STAGE_TRAFFIC_B_PER_CELL_FP64 * cells
                                 if dom == "stage" and args.flux == "lf_weno" else None),
                     "algorithmic_bytes": r["algorithmic_bytes"]},
        "kernels_fp64": {k: {"ms_total": v[0], "launches": v[1]} for k, v in kt.items()},
        "kernel_timing": ("HIP events on the context stream around the launches of every "
                          "%dth step of the timed region" % KERNEL_TIMING_EVERY),
        "autotune_fp64": tuned,
        "roofline_all_fp64": roof,
    }
    if args.flux == "lf_weno":
        fv = {}
        for kind, fpc in FP64_FLOP_PER_CELL.items():
            ms, nl = kt.get(kind, (0.0, 0))
            if nl:
                tfs = fpc * cells / (ms / nl * 1e-3) / 1e12
                fv[kind] = {"flop_per_cell": fpc, "avg_ms": ms / nl, "achieved_TFLOPs": tfs,
                            "peak_TFLOPs": FP64_VECTOR_PEAK_TFLOPS,
                            "frac": tfs / FP64_VECTOR_PEAK_TFLOPS}
        result["fp64_vector"] = fv
    del U, aux
    torch.cuda.empty_cache()

    # ---- the other flux schemes on the headline workload ----
    if not args.no_schemes:
        result["schemes_fp64"] = _schemes_leg(U0, aux0, geom, theta, args.scheme_steps,
                                              args.warmup, not args.no_tune, torch)

    # ---- HDF5 snapshots of the headline state ----
    if not args.no_snapshots:
        result["snapshots"] = _snapshot_leg(U0, aux0, geom, theta, args.snapshot_steps, 200,
                                            torch)
        torch.cuda.empty_cache()

    # ---- mixed vs FP64 on the 8000 x 500 tube ----
    if not args.no_mixed:
        U3, aux3, g3 = cases.mixed_tube(mech)
        n3 = U3.shape[1] * U3.shape[2]
        out = {}
        for prec, dtype in (("mixed", torch.float32), ("fp64", torch.float64)):
            s3 = Simulator(cases.simulation_config(g3, cfl=0.3, precision=prec))
            Ut = torch.as_tensor(U3, dtype=dtype).cuda()
            at = torch.as_tensor(aux3, dtype=dtype).cuda()
            t3 = s3.autotune(Ut, at) if not args.no_tune else None
            e3, k3, d3, f3 = _run(s3, Ut, at, args.mixed_steps, args.warmup, torch)
            out[prec] = {"value": n3 * args.mixed_steps / e3, "ms_per_step": e3 / args.mixed_steps * 1e3,
                         "kernels": {k: {"ms_total": v[0], "launches": v[1]} for k, v in k3.items()},
                         "roofline": _roofline(k3, n3, 4 if prec == "mixed" else 8),
                         "health_flags": f3, "autotune": t3}
            del Ut, at
            torch.cuda.empty_cache()
        out["speedup_mixed_over_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
        out["workload"] = ("8000x500 tube (dx 1.25e-5 m), phi in [0.8,1.2] smooth field, "
                           "40-column 20 atm/2800 K ignition slab, %d CFL steps"
                           % args.mixed_steps)
        result["mixed_tube"] = out

    if not args.no_mixed and args.accuracy_steps > 0:
        result["mixed_accuracy"] = _accuracy(mech, args.accuracy_steps, 100, torch)

    if not args.no_batch:
        result["batch"] = _batch_leg(mech, args.batch_steps, not args.no_tune, torch)
        torch.cuda.empty_cache()
    if not args.no_adjoint:
        result["adjoint"] = _adjoint_leg(mech, args.adjoint_steps, 10, args.fd_directions, torch)
        torch.cuda.empty_cache()

    if not args.no_cpu:
        result["cpu_baseline"] = _cpu_baseline(U0, aux0, geom, theta, args.cpu_steps)
        result["speedup_vs_cpu"] = value / result["cpu_baseline"]["value"]
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
