#!/usr/bin/env python3
"""Throughput of the JANC 2-D reacting-Euler step on one MI355X.

One step = the reference's advance_func (simulation/set_simulator.py:444-447):
CFL dt, three SSP-RK3 stages of WENO5 + global Lax-Friedrichs flux
divergence with Halley temperature updates, point-implicit 9-species /
19-reaction H2-air chemistry and the final temperature solve.

Headline (``value``): the mixed-precision step (FP32 state, split fluxes and
WENO; FP64 divergence, RK combine, chemistry and temperature solve) on the
8000 x 500 (4M-cell) tube with a smooth equivalence-ratio field and a
40-column 20 atm / 2800 K ignition slab; inputs resident in HBM, K steps
timed between two device synchronisations.  Beside it on the same workload:
the FP64 step, the mixed form's error after 1000 steps (against the FP64
step on the full tube and against the FP64 CPU oracle on a crop), the
adjoint's cost per forward step, the batch and annulus legs, the roofline of
the dominant kernel and the CPU oracle timed on this host's cores.

Every leg prints one JSON line of its own (``{"leg": ...}``) as it finishes;
the LAST line is the compact summary (a few KB).  Single GPU only.
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

# Measured HBM bytes per cell per stage-kernel launch (rocprofv3 FETCH_SIZE +
# WRITE_SIZE passes, each scaled by the same counter's reading for a 256 MiB
# device-to-device copy; summaries under profiles/): FP64 = the 4000 x 400
# annulus step (tile variant 3), mixed = the 8000 x 500 tube (variant 1).
# Reads exceed the algorithmic words: each region cell's state is read for
# its primitives and again per variable.
STAGE_TRAFFIC_B_PER_CELL = {"fp64": 459.8, "mixed": 210.4}

# Measured FP64 FLOP per cell per chemistry launch (rocprofv3
# SQ_INSTS_VALU_FLOPS_FP64 x 64 lanes / cells, FMA = 2; summary under
# profiles/): rates, exponentials, LU and temperature solve are FP64 in both
# forms; the mixed form assembles its Jacobian in FP32 (368 FP32 FLOP more).
CHEM_FP64_FLOP_PER_CELL = {"mixed": 2760.2, "fp64": 3249.6}

# Measured VALU wave instructions x 64 per cell per stage-kernel launch
# (rocprofv3 SQ_INSTS_VALU; mixed form, tile variant 1; summary under
# profiles/).  The VALU issue floor of a launch is instructions / 64 wave
# instructions x 4 cycles (64 lanes on a 16-lane SIMD; FP32 without packed
# math issues at the FP64 rate) / (1024 SIMDs x 2.4 GHz).
STAGE_VALU_PER_CELL = {"mixed": 3209.0}
SIMDS = 1024
CLOCK_GHZ = 2.4

# Per-kernel HIP events bracket the launches of every KERNEL_TIMING_EVERY-th
# step of the timed region: events around every launch add ~5 % to a step.
KERNEL_TIMING_EVERY = 5


def emit(leg, obj):
    """One detail line (the summary line comes last)."""
    print(json.dumps(dict({"leg": leg}, **obj), default=float), flush=True)


def _run(sim, U, aux, steps, warmup, torch):
    """Warm up, then time `steps` fused steps; returns (seconds, per-kernel,
    dt history, health flags)."""
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


def _kernels(kt):
    return {k: {"ms_total": v[0], "launches": v[1]} for k, v in kt.items()}


# ------------------------------------------------------------------ CPU legs
# The CPU work runs in worker processes forked before this process touches
# the GPU, so no worker inherits a GPU handle.

_CPU = {}      # state shared with the forked workers (set before the fork)

# FP64 oracle crop of the tube for the mixed form's accuracy check: the
# ignition slab and the phi field's variation are inside it.
ORACLE_CROP = (200, 50)
# CPU baseline: each worker advances its own crop of the tube.
CPU_CROP = (250, 125)


def _tube_bc():
    """Boundary kinds of the tube configurations (janc_amd.cases.detonation_tube)."""
    return {"left_boundary": "slip_wall", "right_boundary": "zero_gradient",
            "bottom_boundary": "slip_wall", "top_boundary": "slip_wall"}


def _cpu_worker(job):
    """Time `steps` CFL steps of the FP64 numpy oracle on one crop."""
    x0, y0, nxc, nyc, steps = job
    from threadpoolctl import threadpool_limits
    from oracle.step import Oracle
    from janc_amd.mechanism import DEFAULT_MECH
    with threadpool_limits(1):
        U = _CPU["U"][:, x0:x0 + nxc, y0:y0 + nyc].copy()
        aux = _CPU["aux"][:, x0:x0 + nxc, y0:y0 + nyc].copy()
        o = Oracle(DEFAULT_MECH, _tube_bc(), _CPU["dx"], _CPU["dx"])
        t = 0.0
        U, aux, t, _ = o.advance_func(U, aux, t, 0.3)          # warm caches
        t0 = time.perf_counter()
        for _ in range(steps):
            U, aux, t, _ = o.advance_func(U, aux, t, 0.3)
        t1 = time.perf_counter()
    return nxc * nyc * steps, t0, t1


def _oracle_crop_worker(job):
    """FP64 oracle on the accuracy crop at a fixed dt; states every `every` steps."""
    nxc, nyc, dt, steps, every = job
    from threadpoolctl import threadpool_limits
    from oracle.step import Oracle
    from janc_amd.mechanism import DEFAULT_MECH
    import numpy as np
    with threadpool_limits(1):
        U = _CPU["U"][:, :nxc, :nyc].copy()
        aux = _CPU["aux"][:, :nxc, :nyc].copy()
        o = Oracle(DEFAULT_MECH, _tube_bc(), _CPU["dx"], _CPU["dx"])
        out = []
        t0 = time.perf_counter()
        for s in range(1, steps + 1):
            U, aux = o.advance_one_step(U, aux, dt)
            if s % every == 0:
                out.append((s, np.array(U), np.array(aux)))
    return out, time.perf_counter() - t0


def _cpu_baseline(pool, nworkers, steps):
    jobs = []
    for w in range(nworkers):
        jobs.append(((w // 4) * CPU_CROP[0], (w % 4) * CPU_CROP[1], CPU_CROP[0], CPU_CROP[1],
                     steps))
    res = pool.map(_cpu_worker, jobs, chunksize=1)
    return _cpu_result(res, nworkers, steps)


def _cpu_result(res, nworkers, steps):
    """The cpu_baseline object from the workers' (cell-updates, start, end)."""
    work = sum(r[0] for r in res)
    wall = max(r[2] for r in res) - min(r[1] for r in res)
    per = [r[0] / (r[2] - r[1]) for r in res]
    return {"value": work / wall, "unit": "cell-updates/s", "cores": nworkers, "kind": "port",
            "per_process_value": sum(per) / len(per),
            "sample": "%d processes (one core each), each %d CFL steps of the FP64 numpy oracle "
                      "on its own %dx%d crop of the 8000x500 tube state (crops tile x<%d, "
                      "y<500; slab in crop 0); value = all crops' cell-updates / wall time "
                      "from the first start to the last end (%.1f s)"
                      % (nworkers, steps, CPU_CROP[0], CPU_CROP[1],
                         CPU_CROP[0] * ((nworkers + 3) // 4), wall)}


# ------------------------------------------------------------------ GPU legs

def _fields(U, aux, inv_M):
    """rho, p, T, Y_k (float64 torch) of a state."""
    import torch
    U = torch.as_tensor(U).double()
    aux = torch.as_tensor(aux).double()
    inv_M = torch.as_tensor(inv_M, device=U.device)
    rho = U[0]
    Y = U[4:] / rho
    Y9 = torch.cat([Y, (1.0 - Y.sum(0))[None]])
    R = (inv_M[:, None, None] * Y9).sum(0)
    return {"rho": rho, "p": rho * R * aux[1], "T": aux[1], "Y": Y}


def _errors(fm, fr, species):
    """Relative L2 and Linf of rho, p, T and every Y_k (fm against fr)."""
    row = {}
    for k in ("rho", "p", "T"):
        d = fm[k] - fr[k]
        row[k] = {"L2": float(d.norm() / fr[k].norm()),
                  "Linf": float(d.abs().max() / fr[k].abs().max())}
    ys = {}
    for i, name in enumerate(species):
        d = fm["Y"][i] - fr["Y"][i]
        ys[name] = {"L2": float(d.norm() / max(float(fr["Y"][i].norm()), 1e-300)),
                    "Linf": float(d.abs().max() / max(float(fr["Y"][i].abs().max()), 1e-300)),
                    "max": float(fr["Y"][i].abs().max())}
    row["Y"] = ys
    return row


def _worst(row):
    """Largest relative L2 of T and over the Y_k of one error row."""
    yk = max(row["Y"], key=lambda s: row["Y"][s]["L2"])
    return {"T_L2": row["T"]["L2"], "rho_L2": row["rho"]["L2"], "p_L2": row["p"]["L2"],
            "Y_L2_max": row["Y"][yk]["L2"], "Y_L2_max_species": yk,
            "T_Linf": row["T"]["Linf"],
            "Y_Linf_max": max(v["Linf"] for v in row["Y"].values())}


def _headline(mech, U3, aux3, g3, steps, warmup, tune, torch, tile=None):
    """Mixed and FP64 steps on the 8000 x 500 tube, `steps` CFL steps each."""
    from janc_amd import cases
    from janc_amd.solver import Simulator
    n3 = U3.shape[1] * U3.shape[2]
    out = {}
    for prec, dtype in (("mixed", torch.float32), ("fp64", torch.float64)):
        s3 = Simulator(cases.simulation_config(g3, cfl=0.3, precision=prec))
        Ut = torch.as_tensor(U3, dtype=dtype).cuda()
        at = torch.as_tensor(aux3, dtype=dtype).cuda()
        if tile is not None:
            s3.set_tile(Ut, tile)
            t3 = {"variant": tile, "forced": True}
        else:
            t3 = s3.autotune(Ut, at) if tune else None
        e3, k3, d3, f3 = _run(s3, Ut, at, steps, warmup, torch)
        out[prec] = {"value": n3 * steps / e3, "ms_per_step": e3 / steps * 1e3,
                     "kernels": _kernels(k3), "kt": k3,
                     "roofline": _roofline(k3, n3, 4 if prec == "mixed" else 8),
                     "health_flags": f3, "dt_first": float(d3[0]), "dt_last": float(d3[-1]),
                     "autotune": t3}
        del Ut, at, s3
        torch.cuda.empty_cache()
    return out


def _accuracy(mech, U3, aux3, g3, steps, every, torch):
    """Mixed vs FP64 step on the 8000 x 500 tube at one fixed dt (the initial
    CFL-0.3 dt): relative L2 and Linf of rho, p, T, Y_k every `every` steps."""
    from janc_amd import cases
    from janc_amd.solver import Simulator
    probe = Simulator(cases.simulation_config(g3, cfl=0.3))
    dt = float(probe.CFL(U3, aux3, 0.3))
    del probe
    inv_M = 1.0 / mech.species_M
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
        row = _errors(_fields(*states["mixed"], inv_M), _fields(*states["fp64"], inv_M),
                      mech.species[:-1])
        row["step"] = done
        report.append(row)
    flags = {p: sims[p].health(states[p][0]) for p in sims}
    return {"workload": "8000x500 tube (dx 1.25e-5 m), fixed dt %.4g s, mixed vs FP64 GPU "
                        "step (the FP64 step is parity-checked against the oracle)" % dt,
            "dt": dt, "steps": steps, "health_flags": flags, "every": report}


def _accuracy_crop_gpu(mech, U3, aux3, g3, dt, steps, every, torch):
    """The mixed step on the oracle's crop at the same fixed dt; host copies
    of the state every `every` steps."""
    from janc_amd import cases
    from janc_amd.solver import Simulator
    nxc, nyc = ORACLE_CROP
    geo = dict(g3)
    sim = Simulator(cases.simulation_config(geo, dt=dt, precision="mixed"))
    U = torch.as_tensor(U3[:, :nxc, :nyc].copy(), dtype=torch.float32).cuda()
    a = torch.as_tensor(aux3[:, :nxc, :nyc].copy(), dtype=torch.float32).cuda()
    out = {}
    done = 0
    while done < steps:
        sim.advance(U, a, every)
        done += every
        out[done] = (U.double().cpu(), a.double().cpu())
    return out, sim.health(U)


def _accuracy_crop_compare(mech, gpu_states, oracle_states, steps, dt, el):
    inv_M = 1.0 / mech.species_M
    rows = []
    for s, Uo, ao in oracle_states:
        Ug, ag = gpu_states[s]
        row = _errors(_fields(Ug, ag, inv_M), _fields(Uo, ao, inv_M), mech.species[:-1])
        row["step"] = s
        rows.append(row)
    return {"workload": "%dx%d corner crop of the 8000x500 tube (slab and phi field "
                        "included), fixed dt %.4g s, %d steps: mixed GPU step vs the FP64 "
                        "numpy oracle (%.1f s on one core)" % (ORACLE_CROP + (dt, steps, el)),
            "steps": steps, "every": rows}


def _annulus_leg(mech, steps, warmup, tune, torch):
    """FP64 step on the 4000 x 400 periodic-in-x detonation annulus."""
    from janc_amd import cases
    from janc_amd.solver import Simulator
    U0, aux0, geom, theta = cases.rdc_annulus(mech)
    nx, ny = U0.shape[1], U0.shape[2]
    sim = Simulator(cases.simulation_config(geom, cfl=0.3, precision="fp64", theta=theta))
    U = torch.as_tensor(U0).cuda()
    aux = torch.as_tensor(aux0).cuda()
    tuned = sim.autotune(U, aux, theta=theta) if tune else None
    el, kt, dts, flags = _run(sim, U, aux, steps, warmup, torch)
    cells = nx * ny
    roof = _roofline(kt, cells, 8)
    out = {"workload": "FP64 plain step, 4000x400 periodic-in-x detonation annulus (ZND-like "
                       "seed 200 columns 30 atm/3000 K, 0.5% T noise), CFL 0.3; bottom "
                       "injection inlet (0.2 MPa / 300 K manifold), top pressure outlet 1 atm",
           "value": cells * steps / el, "ms_per_step": el / steps * 1e3, "steps": steps,
           "cells": cells, "kernels": _kernels(kt), "roofline": roof, "health_flags": flags,
           "dt_first": float(dts[0]), "dt_last": float(dts[-1]), "autotune": tuned}
    tr = STAGE_TRAFFIC_B_PER_CELL["fp64"]
    if "stage" in roof and tr:
        out["stage_traffic_bytes"] = tr * cells
    del U, aux
    torch.cuda.empty_cache()
    return out, (U0, aux0, geom, theta)


# flux_config variants (set_flux_solver, solver_2D/flux.py:19-40)
SCHEMES = {
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
    """The FP64 annulus step with each other flux scheme."""
    from janc_amd import cases
    from janc_amd.solver import Simulator
    cells = U0.shape[1] * U0.shape[2]
    out = {}
    for name, fc in SCHEMES.items():
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
                     "kernels": _kernels(kt), "roofline": _roofline(kt, cells, 8),
                     "health_flags": flags, "autotune": tuned}
        del U, aux
        torch.cuda.empty_cache()
    return out
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

    def J_fwd(Tx, Yx, trace=False, clip=None):
        """J after the rollout; with trace, also the per-step records of
        where the LF maxima are attained and which clip branches are active
        (Simulator.trace); clip = ('record' | 'replay', masks) records the
        chemistry clip's branches or makes the rollout take recorded ones
        (Simulator.clip_branches)."""
        U0, a0 = adjoint.initial_state(mech, Tx, pt, Yx, ut, vt)
        Ur, ar = U0.clone(), a0.clone()
        rec = None
        if trace:
            rec = sim.trace(Ur, torch.zeros((steps, 22), dtype=torch.int64, device=dev))
        if clip is not None:
            sim.clip_branches(Ur, clip[1], clip[0])
        UN, aN = sim.advance(Ur, ar, steps)[:2]
        if clip is not None:
            sim.clip_branches(Ur, None, "off")
        if trace:
            sim.trace(Ur, None)
            return float(J(U0, UN, aN)), rec.cpu().numpy()
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
    # the VJP as adjoint.step runs it: the forward keeps the step's residuals
    # (jh_step_keep: RK stage states and LF maxima) and the backward
    # recomputes nothing (jh_step_vjp_res).  Keeping costs the forward a copy
    # of the stage-3 state (the chemistry runs in place): its cost over a
    # plain one-step call is charged to the adjoint.
    U1r, a1r, res = sim.step_keep(U0, a0, dt)
    sim.step_vjp(U0, a0, dt, W, Wa, known_output=(U1r, a1r), residuals=res)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.advance_one_step(U0, a0, dx, dy, dt)
    torch.cuda.synchronize()
    t_one = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.step_keep(U0, a0, dt)
    torch.cuda.synchronize()
    t_keep = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.step_vjp(U0, a0, dt, W, Wa, known_output=(U1r, a1r), residuals=res)
    torch.cuda.synchronize()
    t_vjpr = (time.perf_counter() - t0) / reps
    t_adj = t_vjpr + max(0.0, t_keep - t_one)
    # central differences on random directions: two-point at the
    # configuration's eps = 1e-6 and at 1e-7, and four-point at 1e-6.  The
    # two-point quotient's own truncation error (chemistry makes J strongly
    # nonlinear in T) falls as eps^2 while the rounding of J grows as 1/eps;
    # the four-point quotient's truncation error falls as eps^4
    # the unperturbed trajectory's branches (where each stage's LF maxima
    # are attained, the active clip branches)
    _, rec0 = J_fwd(Tt, Yt, trace=True)
    # ... and every cell's clip branches along it, replayed by the
    # branch-frozen runs: the function the adjoint differentiates
    masks = torch.zeros((steps, cells), dtype=torch.int16, device=dev)
    J_fwd(Tt, Yt, clip=("record", masks))
    errs = {"1e-06": [], "1e-07": [], "1e-06_4pt": [], "frozen_1e-06": [],
            "frozen_1e-06_4pt": []}
    switches = []
    # record columns: per stage (x index, x count, y index, y count) of the
    # LF maxima; clip hash, clip count; per temperature solve (hash, count)
    # of the cells below the NASA-7 switch
    LF, CLIP, RANGE = slice(0, 12), slice(12, 14), slice(14, 22)
    trunc = []
    for d in range(ndir):
        dT = Tt * torch.tensor(rng.standard_normal(T.shape), device=dev)
        dY = Yt * torch.tensor(rng.standard_normal(Y.shape), device=dev)
        ad = float((Tg.grad * dT).sum() + (Yg.grad * dY).sum())
        Jp = {k: J_fwd(Tt + k * 1e-6 * dT, Yt + k * 1e-6 * dY) for k in (-2, -1, 1, 2)}
        fds = {"1e-06": (Jp[1] - Jp[-1]) / 2e-6,
               "1e-07": (J_fwd(Tt + 1e-7 * dT, Yt + 1e-7 * dY) -
                         J_fwd(Tt - 1e-7 * dT, Yt - 1e-7 * dY)) / 2e-7,
               "1e-06_4pt": (8.0 * (Jp[1] - Jp[-1]) - (Jp[2] - Jp[-2])) / 12e-6}
        # branch-frozen runs (the unperturbed trajectory's clip branches),
        # traced: which other branches the +-eps, +-2 eps runs change
        Jf, recs = {}, {}
        for k in (-2, -1, 1, 2):
            Jf[k], recs[k] = J_fwd(Tt + k * 1e-6 * dT, Yt + k * 1e-6 * dY, trace=True,
                                   clip=("replay", masks))
        fds["frozen_1e-06"] = (Jf[1] - Jf[-1]) / 2e-6
        fds["frozen_1e-06_4pt"] = (8.0 * (Jf[1] - Jf[-1]) - (Jf[2] - Jf[-2])) / 12e-6
        for key, fd in fds.items():
            errs[key].append(abs(fd - ad) / max(abs(fd), abs(ad), 1e-300))
        sw = {}
        for name, cols in (("lf", LF), ("clip", CLIP), ("range", RANGE)):
            sw[name + "_2pt"] = bool(any((recs[k][:, cols] != rec0[:, cols]).any()
                                         for k in (-1, 1)))
            sw[name + "_4pt"] = bool(any((recs[k][:, cols] != rec0[:, cols]).any()
                                         for k in (-2, -1, 1, 2)))
        switches.append(sw)
        # A direction with no switch within +-eps whose two-point error
        # exceeds 1e-6: the quotient again at eps/2 (branch-frozen).  A
        # truncation error (J's third derivative along the direction) falls
        # 4x, and the Richardson combination (4 fd(eps/2) - fd(eps)) / 3
        # removes it; a kink between the traced points would not behave so.
        if not (sw["lf_2pt"] or sw["range_2pt"]) and errs["frozen_1e-06"][-1] > 1e-6:
            Jh = {k: J_fwd(Tt + k * 5e-7 * dT, Yt + k * 5e-7 * dY, clip=("replay", masks))
                  for k in (-1, 1)}
            fd_h = (Jh[1] - Jh[-1]) / 1e-6
            fd_r = (4.0 * fd_h - fds["frozen_1e-06"]) / 3.0
            rel = lambda fd: abs(fd - ad) / max(abs(fd), abs(ad), 1e-300)  # noqa: E731
            trunc.append({"direction": d, "rel_err_eps": errs["frozen_1e-06"][-1],
                          "rel_err_eps_half": rel(fd_h),
                          "ratio": errs["frozen_1e-06"][-1] / max(rel(fd_h), 1e-300),
                          "rel_err_richardson": rel(fd_r)})
    free2 = [i for i, sw in enumerate(switches) if not (sw["lf_2pt"] or sw["range_2pt"])]
    free4 = [i for i, sw in enumerate(switches) if not (sw["lf_4pt"] or sw["range_4pt"])]

    def _stats(key, idx):
        if not idx:
            return {"directions": 0}
        v = [errs[key][i] for i in idx]
        return {"directions": len(idx), "max_rel_err": max(v),
                "median_rel_err": float(np.median(v)), "within_1e-6": bool(max(v) <= 1e-6)}

    fd_free = {"2pt_frozen_eps1e-6": _stats("frozen_1e-06", free2),
               "4pt_frozen_eps1e-6": _stats("frozen_1e-06_4pt", free4),
               "note": "directions whose branch-frozen (clip) runs change no LF-maximum "
                       "cell and move no cell across the NASA-7 switch temperature"}
    fd_free["2pt_truncation_check"] = trunc
    # every direction without a switch within +-eps, each at its two-point
    # quotient, or its Richardson value where the former exceeds 1e-6
    corr = dict((i, errs["frozen_1e-06"][i]) for i in free2)
    for r in trunc:
        corr[r["direction"]] = r["rel_err_richardson"]
    if corr:
        fd_free["2pt_truncation_corrected"] = {
            "directions": len(corr), "max_rel_err": max(corr.values()),
            "median_rel_err": float(np.median(list(corr.values()))),
            "within_1e-6": bool(max(corr.values()) <= 1e-6)}
    per_dir = [dict(sw, rel_err_1e6=errs["1e-06"][i], rel_err_1e7=errs["1e-07"][i],
                    rel_err_4pt=errs["1e-06_4pt"][i], frozen_1e6=errs["frozen_1e-06"][i],
                    frozen_4pt=errs["frozen_1e-06_4pt"][i])
               for i, sw in enumerate(switches)]
    frozen = ({"max_rel_err_2pt": max(errs["frozen_1e-06"]),
               "max_rel_err_4pt": max(errs["frozen_1e-06_4pt"]),
               "median_rel_err_4pt": float(np.median(errs["frozen_1e-06_4pt"])),
               "bar": 1e-6, "within_bar": bool(max(errs["frozen_1e-06_4pt"]) <= 1e-6),
               "note": "finite differences of rollouts that replay the unperturbed "
                       "trajectory's chemistry clip branches"} if ndir else None)
    return {"workload": "1024x256 tube (dx 1.25e-5 m, 51-column 30 atm/3000 K driver, 0.5%% T "
                        "and 3 m/s velocity noise), FP64, fixed dt %.3e s, %d steps, checkpoints "
                        "every %d; J = heat release + pressure mismatch; grad wrt initial T, Y_k"
                        % (dt, steps, every),
            "cells": cells,
            "step_fwd_ms": t_fwd1 * 1e3, "step_vjp_ms": t_adj * 1e3,
            "step_vjp_from_residuals_ms": t_vjpr * 1e3,
            "step_keep_extra_ms": (t_keep - t_one) * 1e3,
            "step_vjp_recompute_stages_ms": t_vjpk * 1e3,
            "step_vjp_recompute_all_ms": t_vjp1 * 1e3,
            "adjoint_over_fwd_step": t_adj / t_fwd1,
            "adjoint_over_fwd_step_recompute_stages": t_vjpk / t_fwd1,
            "rollout_fwd_ms": t_fwd * 1e3,
            "rollout_checkpointed_fwd_ms": t_roll * 1e3,
            "rollout_fwd_plus_adjoint_ms": t_total * 1e3,
            "adjoint_over_fwd_rollout": (t_total - t_roll) / t_fwd,
            "fd_directions": ndir,
            "fd_max_rel_err": {e: max(v) for e, v in errs.items()} if ndir else None,
            "fd_switch_free": fd_free if ndir else None,
            "fd_branch_frozen": frozen,
            "fd_switches": dict({k: sum(sw[k] for sw in switches) for k in switches[0]},
                                clip_branches_active_per_step=float(rec0[:, 13].mean()),
                                cells_below_nasa7_switch_per_solve=float(
                                    rec0[:, 15:22:2].mean())) if ndir else None,
            "fd_per_direction": per_dir if ndir else None,
            "fd_median_rel_err": ({e: float(np.median(v)) for e, v in errs.items()}
                                  if ndir else None)}




def _amr_leg(mech, U3, aux3, g3, dt, steps, warmup, torch):
    """The AMR branch (advance_func_amr, set_simulator.py:496-532) on one
    level: the 8000 x 500 tube, FP64, tiled by same-size blocks, each step
    three block stages with the uniform block grid's ghost refill between them
    and the blocks' chemistry.  Timed: whole steps, and the library's block
    calls alone (torch events on the stream they run on)."""
    from janc_amd import cases
    from janc_amd.amr import BlockSolver, UniformBlockGrid
    nx, ny = U3.shape[1], U3.shape[2]
    bs = BlockSolver(cases.simulation_config(g3, dt=dt, precision="fp64"))
    dx, dy = g3["dx"], g3["dy"]
    out = {"workload": "8000x500 tube (FP64, fixed dt %.4g s) tiled by same-size blocks; "
                       "per step 3 block stages + ghost refills + block chemistry" % dt,
           "steps": steps}
    Ut = torch.as_tensor(U3).cuda()
    at = torch.as_tensor(aux3).cuda()
    for bnx, bny in ((40, 50), (20, 20)):
        grid = UniformBlockGrid(g3["boundary_config"], nx // bnx, ny // bny, bnx, bny)
        blk0 = grid.split(Ut, at)
        lib_ms = [0.0]

        def timed(fn, *a):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            r = fn(*a)
            e1.record()
            ev.append((e0, e1))
            return r

        def step(b):
            g0 = grid.get_ghost_block_data(b)
            gs = g0
            for s in range(3):
                o = timed(bs.stage, s, gs, g0, dx, dy, dt)
                b = grid.update_external_boundary(1, None, o[..., 3:-3, 3:-3])
                if s < 2:
                    gs = grid.get_ghost_block_data(b)
            return timed(bs.chemistry, b.contiguous(), dx, dy, dt)

        b = blk0.clone()
        ev = []
        for _ in range(warmup):
            b = step(b)
        torch.cuda.synchronize()
        ev = []
        t0 = time.perf_counter()
        for _ in range(steps):
            b = step(b)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        lib = sum(e0.elapsed_time(e1) for e0, e1 in ev) / steps
        cells = nx * ny
        out["blocks_%dx%d" % (bnx, bny)] = {
            "n_blk": grid.n_blk, "value": cells * steps / el, "ms_per_step": el / steps * 1e3,
            "library_ms_per_step": lib, "library_value": cells / (lib * 1e-3),
            "health_flags": bs.health()}
        del blk0, b, grid
        torch.cuda.empty_cache()
    return out


def _summary_adjoint(adj):
    """The adjoint leg's headline figures (its own line holds the rest)."""
    keys = ("adjoint_over_fwd_step", "adjoint_over_fwd_step_recompute_stages",
            "adjoint_over_fwd_rollout", "step_fwd_ms", "step_vjp_ms",
            "fd_directions")
    out = {k: adj[k] for k in keys if k in adj}
    free = adj.get("fd_switch_free") or {}
    four = free.get("4pt_frozen_eps1e-6")
    if four:
        out["fd_switch_free_4pt"] = {k: four[k] for k in ("directions", "max_rel_err",
                                                           "within_1e-6")}
    errs = adj.get("fd_max_rel_err") or {}
    if "1e-06" in errs:
        out["fd_max_rel_err_2pt_all"] = errs["1e-06"]
    sw = adj.get("fd_switches") or {}
    if "lf_4pt" in sw:
        out["fd_directions_with_lf_switch"] = sw["lf_4pt"]
    corr = free.get("2pt_truncation_corrected")
    if corr:
        out["fd_switch_free_2pt_truncation_corrected"] = {
            k: corr[k] for k in ("directions", "max_rel_err", "within_1e-6")}
    return out


def build_summary(steps, warmup, nx, ny, mixed, fp64, roofline, roof_valu, chem_fv, acc,
                  crop, adj, bat, amr_out, ann, cpu):
    """The compact last line: the headline and its roofline and CPU baseline
    beside the other legs' headline figures (each leg's own line holds the
    rest).  Legs that did not run are None and leave their keys out."""
    summary = {}
    summary.update({
        "metric": METRIC,
        "value": mixed["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "mixed-precision step (FP32 state, split fluxes, WENO5; FP64 "
                               "divergence, RK combine, chemistry, T solve), 8000x500 tube, "
                               "dx 1.25e-5 m, phi in [0.8,1.2] smooth field, 40-column 20 atm/"
                               "2800 K ignition slab, slip/zero-gradient/slip/slip, CFL 0.3, "
                               "9sp-19r H2-air, LF + WENO5_JS",
                   "nx": nx, "ny": ny, "cells": nx * ny,
                   "precision": "mixed"},
        "fp64_value": fp64["value"],
        "fp64_ms_per_step": fp64["ms_per_step"],
        "speedup_mixed_over_fp64": mixed["value"] / fp64["value"],
        "kernels_mixed_ms": {k: v["ms_total"] / v["launches"]
                             for k, v in mixed["kernels"].items() if v["launches"]},
        "kernels_fp64_ms": {k: v["ms_total"] / v["launches"]
                            for k, v in fp64["kernels"].items() if v["launches"]},
        "health_flags": {"mixed": mixed["health_flags"], "fp64": fp64["health_flags"]},
        "roofline": roofline,
        "roofline_valu": roof_valu,
        "roofline_chem": chem_fv,
    })
    if acc is not None:
        summary["mixed_err_vs_fp64_gpu"] = dict(_worst(acc["every"][-1]),
                                                steps=acc["every"][-1]["step"])
    if crop is not None:
        summary["mixed_err_vs_oracle_crop"] = dict(_worst(crop["every"][-1]),
                                                   steps=crop["every"][-1]["step"],
                                                   crop="%dx%d" % ORACLE_CROP)
    if adj is not None:
        summary["adjoint"] = _summary_adjoint(adj)
        summary["adjoint_over_fwd_step"] = adj["adjoint_over_fwd_step"]
    if bat is not None:
        summary["batch"] = {"value": bat["value"], "speedup_vs_sequential":
                            bat["speedup_vs_sequential"],
                            "bit_identical_to_sequential": bat["bit_identical_to_sequential"]}
    if amr_out is not None:
        summary["amr_blocks_fp64"] = {
            k: {"value": v["value"], "library_value": v["library_value"], "n_blk": v["n_blk"]}
            for k, v in amr_out.items() if k.startswith("blocks_")}
    if ann is not None:
        summary["fp64_annulus_4000x400"] = {"value": ann["value"],
                                            "ms_per_step": ann["ms_per_step"]}
    if cpu is not None:
        summary["cpu_baseline"] = cpu
        summary["speedup_vs_cpu"] = mixed["value"] / cpu["value"]
    return summary


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--accuracy-steps", type=int, default=1000)
    ap.add_argument("--crop-steps", type=int, default=100,
                    help="steps of the mixed-vs-oracle comparison on the tube crop")
    ap.add_argument("--cpu-steps", type=int, default=16)
    ap.add_argument("--cpu-procs", type=int, default=0,
                    help="CPU-baseline processes (default: min(15, usable cores))")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the tile autotuner")
    ap.add_argument("--tile", type=int, default=None,
                    help="force the stage-kernel variant of the tube steps (no autotuning)")
    ap.add_argument("--annulus-steps", type=int, default=100)
    ap.add_argument("--no-annulus", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=50)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=100)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--amr-steps", type=int, default=10)
    ap.add_argument("--no-amr", action="store_true")
    ap.add_argument("--fd-directions", type=int, default=32)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="one 256 MiB device-to-device copy before the timed region "
                         "(reference dispatch for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the mixed and FP64 steps on the tube (for profiler runs)")
    ap.add_argument("--annulus-only", action="store_true",
                    help="only the FP64 annulus step (for profiler runs)")
    ap.add_argument("--scheme-steps", type=int, default=50)
    ap.add_argument("--snapshot-steps", type=int, default=600)
    ap.add_argument("--no-snapshots", action="store_true")
    ap.add_argument("--no-schemes", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the single-block step runs on one GPU; --gpus N > 1 is not "
              "supported (the x-slab decomposition is not built)", file=sys.stderr)
        sys.exit(2)
    only = args.headline_only or args.annulus_only
    if only:
        args.no_cpu = args.no_batch = args.no_adjoint = args.no_amr = True
        args.no_schemes = args.no_snapshots = True
        args.accuracy_steps = args.crop_steps = 0
        args.no_annulus = args.headline_only

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from janc_amd import cases
    from janc_amd.mechanism import Mechanism

    mech = Mechanism()
    U3, aux3, g3 = cases.mixed_tube(mech)
    nx, ny = U3.shape[1], U3.shape[2]
    cells = nx * ny

    # ---- CPU work: worker processes forked before the GPU is touched ----
    pool = None
    cpu = None
    nproc = 0
    if not args.no_cpu:
        import multiprocessing as mp
        _CPU.update(U=U3, aux=aux3, dx=g3["dx"])
        nproc = args.cpu_procs or min(15, len(os.sched_getaffinity(0)))
        pool = mp.get_context("fork").Pool(nproc)
    try:
        if pool is not None:
            cpu = _cpu_baseline(pool, nproc, args.cpu_steps)
            emit("cpu_baseline", cpu)
        _gpu_legs(args, mech, U3, aux3, g3, cells, pool, cpu, only)
    finally:
        if pool is not None:
            pool.close()
            pool.join()


def _gpu_legs(args, mech, U3, aux3, g3, cells, pool, cpu, only):
    import torch
    from janc_amd.solver import Simulator
    tune = not args.no_tune
    if args.calibrate_copy:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # ---- the oracle's crop run for the mixed form's accuracy (cpu_baseline
    # leg: the oracle is the checker) starts now, beside the GPU legs ----
    crop_job = None
    dt_fixed = None
    if pool is not None and args.crop_steps > 0:
        from janc_amd import cases
        probe = Simulator(cases.simulation_config(g3, cfl=0.3))
        dt_fixed = float(probe.CFL(U3, aux3, 0.3))
        del probe
        every = max(1, args.crop_steps // 5)
        crop_job = pool.apply_async(_oracle_crop_worker,
                                    ((ORACLE_CROP[0], ORACLE_CROP[1], dt_fixed,
                                      args.crop_steps, every),))

    if args.annulus_only:
        ann, _ = _annulus_leg(mech, args.steps, args.warmup, tune, torch)
        emit("fp64_annulus", ann)
        print(json.dumps({"metric": METRIC, "value": ann["value"], "unit": "cell-updates/s",
                          "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": ann["ms_per_step"], "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                          "data": "synthetic",
                          "config": {"workload": ann["workload"], "cells": ann["cells"]}}),
              flush=True)
        return

    # ---- headline: mixed and FP64 on the 8000 x 500 tube ----
    head = _headline(mech, U3, aux3, g3, args.steps, args.warmup, tune, torch, args.tile)
    mixed, fp64 = head["mixed"], head["fp64"]
    emit("headline_tube", {k: {kk: vv for kk, vv in v.items() if kk != "kt"}
                           for k, v in head.items()})
    roof = mixed["roofline"]
    kt = mixed["kt"]
    dom = max(roof, key=lambda k: kt[k][0])
    r = roof[dom]
    tr = STAGE_TRAFFIC_B_PER_CELL["mixed"] if dom == "stage" else None
    roofline = {"kernel": dom, "bound": "hbm", "achieved": r["achieved_GBs"],
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": r["frac_hbm"],
                "traffic": tr * cells if tr else None,
                "algorithmic_bytes": r["algorithmic_bytes"], "avg_ms": r["avg_ms"]}
    roof_valu = None
    if "stage" in roof:
        floor_ms = (STAGE_VALU_PER_CELL["mixed"] * cells / 64.0 * 4.0
                    / (SIMDS * CLOCK_GHZ * 1e9) * 1e3)
        roof_valu = {"kernel": "stage", "bound": "valu-issue",
                     "instr_per_cell": STAGE_VALU_PER_CELL["mixed"],
                     "floor_ms": floor_ms, "avg_ms": roof["stage"]["avg_ms"],
                     "frac": floor_ms / roof["stage"]["avg_ms"]}
    chem_fv = None
    if "chem" in kt and kt["chem"][1]:
        ms, nl = kt["chem"]
        tfs = CHEM_FP64_FLOP_PER_CELL["mixed"] * cells / (ms / nl * 1e-3) / 1e12
        chem_fv = {"bound": "fp64-valu", "achieved": tfs, "peak": FP64_VECTOR_PEAK_TFLOPS,
                   "unit": "TFLOP/s", "frac": tfs / FP64_VECTOR_PEAK_TFLOPS,
                   "avg_ms": ms / nl, "flop_per_cell": CHEM_FP64_FLOP_PER_CELL["mixed"]}

    # ---- mixed accuracy: 1000 steps against the FP64 step ----
    acc = None
    if args.accuracy_steps > 0:
        acc = _accuracy(mech, U3, aux3, g3, args.accuracy_steps, 100, torch)
        emit("mixed_accuracy", acc)
        torch.cuda.empty_cache()
    crop_gpu = None
    if crop_job is not None:
        every = max(1, args.crop_steps // 5)
        crop_gpu, crop_flags = _accuracy_crop_gpu(mech, U3, aux3, g3, dt_fixed,
                                                  args.crop_steps, every, torch)

    ann = None
    if not args.no_annulus:
        ann, ann_state = _annulus_leg(mech, args.annulus_steps, args.warmup, tune, torch)
        emit("fp64_annulus", ann)
        if not args.no_schemes:
            emit("schemes_fp64_annulus", _schemes_leg(*ann_state, args.scheme_steps,
                                                      args.warmup, tune, torch))
        if not args.no_snapshots:
            emit("snapshots", _snapshot_leg(*ann_state, args.snapshot_steps, 200, torch))
            torch.cuda.empty_cache()
    bat = None
    if not args.no_batch:
        bat = _batch_leg(mech, args.batch_steps, tune, torch)
        emit("batch", bat)
        torch.cuda.empty_cache()
    amr_out = None
    if not args.no_amr:
        dt_amr = dt_fixed if dt_fixed is not None else head["fp64"]["dt_first"]
        amr_out = _amr_leg(mech, U3, aux3, g3, dt_amr, args.amr_steps, 2, torch)
        emit("amr_blocks", amr_out)
        torch.cuda.empty_cache()
    adj = None
    if not args.no_adjoint:
        adj = _adjoint_leg(mech, args.adjoint_steps, 10, args.fd_directions, torch)
        emit("adjoint", adj)
        torch.cuda.empty_cache()
    crop = None
    if crop_job is not None:
        states, el = crop_job.get()
        crop = _accuracy_crop_compare(mech, crop_gpu, states, args.crop_steps, dt_fixed, el)
        crop["health_flags"] = crop_flags
        emit("mixed_vs_oracle_crop", crop)

    # ---- summary (last line) ----
    print(json.dumps(build_summary(args.steps, args.warmup, U3.shape[1], U3.shape[2], mixed,
                                   fp64, roofline, roof_valu, chem_fv, acc, crop, adj, bat,
                                   amr_out, ann, cpu)), flush=True)


if __name__ == "__main__":
    main()
