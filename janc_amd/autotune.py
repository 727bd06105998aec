"""Tile autotuner for the stage kernel.

The library compiles several stage-kernel variants (tile height along x and
register budget / occupancy, ``jh_tile_info``).  ``tune`` times each variant
for a few steps on a copy of the caller's state and selects the fastest for
that context; results are cached per (GPU, precision, nx, ny) in memory and,
when ``JANC_AMD_TUNE_CACHE`` names a file, in that JSON file.
"""
import ctypes
import json
import os
import time

from . import _lib

_MEM = {}


def variants():
    lib = _lib.load()
    out = []
    for v in range(lib.jh_tile_variants()):
        r, t, w = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(lib.jh_tile_info(v, ctypes.byref(r), ctypes.byref(t), ctypes.byref(w)))
        out.append({"variant": v, "rows": r.value, "threads": t.value, "waves_per_simd": w.value})
    return out


def _key(sim, U):
    """Cache key: GPU, precision, grid, flux scheme and boundary kinds (the
    boundary kernel's cost and the stage kernel's halo sources differ with
    them)."""
    import torch
    dev = torch.cuda.get_device_name(U.device)
    return "%s|%s|%dx%d|flux=%s|bc=%s" % (
        dev, "fp64" if sim.precision == _lib.PREC_FP64 else "mixed", U.shape[1], U.shape[2],
        ",".join(str(v) for v in sim.flux), ",".join(str(k) for k in sim.bc_kinds))


def tunable(sim):
    """Whether the context's stage kernel has variants: the tile variant
    selects the LF flux-splitting kernel (WENO5_JS or MUSCL) with the
    per-cell temperature-solve stop; the Godunov and Steger-Warming kernels
    and the grid-wide-stop kernel have one tile shape each."""
    solver, split = sim.flux[0], sim.flux[1]
    return (solver == _lib.SOLVER_TYPES["flux_splitting"] and split == _lib.SPLIT_METHODS["LF"]
            and sim.halley_stop == "per_cell")


def _load_file():
    path = os.environ.get("JANC_AMD_TUNE_CACHE")
    if path and os.path.exists(path):
        with open(path) as fh:
            return json.load(fh)
    return {}


def _save_file(table):
    path = os.environ.get("JANC_AMD_TUNE_CACHE")
    if path:
        with open(path, "w") as fh:
            json.dump(table, fh, indent=1, sort_keys=True)


def tune(sim, U, aux, steps=8, theta=None, force=False):
    """Pick the fastest stage-kernel variant for sim's context of U's shape.

    Returns {"variant": best, "ms_per_step": {variant: ms}, "cached": bool}.
    The caller's U and aux are not modified."""
    import torch
    U, aux = sim._state(U, aux)
    if not tunable(sim):
        return {"variant": sim.get_tile(U, theta), "tuned": False,
                "reason": "this flux path / temperature-solve stop has one tile shape"}
    key = _key(sim, U)
    table = dict(_load_file())
    table.update(_MEM)
    if not force and key in table:
        best = table[key]["variant"]
        sim.set_tile(U, best, theta)
        return dict(table[key], cached=True)
    # every variant is timed in each of `rounds` interleaved passes and keeps
    # its best: one short sample per variant cannot separate variants within
    # a few per cent of each other
    times = {}
    rounds = 2
    for r in range(rounds):
        for v in variants():
            sim.set_tile(U, v["variant"], theta)
            Ut, at = U.clone(), aux.clone()
            sim.advance(Ut, at, 1, theta)               # warm-up / first-launch costs
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sim.advance(Ut, at, steps, theta)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) / steps * 1e3
            times[v["variant"]] = min(ms, times.get(v["variant"], ms))
            del Ut, at
    best = min(times, key=times.get)
    sim.set_tile(U, best, theta)
    entry = {"variant": best, "ms_per_step": {str(k): v for k, v in times.items()}}
    _MEM[key] = entry
    table[key] = entry
    _save_file(table)
    return dict(entry, cached=False)
