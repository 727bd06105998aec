"""GPU runs of the benchmark configurations at sizes the oracle finishes in
seconds.

* The small CPU configuration (400 x 80 tube) at its literal cell size
  dx = 1e-4 m: FP64 parity with the oracle while the state is physical, the
  GPU step driving T negative like the oracle does (tests/test_oracle_c1.py),
  and the substituted cell size (1.25e-5 m) running the configuration's 200
  steps clean.
* The mixed-precision configuration (8000 x 500 tube, smooth phi field in
  [0.8, 1.2], 40-column 20 atm / 2800 K slab) on its 400 x 50 corner crop
  (slab and phi variation included) at the full tube's CFL-0.3 dt: the mixed
  step against the FP64 oracle every 25 steps for 100 steps, within the
  bounds stated below, and the FP64 step against the oracle over the same
  100 steps: (gamma, T) at the 1e-12 bar, U within 1e-10.  The oracle's own
  state moves by 4.4e-11 (max-norm relative, rho*v: small next to its
  rounding) after 5 steps when its initial state is perturbed by one ulp, so
  over a hundred steps the per-operation rounding differences of two correct
  FP64 implementations reach that size; measured GPU-vs-oracle 5.9e-12 after
  25 steps.
"""
import numpy as np
import pytest

from janc_amd import _lib, cases
from janc_amd.mechanism import Mechanism
from oracle.step import Oracle

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a GPU")]

FP64_TOL = 1e-12
FP64_LONG_U = 1e-10     # U after tens of steps (see the module docstring)


def rel_err(got, ref):
    got = got.detach().cpu().numpy() if hasattr(got, "detach") else np.asarray(got)
    ref = np.asarray(ref)
    d = np.abs(got - ref).reshape(ref.shape[0], -1).max(axis=1)
    s = np.abs(ref).reshape(ref.shape[0], -1).max(axis=1)
    return float(np.max(d / np.maximum(s, 1e-300)))


@pytest.fixture(scope="module")
def mech(mech_path):
    return Mechanism(mech_path)


def _sim(mech_path, g, **kw):
    from janc_amd.solver import Simulator
    return Simulator(cases.simulation_config(g, mech_path, **kw))


# ---------------------------------------------------------------- C1

def test_c1_literal_parity_then_negative_T(mech_path, mech):
    U, aux, g = cases.detonation_tube(mech, dx=1e-4)
    print("C1 literal dx=1e-4", U.shape, flush=True)
    sim = _sim(mech_path, g, cfl=0.3)
    o = Oracle(mech_path, g["boundary_config"], g["dx"], g["dy"])
    Ug, ag = torch.as_tensor(U).cuda(), torch.as_tensor(aux).cuda()
    Ur, ar, t = U, aux, 0.0
    for s in range(10):
        Ug, ag, _, dtg = sim.advance_func(Ug, ag, 0.0)
        Ur, ar, t, dtr = o.advance_func(Ur, ar, t, 0.3)
        assert abs(float(dtg) - dtr) <= 1e-14 * dtr
    assert rel_err(Ug, Ur) <= FP64_TOL
    assert rel_err(ag, ar) <= FP64_TOL
    # the GPU step goes on to a negative temperature within the steps the
    # oracle takes (29) plus a margin; before that every cell is physical
    Tmin = []
    for s in range(10, 40):
        Ug, ag, _, _ = sim.advance_func(Ug, ag, 0.0)
        Tmin.append(float(torch.nan_to_num(ag[1], nan=1e9).min()))
        if Tmin[-1] < 0:
            break
    assert Tmin[-1] < 0, Tmin
    assert min(Tmin[:-1]) > 290.0
    # the device flag names it
    assert sim.health(Ug) & _lib.HEALTH_FLAGS["nonpositive_T"]
    assert 25 <= 10 + len(Tmin) <= 35


def test_c1_substituted_cell_size_runs_clean(mech_path, mech):
    U, aux, g = cases.detonation_tube(mech)
    assert g["dx"] == cases.TUBE_DX
    sim = _sim(mech_path, g, cfl=0.3)
    Ug, ag = torch.as_tensor(U).cuda(), torch.as_tensor(aux).cuda()
    sim.advance(Ug, ag, 200)
    assert sim.health(Ug) == 0
    T = ag[1]
    assert bool(torch.isfinite(Ug).all()) and bool(torch.isfinite(ag).all())
    assert float(T.min()) > 290.0 and float(T.max()) < 4000.0


# ---------------------------------------------------------------- C3 crop

C3_CROP = (400, 50)
C3_STEPS = 100
C3_EVERY = 25
# Bounds of the mixed step against the FP64 oracle after 100 steps on the
# crop (relative L2 of the field; Linf relative to the field's max |.|).
MIXED_L2 = 1e-5
MIXED_LINF = 1e-4


def _c3_dt(mech):
    # the full tube's CFL-0.3 dt: at rest, the largest |u| + a is in the
    # 2800 K slab (columns 0-39 over the whole height)
    U, aux, g = cases.mixed_tube(mech, crop=(40, 500))
    R = (U[4:] / U[0])
    Y9 = np.concatenate([R, (1.0 - R.sum(0))[None]])
    Rm = np.sum((1.0 / mech.species_M)[:, None, None] * Y9, axis=0)
    p = U[0] * Rm * aux[1]
    a = np.sqrt(aux[0] * p / U[0])
    u, v = U[1] / U[0], U[2] / U[0]
    return 0.3 * np.min(1.0 / ((np.abs(u) + a) / g["dx"] + (np.abs(v) + a) / g["dy"]))


def _fields(U, aux, mech):
    U = np.asarray(U, dtype=np.float64)
    aux = np.asarray(aux, dtype=np.float64)
    Y = U[4:] / U[0]
    Y9 = np.concatenate([Y, (1.0 - Y.sum(0))[None]])
    R = np.sum((1.0 / mech.species_M)[:, None, None] * Y9, axis=0)
    return {"rho": U[0], "p": U[0] * R * aux[1], "T": aux[1], "Y": Y}


def test_c3_crop_mixed_and_fp64_vs_oracle(mech_path, mech):
    U, aux, g = cases.mixed_tube(mech, crop=C3_CROP)
    dt = _c3_dt(mech)
    print("C3 crop", U.shape, "dt %.4e" % dt, flush=True)
    o = Oracle(mech_path, g["boundary_config"], g["dx"], g["dy"])
    sm = _sim(mech_path, g, dt=dt, precision="mixed")
    sd = _sim(mech_path, g, dt=dt, precision="fp64")
    Um = torch.as_tensor(U, dtype=torch.float32).cuda()
    am = torch.as_tensor(aux, dtype=torch.float32).cuda()
    Ud = torch.as_tensor(U).cuda()
    ad = torch.as_tensor(aux).cuda()
    Ur, ar = U, aux
    worst = {}
    for s in range(C3_EVERY, C3_STEPS + 1, C3_EVERY):
        sm.advance(Um, am, C3_EVERY)
        sd.advance(Ud, ad, C3_EVERY)
        for _ in range(C3_EVERY):
            Ur, ar = o.advance_one_step(Ur, ar, dt)
        fd = rel_err(Ud, Ur), rel_err(ad, ar)
        fm = _fields(Um.cpu().numpy(), am.cpu().numpy(), mech)
        fr = _fields(Ur, ar, mech)
        row = {}
        for k in ("rho", "p", "T"):
            d = fm[k] - fr[k]
            row[k] = (np.linalg.norm(d) / np.linalg.norm(fr[k]),
                      np.abs(d).max() / np.abs(fr[k]).max())
        for i, name in enumerate(mech.species[:-1]):
            d = fm["Y"][i] - fr["Y"][i]
            row["Y_" + name] = (np.linalg.norm(d) / np.linalg.norm(fr["Y"][i]),
                                np.abs(d).max() / np.abs(fr["Y"][i]).max())
        print("step %d fp64 %.2e %.2e | mixed " % ((s,) + fd) +
              " ".join("%s %.1e/%.1e" % (k, *v) for k, v in row.items()), flush=True)
        worst = row
        assert fd[0] <= FP64_LONG_U and fd[1] <= FP64_TOL
    assert sm.health(Um) == 0 and sd.health(Ud) == 0
    for k, (l2, linf) in worst.items():
        assert l2 <= MIXED_L2, (k, l2)
        assert linf <= MIXED_LINF, (k, linf)
