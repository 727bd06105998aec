"""The small CPU configuration (400 x 80 tube, 20-column 30 atm / 3000 K
unburned driver, CFL 0.3) at its literal cell size dx = 1e-4 m.

The reference algorithm itself (the FP64 oracle: first-order point-implicit
chemistry, reaction_model.py:183-193, then the per-species clip) drives the
temperature negative inside the driver after 29 steps at this cell size,
which is why the tube configurations run at the reference's own
detonation-tube cell size, 1.25e-5 m (examples/janc_basic_example1.ipynb:
0.05 m / 4000 cells).  The GPU side of the same statement (parity on the
literal state before it fails, the GPU going negative too, and the
substituted cell size running clean for the configuration's 200 steps) is in
tests/test_gpu_configs.py.
"""
import numpy as np

from janc_amd import cases
from janc_amd.mechanism import Mechanism
from oracle.step import Oracle

# first step after which the oracle's minimum temperature is negative
C1_LITERAL_FAILS_AT = 29


def test_c1_literal_cell_size_drives_T_negative(mech_path):
    mech = Mechanism(mech_path)
    U, aux, g = cases.detonation_tube(mech, dx=1e-4)
    assert U.shape == (12, 400, 80) and g["dx"] == g["dy"] == 1e-4
    o = Oracle(mech_path, g["boundary_config"], g["dx"], g["dy"])
    t = 0.0
    Tmin = []
    for _ in range(C1_LITERAL_FAILS_AT):
        with np.errstate(all="ignore"):
            U, aux, t, _ = o.advance_func(U, aux, t, 0.3)
        Tmin.append(float(np.nanmin(aux[1])))
    # physical (fresh gas stays near its 300 K) until the last step, then
    # a cell of the driver goes far below zero
    assert min(Tmin[:-1]) > 290.0
    assert Tmin[-1] < -100.0
    i, j = np.unravel_index(np.nanargmin(aux[1]), aux[1].shape)
    assert i < 20                  # inside the 20-column driver
