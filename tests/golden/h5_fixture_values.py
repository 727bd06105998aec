"""Closed-form contents of tests/golden/h5py_snapshots.h5 (see
make_h5_fixture.py): small arrays of the reference's U shape (12, nx, ny)
in float64 and float32, with a few chunk-edge sizes."""
import numpy as np

NSNAP = 11          # more than one symbol-table node (8 entries) in the root


def fixture_arrays(k):
    nx, ny = 5 + k % 3, 7 + 2 * (k % 2)
    i, j, v = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(12), indexing="ij")
    u = (np.sin(0.3 * i + 0.7 * j + 0.11 * v + k) * 10.0 ** (v % 5)).transpose(2, 0, 1)
    out = {"u": np.ascontiguousarray(u)}
    if k % 4 == 0:
        out["u32"] = out["u"].astype(np.float32)
    if k == 2:
        out["plain"] = np.arange(60, dtype=np.float64).reshape(3, 4, 5) * 0.5
        out["shuffled"] = np.arange(24, dtype=np.float64).reshape(2, 3, 4) ** 2
    return out


def fixture_attrs(k):
    return 1.25e-6 * k, 10 * k
