"""Write tests/golden/h5py_snapshots.h5 with h5py (libhdf5), the way the
reference's H5Saver does (simulation/set_simulator.py:155-175): one group
save_step{k} per snapshot, attributes time (float) and step (int), a gzip
dataset u.  Run with an interpreter that has h5py (e.g. /opt/conda/bin/python3.9):

    python tests/golden/make_h5_fixture.py

The values are the closed forms of fixture_arrays() (shared with the test),
so the fixture holds inputs only; the test recomputes the expected arrays.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from h5_fixture_values import NSNAP, fixture_arrays, fixture_attrs  # noqa: E402


def main():
    import h5py
    path = os.path.join(HERE, "h5py_snapshots.h5")
    if os.path.exists(path):
        os.remove(path)
    f = h5py.File(path, "a")
    for k in range(NSNAP):
        key = "save_step%d" % k
        if key in f:
            del f[key]
        grp = f.create_group(key)
        t, step = fixture_attrs(k)
        grp.attrs["time"] = t
        grp.attrs["step"] = step
        for name, arr in fixture_arrays(k).items():
            if name == "plain":                     # contiguous, no filter
                grp.create_dataset(name, data=arr)
            elif name == "shuffled":                # shuffle + gzip
                grp.create_dataset(name, data=arr, compression="gzip", shuffle=True)
            else:
                grp.create_dataset(name, data=arr, compression="gzip")
        f.flush()
    # re-save one snapshot as H5Saver.save does for an existing key
    del f["save_step1"]
    grp = f.create_group("save_step1")
    t, step = fixture_attrs(1)
    grp.attrs["time"] = t
    grp.attrs["step"] = step
    for name, arr in fixture_arrays(1).items():
        grp.create_dataset(name, data=arr, compression="gzip")
    f.close()
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
