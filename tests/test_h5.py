"""HDF5 snapshots (janc_amd.h5) against libhdf5: the reader on a file h5py
wrote the way the reference's H5Saver does (tests/golden/h5py_snapshots.h5,
made by tests/golden/make_h5_fixture.py), the writer read back by itself and,
where an interpreter with h5py exists on this machine, by h5py."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from janc_amd.h5 import H5File, H5Saver

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from h5_fixture_values import NSNAP, fixture_arrays, fixture_attrs  # noqa: E402

FIXTURE = os.path.join(HERE, "golden", "h5py_snapshots.h5")


def _h5py_python():
    for p in ("/opt/conda/bin/python3.9", "/opt/conda/bin/python3", shutil.which("python3")):
        if p and os.path.exists(p):
            r = subprocess.run([p, "-c", "import h5py"], capture_output=True)
            if r.returncode == 0:
                return p
    return None


def test_reads_h5py_written_snapshots():
    f = H5File(FIXTURE, "r")
    assert sorted(f.keys()) == sorted("save_step%d" % k for k in range(NSNAP))
    for k in range(NSNAP):
        data, attrs = f.read_group("save_step%d" % k)
        t, step = fixture_attrs(k)
        assert attrs["time"] == t and attrs["step"] == step
        exp = fixture_arrays(k)
        assert sorted(data) == sorted(exp)
        for name, arr in exp.items():
            assert data[name].dtype == arr.dtype
            np.testing.assert_array_equal(data[name], arr)
    f.close()


def _snapshots(n, rng):
    out = []
    for k in range(n):
        nx, ny = 3 + k % 5, 4 + k % 7
        out.append(("save_step%d" % k, {"u": rng.standard_normal((12, nx, ny))},
                    {"time": np.float64(k * 1e-7), "step": np.int64(k)}))
    return out


def test_round_trip_many_groups_and_chunks(tmp_path):
    """300 groups (a two-level group B-tree) and a dataset of 240 chunks
    (a two-level chunk B-tree), ragged edge chunks, float32 and float64."""
    rng = np.random.default_rng(0)
    path = str(tmp_path / "a.h5")
    f = H5File(path, "w", chunk_bytes=512)
    snaps = _snapshots(300, rng)
    for key, data, attrs in snaps:
        f.write_group(key, data, attrs)
    big = rng.standard_normal((12, 41, 29))
    f.write_group("big", {"u": big, "u32": big.astype(np.float32)}, {"step": np.int64(7)})
    f.close()
    g = H5File(path, "r")
    assert len(g.keys()) == 301
    for key, data, attrs in snaps[::37]:
        d, a = g.read_group(key)
        np.testing.assert_array_equal(d["u"], data["u"])
        assert a == attrs
    d, a = g.read_group("big")
    np.testing.assert_array_equal(d["u"], big)
    np.testing.assert_array_equal(d["u32"], big.astype(np.float32))
    g.close()


def test_append_and_replace(tmp_path):
    rng = np.random.default_rng(1)
    path = str(tmp_path / "b.h5")
    f = H5File(path, "a")
    f.write_group("save_step0", {"u": np.ones((12, 3, 4))}, {"time": np.float64(0.0)})
    f.close()
    f = H5File(path, "a")                       # reopen, add and replace
    u1 = rng.standard_normal((12, 5, 6))
    f.write_group("save_step1", {"u": u1}, {"time": np.float64(1.0)})
    f.write_group("save_step0", {"u": 2 * np.ones((12, 3, 4))}, {"time": np.float64(0.5)})
    f.close()
    g = H5File(path, "r")
    assert sorted(g.keys()) == ["save_step0", "save_step1"]
    d, a = g.read_group("save_step0")
    assert a["time"] == 0.5 and np.all(d["u"] == 2.0)
    np.testing.assert_array_equal(g.read_group("save_step1")[0]["u"], u1)
    g.close()


def test_appends_to_h5py_file(tmp_path):
    path = str(tmp_path / "c.h5")
    shutil.copy(FIXTURE, path)
    f = H5File(path, "a")
    extra = np.arange(12 * 4 * 5, dtype=np.float64).reshape(12, 4, 5)
    f.write_group("save_step%d" % NSNAP, {"u": extra}, {"time": np.float64(9.0),
                                                        "step": np.int64(99)})
    f.close()
    g = H5File(path, "r")
    assert len(g.keys()) == NSNAP + 1
    np.testing.assert_array_equal(g.read_group("save_step3")[0]["u"], fixture_arrays(3)["u"])
    np.testing.assert_array_equal(g.read_group("save_step%d" % NSNAP)[0]["u"], extra)
    g.close()


def test_saver_api_on_host_arrays(tmp_path):
    path = str(tmp_path / "d.h5")
    s = H5Saver(path, "2D")
    U = np.random.default_rng(2).standard_normal((12, 6, 5))
    aux = np.stack([np.full((6, 5), 1.3), np.full((6, 5), 900.0)])
    s.save(0, 0.0, 0, u=U)
    s.save(1, 2e-7, 10, u=2 * U, aux=aux)
    assert sorted(s.list_snapshots()) == ["save_step0", "save_step1"]
    t, step, U1, a1 = s.load_state("save_step1")
    assert (t, step) == (2e-7, 10)
    np.testing.assert_array_equal(U1, 2 * U)
    np.testing.assert_array_equal(a1, aux)
    s.close()


@pytest.mark.skipif(_h5py_python() is None, reason="no interpreter with h5py on this machine")
def test_h5py_reads_our_files(tmp_path):
    """Whole and partial reads (libhdf5 looks partial reads up through the
    chunk B-tree keys) of 270 groups (two-level group B-tree); the larger
    datasets have more than 64 chunks (two-level chunk B-tree)."""
    rng = np.random.default_rng(3)
    path = str(tmp_path / "e.h5")
    f = H5File(path, "w", chunk_bytes=300)
    arrays = {}
    for k in range(270):
        u = rng.standard_normal((12, 9 + k % 23, 13))
        arrays[k] = u
        f.write_group("save_step%d" % k, {"u": u, "u32": u.astype(np.float32)},
                      {"time": np.float64(k * 0.5), "step": np.int64(3 * k)})
    f.close()
    out = str(tmp_path / "e.npz")
    script = (
        "import h5py, numpy as np, sys\n"
        "f = h5py.File(sys.argv[1], 'r')\n"
        "d = {}\n"
        "for k in f.keys():\n"
        "    g = f[k]\n"
        "    d[k + '/u'] = g['u'][:]\n"
        "    d[k + '/part'] = g['u'][5, 4:7, 2:11]\n"
        "    d[k + '/u32'] = g['u32'][:]\n"
        "    d[k + '/time'] = np.float64(g.attrs['time'])\n"
        "    d[k + '/step'] = np.int64(g.attrs['step'])\n"
        "    assert g['u'].compression == 'gzip'\n"
        "np.savez(sys.argv[2], **d)\n")
    subprocess.run([_h5py_python(), "-c", script, path, out], check=True)
    d = np.load(out)
    for k, u in arrays.items():
        key = "save_step%d" % k
        np.testing.assert_array_equal(d[key + "/u"], u)
        np.testing.assert_array_equal(d[key + "/part"], u[5, 4:7, 2:11])
        np.testing.assert_array_equal(d[key + "/u32"], u.astype(np.float32))
        assert float(d[key + "/time"]) == k * 0.5 and int(d[key + "/step"]) == 3 * k
