"""The ctypes DLPack producer/consumer (tests/dlpack_ctypes.py) against
torch on host memory: struct layout, zero-copy import, the deleter's
release, and export.  The GPU form (kDLROCM memory into the solver) is in
test_gpu_dlpack.py."""
import gc

import numpy as np
import pytest

from dlpack_ctypes import Producer, consume, kDLCPU

torch = pytest.importorskip("torch")


def test_host_producer_into_torch():
    v = np.random.default_rng(3).standard_normal((12, 5, 7))
    p = Producer(v)
    t = torch.from_dlpack(p)
    assert t.dtype == torch.float64 and tuple(t.shape) == v.shape
    assert t.data_ptr() == p.ptr                      # zero copy
    assert np.array_equal(t.numpy(), v)
    t[0, 0, 0] = 42.0                                 # shared memory
    assert p.read()[0, 0, 0] == 42.0
    assert not p.released()
    del t
    gc.collect()
    assert p.released()                               # the consumer called the deleter


def test_torch_capsule_into_consumer():
    t = torch.arange(24, dtype=torch.float64).reshape(2, 3, 4) * 0.5
    out, (dtype, did, ptr) = consume(torch.utils.dlpack.to_dlpack(t))
    assert dtype == kDLCPU and ptr == t.data_ptr()
    assert np.array_equal(out, t.numpy())
