"""A DLPack producer and consumer written with ctypes only (no torch, no
numpy DLPack support): the non-torch side of the DLPack bridge tests.

``Producer`` exports a C-contiguous float64 buffer -- host memory (kDLCPU) or
device memory from hipMalloc (kDLROCM) -- as a legacy ``dltensor`` capsule
(DLPack 0.8 structs: DLDevice, DLDataType, DLTensor, DLManagedTensor).
``consume`` reads a ``dltensor`` capsule that another library produced (the
data pointer, shape, strides, dtype and device of its DLManagedTensor) and
copies the values into a numpy array, then hands the tensor back to its
owner through the deleter, as a consumer must."""
import ctypes

import numpy as np

kDLCPU, kDLCUDA, kDLROCM = 1, 2, 10
kDLFloat = 2


class DLDevice(ctypes.Structure):
    _fields_ = [("device_type", ctypes.c_int32), ("device_id", ctypes.c_int32)]


class DLDataType(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("bits", ctypes.c_uint8), ("lanes", ctypes.c_uint16)]


class DLTensor(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("device", DLDevice), ("ndim", ctypes.c_int32),
                ("dtype", DLDataType), ("shape", ctypes.POINTER(ctypes.c_int64)),
                ("strides", ctypes.POINTER(ctypes.c_int64)), ("byte_offset", ctypes.c_uint64)]


class DLManagedTensor(ctypes.Structure):
    pass


_Deleter = ctypes.CFUNCTYPE(None, ctypes.POINTER(DLManagedTensor))
DLManagedTensor._fields_ = [("dl_tensor", DLTensor), ("manager_ctx", ctypes.c_void_p),
                            ("deleter", _Deleter)]

_api = ctypes.pythonapi
_api.PyCapsule_New.restype = ctypes.py_object
_api.PyCapsule_New.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
_api.PyCapsule_GetPointer.restype = ctypes.c_void_p
_api.PyCapsule_GetPointer.argtypes = [ctypes.py_object, ctypes.c_char_p]
_api.PyCapsule_SetName.restype = ctypes.c_int
_api.PyCapsule_SetName.argtypes = [ctypes.py_object, ctypes.c_char_p]

_hip = None


def hip():
    """libamdhip64 through ctypes (hipMalloc, hipMemcpy, hipFree,
    hipDeviceSynchronize)."""
    global _hip
    if _hip is None:
        h = ctypes.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipFree.argtypes = [ctypes.c_void_p]
        _hip = h
    return _hip


H2D, D2H = 1, 2
_released = []        # manager_ctx values the consumers handed back


@_Deleter
def _on_release(mt):
    _released.append(mt.contents.manager_ctx)


class Producer:
    """A C-contiguous float64 array in host memory (device=None) or in
    hipMalloc memory (device = GPU ordinal), exported through __dlpack__."""

    _next_id = 1

    def __init__(self, values, device=None):
        values = np.ascontiguousarray(values, dtype=np.float64)
        self.shape = values.shape
        self.nbytes = values.nbytes
        self.device = device
        self.id = Producer._next_id
        Producer._next_id += 1
        if device is None:
            self._host = values.copy()
            self.ptr = self._host.ctypes.data
        else:
            p = ctypes.c_void_p()
            rc = hip().hipMalloc(ctypes.byref(p), self.nbytes)
            if rc != 0:
                raise RuntimeError("hipMalloc failed: %d" % rc)
            self.ptr = p.value
            rc = hip().hipMemcpy(self.ptr, values.ctypes.data, self.nbytes, H2D)
            if rc != 0:
                raise RuntimeError("hipMemcpy failed: %d" % rc)
        nd = len(self.shape)
        self._shape = (ctypes.c_int64 * nd)(*self.shape)
        self._mt = DLManagedTensor()
        t = self._mt.dl_tensor
        t.data = self.ptr
        t.device = DLDevice(kDLCPU if device is None else kDLROCM, 0 if device is None else device)
        t.ndim = nd
        t.dtype = DLDataType(kDLFloat, 64, 1)
        t.shape = self._shape
        t.strides = None            # compact row-major
        t.byte_offset = 0
        self._mt.manager_ctx = self.id
        self._mt.deleter = _on_release

    def __dlpack_device__(self):
        return (kDLCPU, 0) if self.device is None else (kDLROCM, self.device)

    def __dlpack__(self, stream=None):
        # the values were written by a synchronous hipMemcpy: nothing pending
        return _api.PyCapsule_New(ctypes.addressof(self._mt), b"dltensor", None)

    def released(self):
        return self.id in _released

    def read(self):
        out = np.empty(self.shape, dtype=np.float64)
        if self.device is None:
            out[...] = self._host
        else:
            rc = hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H)
            if rc != 0:
                raise RuntimeError("hipMemcpy failed: %d" % rc)
        return out

    def free(self):
        if self.device is not None and self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = 0


def consume(capsule):
    """Values of a legacy ``dltensor`` capsule (float64, CPU or ROCm memory)
    as a numpy array, plus the capsule's (device_type, device_id, data
    pointer).  The tensor is released through its deleter."""
    addr = _api.PyCapsule_GetPointer(capsule, b"dltensor")
    mt = DLManagedTensor.from_address(addr)
    t = mt.dl_tensor
    if (t.dtype.code, t.dtype.bits, t.dtype.lanes) != (kDLFloat, 64, 1):
        raise TypeError("float64 expected")
    shape = tuple(t.shape[k] for k in range(t.ndim))
    if t.strides:
        want, s = [], 1
        for d in reversed(shape):
            want.append(s)
            s *= d
        if tuple(t.strides[k] for k in range(t.ndim)) != tuple(reversed(want)):
            raise ValueError("compact row-major tensor expected")
    nbytes = 8 * int(np.prod(shape))
    src = t.data + t.byte_offset
    out = np.empty(shape, dtype=np.float64)
    dev = (t.device.device_type, t.device.device_id, src)
    if t.device.device_type == kDLCPU:
        ctypes.memmove(out.ctypes.data, src, nbytes)
    elif t.device.device_type in (kDLROCM, kDLCUDA):
        rc = hip().hipMemcpy(out.ctypes.data, src, nbytes, D2H)
        if rc != 0:
            raise RuntimeError("hipMemcpy failed: %d" % rc)
    else:
        raise ValueError("device type %d" % t.device.device_type)
    _api.PyCapsule_SetName(capsule, b"used_dltensor")
    if mt.deleter:
        mt.deleter(ctypes.cast(addr, ctypes.POINTER(DLManagedTensor)))
    return out, dev
