"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on CPU.

Both the HIP library and the CPU oracle are loaded from the tree; neither
falls back to the other."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    import oracle_bmx
    oracle_bmx.lib()
    return oracle_bmx


@pytest.fixture(scope="session")
def bmx():
    import boltzmannmfx_amd
    boltzmannmfx_amd.load_library()
    return boltzmannmfx_amd


@pytest.fixture(scope="session")
def gpu(bmx):
    """A HIP device; a GPU test without one fails loudly rather than skipping."""
    ctx = bmx.Context(0)
    ctx.close()
    return True


class _DLTensorView:
    """DLPack producer over a HipArray (device type kDLROCM), for checking
    the binding's DLPack path with the same HIP runtime the library uses (a
    framework in the same process may bring its own runtime)."""

    def __init__(self, arr):
        import ctypes as C
        from boltzmannmfx_amd import _DLTensor
        self._arr = arr
        self._shape = (C.c_int64 * len(arr.shape))(*arr.shape)
        t = _DLTensor()
        t.data = arr.ptr.value
        t.device.device_type = 10
        t.device.device_id = 0
        t.ndim = len(arr.shape)
        t.dtype.code, t.dtype.bits, t.dtype.lanes = 2, 64, 1
        t.shape = C.cast(self._shape, C.POINTER(C.c_int64))
        t.strides = None
        t.byte_offset = 0
        # DLManagedTensor: the DLTensor, then manager_ctx and deleter (unset)
        self._managed = (C.c_byte * (C.sizeof(_DLTensor) + 2 * C.sizeof(C.c_void_p)))()
        C.memmove(self._managed, C.byref(t), C.sizeof(_DLTensor))

    def __dlpack_device__(self):
        return (10, 0)

    def __dlpack__(self, stream=None):
        import ctypes as C
        new = C.pythonapi.PyCapsule_New
        new.restype = C.py_object
        new.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
        return new(C.addressof(self._managed), b"dltensor", None)


class HipArray:
    """float64 device memory (hipMalloc through ctypes) exposing
    __cuda_array_interface__, for passing device meshes to the C ABI without
    a framework."""

    _hip = None

    @classmethod
    def _lib(cls):
        import ctypes as C
        if cls._hip is None:
            # the runtime the library links (torch may have loaded another)
            for name in ("/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"):
                try:
                    cls._hip = C.CDLL(name)
                    break
                except OSError:
                    continue
        return cls._hip

    def __init__(self, shape):
        import ctypes as C
        import numpy as np
        self.shape = tuple(int(v) for v in shape)
        self.nbytes = 8 * int(np.prod(self.shape))
        self.ptr = C.c_void_p()
        assert self._lib().hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)) == 0

    def dlpack(self):
        """This array seen only through the DLPack protocol."""
        return _DLTensorView(self)

    @classmethod
    def upload(cls, a):
        import ctypes as C
        import numpy as np
        a = np.ascontiguousarray(a, np.float64)
        d = cls(a.shape)
        assert cls._lib().hipMemcpy(d.ptr, a.ctypes.data_as(C.c_void_p),
                                    C.c_size_t(d.nbytes), 1) == 0   # host to device
        return d

    def download(self):
        import ctypes as C
        import numpy as np
        out = np.empty(self.shape)
        assert self._lib().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr,
                                     C.c_size_t(self.nbytes), 2) == 0  # device to host
        return out

    @property
    def __cuda_array_interface__(self):
        return {"shape": self.shape, "typestr": "<f8", "data": (self.ptr.value, False),
                "version": 2, "strides": None}

    def free(self):
        if self.ptr:
            self._lib().hipFree(self.ptr)
            self.ptr = None
