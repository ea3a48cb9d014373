"""MI355X implementation of SDE's imaging hot path.

The compute lives in libsdegpu.so (hand-written HIP for gfx950, built by
``sde_amd.build``).  This package is the host-side mirror of the reference
interface:

  * ``sde_amd.grd``      the pixel-level routines with the reference's
                         argument lists (GRDH2D, GRDDEH2D, GRDUWT2D, ...),
                         called through the Fortran ABI of the library;
  * ``sde_amd.imaging``  uvmap / vismodel drivers on a device-resident plan;
  * ``sde_amd.synth``    seeded synthetic observations of the benchmark
                         workloads.

There is no CPU fallback: importing a compute module without the built
library raises.
"""
import ctypes as _C
import os as _os

_HERE = _os.path.dirname(_os.path.abspath(__file__))
LIBPATH = _os.path.join(_HERE, "libsdegpu.so")

_lib = None


class SdeError(RuntimeError):
    """Raised when the library reports an error (the reference's ERRREPOR)."""


def lib():
    """The loaded libsdegpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not _os.path.exists(LIBPATH):
            raise ImportError("libsdegpu.so is not built (run python -m sde_amd.build): "
                              "the HIP implementation is required, there is no CPU path")
        _lib = _C.CDLL(LIBPATH)
        _declare(_lib)
    return _lib


def _declare(L):
    L.sde_error_flag.restype = _C.c_int
    L.sde_last_error.restype = _C.c_char_p
    L.sde_last_warning.restype = _C.c_char_p
    L.sde_device_count.restype = _C.c_int
    L.sde_version.restype = _C.c_char_p
    L.sde_plan_device_bytes.restype = _C.c_int64
    L.sde_plan_estimate_bytes.restype = _C.c_int64
    L.sde_grdh23d_device.restype = _C.c_int
    L.sde_vistoimg_slowz_device.restype = _C.c_int
    L.sde_imgtovis_slowz_device.restype = _C.c_int64
    L.sde_device_alloc.restype = _C.c_void_p
    L.sde_device_alloc.argtypes = [_C.c_size_t]
    L.sde_device_free.argtypes = [_C.c_void_p]
    L.sde_copy.argtypes = [_C.c_void_p, _C.c_void_p, _C.c_size_t]
    L.sde_grddeh23_device.restype = _C.c_int64
    L.sde_grdh2d_device_noalloc.restype = _C.c_int
    L.sde_noalloc_scratch_bytes.restype = _C.c_size_t
    L.sde_sort_pairs_device.restype = _C.c_int
    L.sde_slowz_staged_blocks.restype = _C.c_longlong
    L.sde_host_alloc.restype = _C.c_void_p
    L.sde_host_alloc.argtypes = [_C.c_size_t]
    L.sde_host_free.argtypes = [_C.c_void_p]


def pinned_empty(shape, dtype):
    """A numpy array in pinned (page-locked) host memory, for asynchronous
    transfers; the memory is freed when the last view of it goes."""
    import weakref
    import numpy as np
    dtype = np.dtype(dtype)
    n = int(np.prod(shape)) * dtype.itemsize
    ptr = lib().sde_host_alloc(n)
    if not ptr:
        raise MemoryError("sde_host_alloc: %d bytes: %s" % (n, last_error()))
    buf = (_C.c_char * max(n, 1)).from_address(ptr)
    weakref.finalize(buf, lib().sde_host_free, ptr)
    return np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)


class DeviceArray:
    """An array in device memory of the library's HIP runtime, for the
    *_device entry points (torch's bundled HIP runtime is a separate one
    here: its tensors are not handed to the library)."""

    def __init__(self, shape, dtype):
        import weakref
        import numpy as np
        self.shape = tuple(int(x) for x in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = lib().sde_device_alloc(self.nbytes)
        if not self.ptr:
            raise MemoryError("sde_device_alloc: %d bytes: %s" % (self.nbytes, last_error()))
        self._fin = weakref.finalize(self, lib().sde_device_free, _C.c_void_p(self.ptr))

    @classmethod
    def from_numpy(cls, a):
        import numpy as np
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        import numpy as np
        a = np.ascontiguousarray(a, self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        if lib().sde_copy(_C.c_void_p(self.ptr), a.ctypes.data_as(_C.c_void_p), self.nbytes) != 0:
            raise SdeError("sde_copy: " + last_error())

    def copy_from(self, other):
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        if lib().sde_copy(_C.c_void_p(self.ptr), _C.c_void_p(other.ptr), self.nbytes) != 0:
            raise SdeError("sde_copy: " + last_error())

    def numpy(self):
        import numpy as np
        out = np.empty(self.shape, self.dtype)
        if lib().sde_copy(out.ctypes.data_as(_C.c_void_p), _C.c_void_p(self.ptr), self.nbytes) != 0:
            raise SdeError("sde_copy: " + last_error())
        return out

    def free(self):
        self._fin()


def device_count():
    return lib().sde_device_count()


def error_flag():
    return bool(lib().sde_error_flag())


def clear_error():
    lib().sde_error_clear()


def last_error():
    return lib().sde_last_error().decode()


def last_warning():
    return lib().sde_last_warning().decode()


def check_error(routine=""):
    """Raise SdeError if the sticky ERROR flag is set (and clear it)."""
    if lib().sde_error_flag():
        msg = last_error()
        clear_error()
        raise SdeError(msg or ("error in " + routine))
