"""Device buffers for GPU tests through the HIP runtime that librt_hip.so
itself is linked against (found in /proc/self/maps once the library is
loaded), so a test never mixes that runtime with the copy torch bundles."""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        import ray_tracing_amd as rt
        rt.lib()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line and "/opt/rocm" in line:
                    path = line.split()[-1]
                    break
        if path is None:
            raise RuntimeError("librt_hip.so's HIP runtime is not mapped")
        H = C.CDLL(path)
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        H.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        H.hipDeviceSynchronize.argtypes = []
        _hip = H
    return _hip


def sync():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def _ok(e, what):
    if e != 0:
        raise RuntimeError(f"{what} failed with HIP error {e}")


class DeviceArray:
    """A device allocation of nbytes; upload/download through host numpy."""

    H2D, D2H, D2D = 1, 2, 3

    def __init__(self, nbytes: int):
        self.nbytes = nbytes
        self.p = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(self.p), max(nbytes, 1)), "hipMalloc")

    @classmethod
    def of(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        d.upload(a)
        return d

    @property
    def ptr(self) -> int:
        return self.p.value

    def upload(self, a: np.ndarray, offset: int = 0):
        a = np.ascontiguousarray(a)
        _ok(hip().hipMemcpy(C.c_void_p(self.ptr + offset), a.ctypes.data_as(C.c_void_p),
                            a.nbytes, self.H2D), "hipMemcpy H2D")
        sync()      # the library's kernels run on a non-blocking stream

    def copy_within(self, dst_offset: int, src_offset: int, nbytes: int):
        _ok(hip().hipMemcpy(C.c_void_p(self.ptr + dst_offset), C.c_void_p(self.ptr + src_offset),
                            nbytes, self.D2D), "hipMemcpy D2D")
        sync()

    def download(self, shape, dtype, offset: int = 0) -> np.ndarray:
        out = np.empty(shape, dtype)
        _ok(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr + offset),
                            out.nbytes, self.D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
