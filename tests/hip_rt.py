"""Device buffers for GPU tests through the HIP runtime already loaded with the library (no second
runtime in the process)."""
import ctypes as C

import numpy as np


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime already loaded with the library."""
    H2D, D2H = 1, 2

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]

    def put(self, a: np.ndarray) -> C.c_void_p:
        a = np.ascontiguousarray(a)
        ptr = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(ptr), max(a.nbytes, 1)) == 0
        assert self.rt.hipMemcpy(ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, self.H2D) == 0
        return ptr

    def get(self, ptr: C.c_void_p, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), ptr, out.nbytes, self.D2H) == 0
        return out

    def write(self, ptr, a: np.ndarray):
        """Copy a host array into existing device memory."""
        a = np.ascontiguousarray(a)
        assert self.rt.hipMemcpy(C.c_void_p(ptr) if isinstance(ptr, int) else ptr, a.ctypes.data_as(C.c_void_p),
                                 a.nbytes, self.H2D) == 0

    def read(self, ptr, like: np.ndarray) -> np.ndarray:
        return self.get(C.c_void_p(ptr) if isinstance(ptr, int) else ptr, like)

    def free(self, *ptrs):
        for q in ptrs:
            assert self.rt.hipFree(q) == 0
