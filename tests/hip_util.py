"""Device buffers for GPU tests through the HIP runtime's C API (ctypes):
the tests own the device memory they hand to device-pointer entry points."""
import ctypes
import os



class Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime (the test owns the
    device buffers it hands to the entry point)."""

    def __init__(self):
        path = "libamdhip64.so"
        if os.path.exists("/opt/rocm/lib/libamdhip64.so"):
            path = "/opt/rocm/lib/libamdhip64.so"
        self.L = ctypes.CDLL(path)
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_int]
        self.L.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t),
                                         ctypes.POINTER(ctypes.c_size_t)]
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(int(nbytes), 1)) == 0
        self.bufs.append(p.value)
        return p.value

    def h2d(self, dst, a):
        assert self.L.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1) == 0

    def d2h(self, a, src):
        assert self.L.hipMemcpy(a.ctypes.data, src, a.nbytes, 2) == 0

    def free_bytes(self):
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        assert self.L.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
        return f.value

    def close(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []
