"""Device buffers for GPU tests that drive the plan API directly."""
import ctypes

import numpy as np


class DeviceArray:
    """int32 device buffer from the HIP runtime the library links against
    (libamdhip64.so.7, already loaded by it, so one runtime instance)."""

    def __init__(self, shape):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 4
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(self.nbytes)) == 0

    def get(self):
        out = np.empty(self.shape, np.int32)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(self.nbytes), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)
