"""Host array + device buffer pair (the API of reference ``nofluidx3d/util/Bridge.py:16-26``).

``arrayC`` is the host numpy array in ``datatype`` (as in the reference).  ``buf`` is a torch
tensor on the active HIP device.  Floating-point state is held in FP32 on the device
(``device_dtype`` defaults to float32 for float64 host data); unsigned 32-bit indices are held
bit-for-bit in an int32 tensor.
"""

from __future__ import annotations

import numpy as np

from nofluidx3d_amd.openCL import getCommandQueue, getContext

_DEVICE_DTYPE = {
    np.dtype(np.float64): np.float32,
    np.dtype(np.float32): np.float32,
    np.dtype(np.uint32): np.uint32,
    np.dtype(np.int32): np.int32,
}


def _torch_from(arr: np.ndarray):
    import torch

    if arr.dtype == np.uint32:
        arr = arr.view(np.int32)
    return torch.from_numpy(np.ascontiguousarray(arr))


class Bridge:
    def __init__(self, startarray, datatype, device_dtype=None):
        import torch

        self.arrayC = np.asarray(startarray).astype(datatype)
        self.device_dtype = np.dtype(device_dtype or _DEVICE_DTYPE.get(self.arrayC.dtype, self.arrayC.dtype))
        ctx = getContext()
        if ctx is None:
            raise RuntimeError("initializeOpenCLObjects() must run before a Bridge is created")
        self.buf = _torch_from(self.arrayC.astype(self.device_dtype)).to(ctx.torch_device)
        torch.cuda.synchronize(ctx.torch_device)

    @property
    def ptr(self) -> int:
        return self.buf.data_ptr()

    def readFromGPU(self):
        getCommandQueue().finish()
        host = self.buf.cpu().numpy()
        if self.device_dtype == np.uint32:
            host = host.view(np.uint32)
        self.arrayC[...] = host.astype(self.arrayC.dtype)
        return self.arrayC

    def writeToGPU(self):
        import torch

        getCommandQueue().finish()
        self.buf.copy_(_torch_from(self.arrayC.astype(self.device_dtype)))
        torch.cuda.synchronize(self.buf.device)
