"""Device, context and queue singletons (the API of reference ``nofluidx3d/openCL.py:17-50``).

``initializeOpenCLObjects(deviceId)`` selects HIP device ``deviceId`` and creates the libnfx
context with its own default-priority stream; ``getCommandQueue().finish()`` waits for it.  As
in the reference, a missing device id prints an error and exits with status 1
(``openCL.py:30-32``).
"""

from __future__ import annotations

import ctypes

from nofluidx3d_amd import _lib


class Device:
    def __init__(self, device_id: int, name: str):
        self.id = device_id
        self.name = name

    def __repr__(self):
        return f"Device({self.id}, {self.name!r})"


class Context:
    """Owns the libnfx context (``nfx_ctx*``) and the torch device its buffers live on."""

    def __init__(self, device_id: int):
        import torch

        self.device_id = device_id
        self.torch_device = torch.device("cuda", device_id)
        torch.cuda.set_device(device_id)
        h = ctypes.c_void_p()
        _lib.call("nfx_init", device_id, None, ctypes.byref(h))
        self.handle = h.value
        buf = ctypes.create_string_buffer(256)
        _lib.call("nfx_device_name", self.handle, buf, 256)
        self.name = buf.value.decode()

    def release(self):
        if self.handle:
            _lib.lib().nfx_destroy(self.handle)
            self.handle = None


class CommandQueue:
    """In-order queue = the context's stream (reference ``openCL.py:34``)."""

    def __init__(self, context: Context):
        self.context = context

    def finish(self):
        import torch

        _lib.call("nfx_stream_sync", self.context.handle)
        torch.cuda.synchronize(self.context.torch_device)


device: Device = None
context: Context = None
commandQueue: CommandQueue = None


def initializeOpenCLObjects(deviceId: int):
    global device, context, commandQueue
    if context is not None and context.device_id == deviceId:
        return
    count = ctypes.c_int(0)
    _lib.call("nfx_device_count", ctypes.byref(count))
    if deviceId < 0 or deviceId >= count.value:
        print(f"ERROR: Device ID {deviceId} does not exist!\n")
        exit(1)
    if context is not None:
        context.release()
    context = Context(deviceId)
    device = Device(deviceId, context.name)
    commandQueue = CommandQueue(context)


def getDevice() -> Device:
    return device


def getContext() -> Context:
    return context


def getCommandQueue() -> CommandQueue:
    return commandQueue


def handle():
    """The ``nfx_ctx*`` of the active context (raises if none)."""
    if context is None:
        raise _lib.NfxError("initializeOpenCLObjects() has not been called")
    return context.handle


__all__ = ["initializeOpenCLObjects", "getDevice", "getContext", "getCommandQueue"]
