"""rocm device object and the process-global active device
(sqaodpy/sqaod/cuda/device.py:1-37)."""
import atexit
import sys

from ._cext import rocm_device


class Device:
    _cext = rocm_device

    def __init__(self, devno=0):
        self._cobj = rocm_device.new(devno)

    def __del__(self):
        if getattr(self, '_cobj', None) is not None:
            self.finalize()

    def finalize(self):
        if self._cobj is not None:
            rocm_device.delete(self._cobj)
            self._cobj = None

    def synchronize(self):
        rocm_device.synchronize(self._cobj)


def create_device(devno=0):
    return Device(devno)


this_module = sys.modules[__name__]
this_module.active_device = None


def get_active_device():
    """The active device, created on first use (the reference creates it at import)."""
    if this_module.active_device is None:
        this_module.active_device = create_device()
    return this_module.active_device


def unload():
    this_module.active_device = None


atexit.register(unload)
