"""MI355X-native focused-ultrasound kernels (Westervelt, Pennes bioheat, Gilmore).

The compute path is the HIP library ``libbubbledynamics.so`` behind the C ABI
declared in ``include/bubbledynamics.h``; this package is its host-side
binding.  Importing the package does not touch the GPU.
"""
from . import _ffi
from ._ffi import BDError, WesterveltParams, load

__all__ = ["BDError", "WesterveltParams", "load", "_ffi"]
