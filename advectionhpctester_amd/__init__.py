"""MI355X-native MPDATA advection step (EULAG dwarf), drop-in for the
reference's advection driver call.

Submodules
----------
mpdata    GPU host interface (``Domain``) over libmpdata_hip.so
testdata  initial conditions and synthetic flows of the tests/benchmarks
_lib      ctypes binding of the C ABI (include/mpdata_hip.h)

``mpdata`` loads the HIP library on import and raises if it is missing.
"""
__all__ = ["mpdata", "testdata"]
__version__ = "0.1.0"
