"""MI355X-native (gfx950) DPM2D deformable-ring vertex engine.

The compute path is libdpmring.so (hand-written HIP kernels behind the C ABI of
include/dpmring.h); `cudaring_amd.dpm.DPMSet` binds it from Python and the C++
classes under cudaring_amd/host mirror the reference's DPM2D / FIRE /
SimInterface API for C++ drivers.
"""
from .dpm import DPMSet, DPMError, load_library  # noqa: F401
