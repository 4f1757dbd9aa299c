"""metakonga_amd — MI355X-native DEM contact hot path (xdynamics DEM step).

The compute path is libxdem.so (HIP for gfx950, C ABI in include/xdem.h).
This package holds the ctypes binding (abi.py), the host-side mirror of the
reference's grid/integrator/contact objects (xdem.py), the seeded scene
generators (scenes.py) and the build recipe (build.py).
"""
__all__ = ["abi", "scenes", "xdem", "build"]
