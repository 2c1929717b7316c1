"""MI355X-native (gfx950 HIP) 3-D elastic FWI propagation path.

Submodules:
  _native     ctypes binding of libefwi.so (fails loudly if it is missing)
  propagator  device-resident context API (GridSpec, ShotSpec, Propagator)
  io          the para/survey JSON and Shot_*.bin contract of the reference
  synth       seeded synthetic workloads
"""
from .propagator import GridSpec, ShotSpec, Propagator  # noqa: F401

__all__ = ["GridSpec", "ShotSpec", "Propagator"]
