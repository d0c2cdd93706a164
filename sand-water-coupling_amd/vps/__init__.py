"""Python binding of libvps, the MI355X virtual-particle SPH projection.

The library itself is C++/HIP behind the C ABI in include/vps.h; this module
is a thin ctypes layer used by tests and the benchmark.  ``Projection``
mirrors the reference's DualParticleSphLiquid / VirtualParticle call order
(Cores/Physics/DualParticleSphLiquid.cpp:78-126).
"""
from .core import DemSand, Projection, VpsError, load_library  # noqa: F401
from . import scenes  # noqa: F401
