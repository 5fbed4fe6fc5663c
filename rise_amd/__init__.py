"""MI355X (gfx950) implementation of RISE's ray-cast and spectral
path-tracing hot path.

The compute path is the C ABI of ``librise_hip.so`` (``include/rise_hip.h``),
hand-written HIP kernels behind it.  This package holds the build recipe,
the ctypes binding, seeded workload generators, and a Python mirror of the
reference's object-manager interface.
"""

__all__ = ["raycast", "scenes"]
