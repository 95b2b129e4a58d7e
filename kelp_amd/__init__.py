"""kelp_amd -- MI355X-native particle time-stepping for KELP.

The hot path (localization, interpolation, Euler/RK2/RK4 stepping, random
walk, beaching) runs in hand-written HIP kernels for gfx950
(kelp_amd/csrc/*.hip) behind a C ABI (include/kelp_hip.h) loaded with
ctypes.  The Python modules mirror the reference's host surface: the mesh
container and its grid hash (mesh.py), the particle state (particles.py),
the integrator registry (integrators.py) and the run loop (tracking.py).
"""
__version__ = "0.1.0"
