"""MI355X-native (gfx950 HIP) implementation of the TOP3D_XL topology-optimisation hot path.

The compute lives in ``libt3x.so`` (C ABI declared in ``include/t3x.h``);
``t3x`` holds the ctypes binding and the ``TOP3D_XL`` drop-in mirror.
"""
from .t3x import (TOP3D_XL, Problem, Server, T3XError, load_library, default_params, footprint,  # noqa: F401
                  stream_copy_ms, oc_noalloc, device_synchronize, EXPORTED)
from . import topvoxel  # noqa: F401,E402
