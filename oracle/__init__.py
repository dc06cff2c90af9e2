"""FP64 CPU oracle for the TOP3D_XL hot path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker.  The product
(``top3d_xl_amd``) never imports it and has no CPU fallback.

``t3x_oracle.cpp`` restates TOP3D_XL.m (TO/PIO loop bodies and callees) in
C++/OpenMP using the reference's own compact-index data structures.  This
module is a thin ctypes binding to ``liboracle.so`` built by ``Makefile``.
"""
from .binding import Oracle, build_oracle, cuboid_bc_oracle  # noqa: F401
