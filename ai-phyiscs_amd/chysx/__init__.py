"""ChysX implicit cloth step on AMD Instinct MI355X (gfx950).

Drop-in for the reference's cloth path (``chysx/__init__.py:36-51``):
:class:`ClothMaterial` and :class:`ClothSimulator` keep the reference's
pybind11 surface (``set_material``, ``set_external_buffers`` on raw device
pointers, ``step(dt)``, pins, springs, stretch, ``debug_dump_last_solve``).
The step runs in hand-written HIP kernels for gfx950 behind the C ABI of
``libchysx_amd.so`` (``include/chysx_cloth.h``).

Additions: ``set_pcg_tolerance`` (stop at a true relative residual),
``set_springs_active``, ``set_external_force_active``, per-element stiffness,
``last_solve_stats``, warm-start control, :class:`ClothBatch` (many cloths in
one launch) and :mod:`chysx.warmstart` (learned PCG initial guess).

Importing fails loudly when the native module has not been built.
"""

from __future__ import annotations

try:
    # One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64 /
    # libhsa-runtime64 under the same sonames.  Loading it first makes this
    # library bind to that runtime, so device pointers from torch tensors are
    # valid here; loading /opt/rocm's copy first would start a second runtime.
    import torch  # noqa: F401
except ImportError:  # no PyTorch: the system ROCm runtime is used
    pass

from . import _chysx_native

ClothMaterial = _chysx_native.ClothMaterial
ClothSimulator = _chysx_native.ClothSimulator
ClothBatch = _chysx_native.ClothBatch
topology_query = _chysx_native.topology_query

__version__ = "0.1.0"

__all__ = [
    "ClothBatch",
    "ClothMaterial",
    "ClothSimulator",
    "topology_query",
    "__version__",
]
