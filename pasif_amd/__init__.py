"""MI355X (gfx950) implementation of PASIf's batched RK4 forward integrator
and its checkpointed adjoint, behind PASIf's own Python API.

``from pasif_amd import PASIf, coo_tensor`` is the drop-in for the
reference's ``from PASIf import *``.  The compute runs in
``libpasif_hip.so`` (hand-written HIP for gfx950, C ABI in
``include/pasif.h``); build it with ``python -m pasif_amd.build``.
"""
from .PASIf import PASIf, coo_tensor  # noqa: F401
from ._lib import PasifError  # noqa: F401

__version__ = "0.1.0"
