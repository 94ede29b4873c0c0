"""CPU oracle for the PASIf RK4 forward / adjoint hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path (``pasif_amd``) may
import this package; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.
"""
from .pasif_oracle import *  # noqa: F401,F403
