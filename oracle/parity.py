"""Parity metric and oracle runs at full workload sizes.

TEST INFRASTRUCTURE ONLY (the checker): used by ``tests/``, the golden
fixture script ``tests/golden/make_fullsize.py`` and the parity / CPU-baseline
legs of ``bench.py``.

Metric.  Per output ``i`` of system ``s``::

    e_i = |g_i - o_i| / max(|o_i|, FLOOR * max_{j in s} |o_j|)

with ``FLOOR = 1e-4``: outputs that sit near zero (a state caught at a zero
crossing at the final step, a gradient component four or more orders below
the system's largest) are measured against the system's own scale.  Why that
floor: perturbing every coefficient of the FP64 oracle by one ULP (a change
smaller than any implementation's rounding, and the reference's atomic
summation order is not fixed either, ``src/kernels.cu:52-61``) moves the
oracle's own outputs by up to 7.5e-10 per output without a floor on the
4,096 two-DOF systems x 50,000 steps (at x1 = 1.7e-6 of an oscillation of
amplitude 0.066), and by at most 5.2e-11 with it; on 256 chains of 16 DOF x
20,000 steps by 1.3e-9 (gradients) / 1.4e-10 (states) without it and 5.9e-12
/ 2.4e-11 with it.  Below the floor the outputs are not determined to 1e-10
by the problem itself.  ``sensitivity`` reproduces that measurement.
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import Callable, Optional, Sequence

import numpy as np

from . import pasif_oracle as O

FLOOR = 1e-4
TOL64 = 1e-10


@dataclass
class ParityReport:
    max_rel: float          # max e_i under the per-system floor
    argmax_system: int
    argmax_output: int      # index inside its system
    value: float            # o_i at the argmax
    system_scale: float     # max |o_j| of that system
    unfloored_max: float    # max |g_i - o_i| / |o_i| (no floor)
    unfloored_system: int
    unfloored_output: int
    per_system: np.ndarray  # max e_i of every system

    def as_dict(self) -> dict:
        return {"max_rel_err": self.max_rel, "argmax": {
            "system": self.argmax_system, "output": self.argmax_output,
            "oracle_value": self.value, "system_max_abs": self.system_scale},
            "unfloored_max_rel_err": self.unfloored_max,
            "unfloored_argmax": {"system": self.unfloored_system,
                                 "output": self.unfloored_output},
            "floor": f"{FLOOR:g} x max |o| of the same system"}


def _offsets(sizes: Sequence[int]) -> np.ndarray:
    return np.concatenate([[0], np.cumsum(np.asarray(sizes, dtype=np.int64))])


def system_scale(o: np.ndarray, sizes: Sequence[int]) -> np.ndarray:
    """``max_{j in s} |o_j|`` broadcast to every output of system ``s``."""
    o = np.abs(np.asarray(o, dtype=np.float64).reshape(-1))
    off = _offsets(sizes)
    if off[-1] != o.size:
        raise ValueError(f"system sizes cover {off[-1]} outputs, got {o.size}")
    sizes = np.diff(off)
    mx = np.maximum.reduceat(o, off[:-1]) if o.size else np.zeros(0)
    return np.repeat(mx, sizes)


def rel_err(g, o, sizes: Sequence[int], floor: float = FLOOR) -> np.ndarray:
    """Per-output relative error under the per-system near-zero floor."""
    g = np.asarray(g, dtype=np.float64).reshape(-1)
    o = np.asarray(o, dtype=np.float64).reshape(-1)
    if g.shape != o.shape:
        raise ValueError(f"shape mismatch {g.shape} vs {o.shape}")
    den = np.maximum(np.abs(o), floor * system_scale(o, sizes))
    return np.abs(g - o) / np.maximum(den, 1e-300)


def report(g, o, sizes: Sequence[int], floor: float = FLOOR) -> ParityReport:
    g = np.asarray(g, dtype=np.float64).reshape(-1)
    o = np.asarray(o, dtype=np.float64).reshape(-1)
    e = rel_err(g, o, sizes, floor)
    off = _offsets(sizes)
    sys_of = np.repeat(np.arange(len(off) - 1), np.diff(off))
    i = int(np.argmax(e))
    scale = system_scale(o, sizes)
    eu = np.abs(g - o) / np.maximum(np.abs(o), 1e-300)
    j = int(np.argmax(eu))
    per = np.maximum.reduceat(e, off[:-1])
    return ParityReport(float(e[i]), int(sys_of[i]), int(i - off[sys_of[i]]), float(o[i]),
                        float(scale[i]), float(eu[j]), int(sys_of[j]),
                        int(j - off[sys_of[j]]), per)


def ulp_perturbed(asm: "O.Assembled", seed: int) -> "O.Assembled":
    """The same system with every coefficient (A, Gamma, Lambda, Psi values)
    moved by one ULP of its own precision, up or down at random."""
    rng = np.random.default_rng(seed)

    def bump(v):
        v = np.asarray(v)
        if v.size == 0:
            return v.copy()
        to = np.where(rng.integers(0, 2, v.shape) == 1, np.inf, -np.inf).astype(v.dtype)
        return np.nextafter(v, to)

    A = asm.A.copy()
    A.data = bump(A.data)
    return O.Assembled(n=asm.n, A=A, gamma_idx=asm.gamma_idx, gamma_val=bump(asm.gamma_val),
                       lam_idx=asm.lam_idx, lam_val=bump(asm.lam_val), psi_idx=asm.psi_idx,
                       psi_val=bump(asm.psi_val), force=asm.force, q0=asm.q0,
                       n_set=asm.n_set, sizes=asm.sizes)


def sensitivity(o, o_perturbed, sizes, floor: float = FLOOR) -> dict:
    """How far the oracle's own outputs move under the one-ULP perturbation,
    without and with the floor."""
    return {"unfloored": float(np.max(rel_err(o_perturbed, o, sizes, 0.0))),
            "floored": float(np.max(rel_err(o_perturbed, o, sizes, floor)))}


def run_sharded(fn: Callable, shards: Sequence, processes: Optional[int] = None) -> list:
    """``[fn(s) for s in shards]`` over a process pool (the oracle is
    single-threaded NumPy and SciPy sparse; one process per host core)."""
    import multiprocessing as mp
    processes = processes or os.cpu_count() or 1
    if processes <= 1 or len(shards) <= 1:
        return [fn(s) for s in shards]
    ctx = mp.get_context("fork")
    with ctx.Pool(min(processes, len(shards))) as pool:
        return pool.map(fn, shards)
