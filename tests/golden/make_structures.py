"""Writes ``fullsize_structures.npz``: a stratified sample of the
100,000-system heterogeneous batch with 66 sparsity structures
(``workloads.heterogeneous(variants_per_size=12)``, seed 3): for every DOF
size the two systems with the most steps and two drawn at random (ids
``ids``, variants ``variants``), their final states ``q64`` and energy
gradients ``g64`` (checkpoint every 256 steps) from the FP64 oracle.

    python tests/golden/make_structures.py        (about 4 minutes on 8 cores)

Deterministic: rerunning reproduces the file bit for bit
(``tests/test_fullsize_fixtures.py`` recomputes one system and compares).
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from oracle import parity as PAR  # noqa: E402
from pasif_amd import adjoint as ADJ  # noqa: E402
from pasif_amd import workloads as W  # noqa: E402
from tests.golden.make_fullsize import heterogeneous_sub  # noqa: E402

OUT_FILE = os.path.join(HERE, "fullsize_structures.npz")
VARIANTS = 12

_W = {}


def batch():
    """The batch, built once per process (shard workers fork from a parent
    that built it already)."""
    if "w" not in _W:
        _W["w"] = W.heterogeneous(variants_per_size=VARIANTS)
    return _W["w"]


def sample_ids(w) -> np.ndarray:
    """Per DOF size: the two systems with the most steps and two at random."""
    rng = np.random.default_rng(12)
    dofs = np.asarray(w.meta["dofs"])
    steps = np.asarray(w.steps_of_system)
    ids = []
    for d in np.unique(dofs):
        members = np.nonzero(dofs == d)[0]
        longest = members[np.argsort(-steps[members], kind="stable")[:2]]
        ids += list(longest) + [int(x) for x in rng.choice(members, 2, replace=False)]
    return np.unique(np.asarray(ids, dtype=np.int64))


def oracle_system(w, sid):
    """FP64 oracle final state and energy gradient of system ``sid``."""
    sub = heterogeneous_sub(w, np.array([sid]))
    J = ADJ.energy_gradient_inputs(sub)
    fwd = O.assemble_forward(sub.vecM, sub.vecB, sub.vecK, sub.vecGamma, sub.vecLambda,
                             sub.vecForce, sub.vecQ0)
    adj = O.assemble_adjoint(*J.as_args(), n_fwd=fwd.n)
    lay = O.mapped_layout(fwd.sizes, sub.exc_of_system, sub.steps_of_system)
    F = O.Forcing(np.asarray(sub.excitations), sub.interp)
    q = O.rk4_forward(fwd, F, sub.sample_rate, lay).q[0]
    asteps = np.repeat(sub.steps_of_system, adj.sizes)[None, :]
    g = O.adjoint_gradient(fwd, adj, F, sub.sample_rate, lay, asteps, checkpoint_every=256)[0]
    return q, g


def shard(sid):
    return oracle_system(batch(), int(sid))


def main():
    t0 = time.perf_counter()
    w = batch()
    ids = sample_ids(w)
    res = PAR.run_sharded(shard, [int(i) for i in ids])
    np.savez_compressed(OUT_FILE, ids=ids, variants=np.asarray(w.meta["variants"])[ids],
                        q64=np.concatenate([r[0] for r in res]),
                        g64=np.concatenate([r[1] for r in res]))
    print(f"structures: {len(ids)} systems, {time.perf_counter() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main()
