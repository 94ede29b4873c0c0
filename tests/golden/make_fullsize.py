"""Writes the full-size parity fixtures from the FP64 / FP32 NumPy oracle.

    python tests/golden/make_fullsize.py          (about 6 minutes on 8 cores)

* ``fullsize_two_dof.npz``: the 4,096 two-DOF systems x 50,000 steps
  (``workloads.two_dof()``, seed 1): final states ``o`` (FP64 oracle),
  ``op`` (FP64 oracle with every coefficient moved one ULP, the sensitivity
  probe), ``o32`` / ``p32`` (the FP32 oracle, plain and one-FP32-ULP
  perturbed).
* ``fullsize_chains.npz``: 256 of the 8,192 16-DOF chains x 20,000 steps
  (``workloads.chains()``, seed 2), ids ``ids``, including chain 3088 (the
  largest lane-vs-tile gradient gap on the GPU): final states ``oq`` and
  gradients ``og`` (FP64 oracle, energy gradient, checkpoint every 256 steps)
  and the one-ULP perturbed ``pq`` / ``pg``.

* ``fullsize_heterogeneous.npz``: a stratified sample of the 100,000-system
  heterogeneous batch (``workloads.heterogeneous()``, seed 3): for every DOF
  size the two systems with the most steps and one drawn at random (ids
  ``ids``), their final states and energy gradients from the FP64 oracle
  (``q64`` / ``g64``) and the FP32 oracle (``q32`` / ``g32``).

The oracle is deterministic and the workloads are seeded, so rerunning the
script reproduces the files bit for bit (``tests/test_fullsize_fixtures.py``
recomputes a few systems and compares).
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

TWO_DOF_FILE = os.path.join(HERE, "fullsize_two_dof.npz")
CHAINS_FILE = os.path.join(HERE, "fullsize_chains.npz")
HET_FILE = os.path.join(HERE, "fullsize_heterogeneous.npz")
CHAIN_SAMPLE = 256
CHAIN_MUST = (3088,)


def chain_ids() -> np.ndarray:
    rng = np.random.default_rng(7)
    pick = rng.choice(8192, CHAIN_SAMPLE, replace=False)
    ids = np.unique(np.concatenate([np.asarray(CHAIN_MUST), pick]))
    return np.sort(ids)[:CHAIN_SAMPLE] if len(ids) > CHAIN_SAMPLE else ids


def two_dof_shard(args):
    """Oracle final states of systems lo..hi-1 of ``two_dof()``."""
    lo, hi, dtype, perturb = args
    w = W.two_dof()
    sl = slice(lo, hi)
    asm = O.assemble_forward(w.vecM[sl], w.vecB[sl], w.vecK[sl], w.vecGamma[sl],
                             w.vecLambda[sl], w.vecForce[sl], w.vecQ0[sl]).astype(dtype)
    if perturb:
        asm = PAR.ulp_perturbed(asm, seed=lo)
    n = hi - lo
    lay = O.mapped_layout(asm.sizes, np.arange(n), np.full(n, w.numsteps))
    r = O.rk4_forward(asm, O.Forcing(w.excitations[sl]), w.sample_rate, lay, dtype=dtype)
    return r.q.reshape(-1).astype(np.float64)


def chains_shard(args):
    """Oracle final states and energy gradients of the listed chains."""
    ids, perturb = args
    w = W.chains()
    J = ADJ.energy_gradient_inputs(w)
    sub = lambda seq: [seq[i] for i in ids]  # noqa: E731
    fwd = O.assemble_forward(sub(w.vecM), sub(w.vecB), sub(w.vecK), sub(w.vecGamma),
                             sub(w.vecLambda), sub(w.vecForce), sub(w.vecQ0))
    adj = O.assemble_adjoint(*[sub(a) for a in J.as_args()], n_fwd=fwd.n)
    if perturb:
        fwd = PAR.ulp_perturbed(fwd, seed=int(ids[0]))
        adj = PAR.ulp_perturbed(adj, seed=int(ids[0]) + 1)
    n = len(ids)
    lay = O.mapped_layout(fwd.sizes, np.arange(n), np.full(n, w.numsteps))
    F = O.Forcing(w.excitations[ids], w.interp)
    oq = O.rk4_forward(fwd, F, w.sample_rate, lay).q[0]
    og = O.adjoint_gradient(fwd, adj, F, w.sample_rate, lay, np.full((1, adj.n), w.numsteps),
                            checkpoint_every=256)[0]
    return oq, og


def heterogeneous_ids(w) -> np.ndarray:
    """Per DOF size: the two systems with the most steps and one at random."""
    rng = np.random.default_rng(11)
    dofs = np.asarray(w.meta["dofs"])
    steps = np.asarray(w.steps_of_system)
    ids = []
    for d in np.unique(dofs):
        members = np.nonzero(dofs == d)[0]
        longest = members[np.argsort(-steps[members], kind="stable")[:2]]
        ids += list(longest) + [int(rng.choice(members))]
    return np.unique(np.asarray(ids, dtype=np.int64))


def heterogeneous_sub(w, ids):
    """The workload restricted to systems ``ids`` (same signals, steps)."""
    pick = lambda seq: [seq[i] for i in ids]  # noqa: E731
    return W.Workload("heterogeneous-sample", pick(w.vecM), pick(w.vecB), pick(w.vecK),
                      pick(w.vecGamma), pick(w.vecLambda), pick(w.vecForce), pick(w.vecQ0),
                      w.excitations, w.sample_rate, int(np.max(w.steps_of_system[ids])),
                      exc_of_system=np.asarray(w.exc_of_system)[ids],
                      steps_of_system=np.asarray(w.steps_of_system)[ids], interp=w.interp,
                      meta={"dofs": np.asarray(w.meta["dofs"])[ids]})


_HET = {}


def _heterogeneous():
    """The batch, built once per process (shard workers fork from a parent
    that built it already)."""
    if "w" not in _HET:
        _HET["w"] = W.heterogeneous()
    return _HET["w"]


def heterogeneous_shard(args):
    """Oracle states and energy gradients of one system of the batch sample."""
    sid, dtype = args
    w = _heterogeneous()
    sub = heterogeneous_sub(w, np.array([sid]))
    J = ADJ.energy_gradient_inputs(sub)
    fwd = O.assemble_forward(sub.vecM, sub.vecB, sub.vecK, sub.vecGamma, sub.vecLambda,
                             sub.vecForce, sub.vecQ0)
    adj = O.assemble_adjoint(*J.as_args(), n_fwd=fwd.n)
    lay = O.mapped_layout(fwd.sizes, sub.exc_of_system, sub.steps_of_system)
    F = O.Forcing(np.asarray(sub.excitations), sub.interp)
    q = O.rk4_forward(fwd, F, sub.sample_rate, lay, dtype=dtype).q[0]
    asteps = np.repeat(sub.steps_of_system, adj.sizes)[None, :]
    g = O.adjoint_gradient(fwd, adj, F, sub.sample_rate, lay, asteps, dtype=dtype,
                           checkpoint_every=256)[0]
    return q.astype(np.float64), g.astype(np.float64)


def two_dof_bounds(n_systems=4096, shards=32):
    b = np.linspace(0, n_systems, shards + 1).astype(int)
    return [(int(b[i]), int(b[i + 1])) for i in range(shards)]


def main():
    t0 = time.perf_counter()
    out = {}
    for key, dtype, perturb in (("o", np.float64, False), ("op", np.float64, True),
                                ("o32", np.float32, False), ("p32", np.float32, True)):
        res = PAR.run_sharded(two_dof_shard, [(lo, hi, dtype, perturb)
                                              for lo, hi in two_dof_bounds()])
        out[key] = np.concatenate(res)
        print(f"two-DOF {key}: {time.perf_counter() - t0:.0f} s", flush=True)
    np.savez_compressed(TWO_DOF_FILE, **out)
    ids = chain_ids()
    parts = np.array_split(ids, 16)
    res = PAR.run_sharded(chains_shard, [(p, False) for p in parts])
    resp = PAR.run_sharded(chains_shard, [(p, True) for p in parts])
    np.savez_compressed(CHAINS_FILE, ids=ids,
                        oq=np.concatenate([r[0] for r in res]),
                        og=np.concatenate([r[1] for r in res]),
                        pq=np.concatenate([r[0] for r in resp]),
                        pg=np.concatenate([r[1] for r in resp]))
    print(f"chains: {time.perf_counter() - t0:.0f} s", flush=True)
    ids = heterogeneous_ids(_heterogeneous())
    out = {"ids": ids}
    for tag, dtype in (("64", np.float64), ("32", np.float32)):
        res = PAR.run_sharded(heterogeneous_shard, [(int(i), dtype) for i in ids])
        out["q" + tag] = np.concatenate([r[0] for r in res])
        out["g" + tag] = np.concatenate([r[1] for r in res])
        print(f"heterogeneous FP{tag}: {time.perf_counter() - t0:.0f} s", flush=True)
    np.savez_compressed(HET_FILE, **out)


if __name__ == "__main__":
    main()
