"""The concatenated adjoint assembler (``jacobian_blocks``) hands the driver
exactly what ``setJacobian`` unfolds from the per-system inputs of
``jacobian_inputs``: the same arrays, entry for entry, for chains of every
size in one batch and for mixed parameter kinds."""
import os

import numpy as np

from pasif_amd import adjoint as ADJ
from pasif_amd import workloads as W


def host_solver(w):
    os.environ["LOCAL_RANK"] = "-1"
    try:
        return W.make_solver(w)
    finally:
        del os.environ["LOCAL_RANK"]


def _state(p):
    K = p.jacobian_K.tocoo()
    out = {"sizes": np.asarray(p.jacobian_sizes), "K": (K.row, K.col, K.data, K.shape),
           "f": p.jacobian_forcePattern, "q0": p.jacobian_initialConditions}
    for name in ("Gamma", "Lambda", "Psi"):
        t = getattr(p, "jacobian_" + name)
        out[name] = (list(t.dimensions), np.asarray(t.indices), np.asarray(t.val))
    return out


def _same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        x, y = a[k], b[k]
        if isinstance(x, tuple):
            for u, v in zip(x, y):
                assert np.array_equal(np.asarray(u), np.asarray(v)), k
        else:
            assert np.array_equal(x, y), k


def test_energy_gradient_blocks_equal_per_system_inputs():
    w = W.heterogeneous(40, 30, 300, steps_per_sample=5, n_signals=3,
                        dof_choices=(1, 2, 4, 8), dof_probs=(1, 1, 1, 1))
    a = host_solver(w)
    a.setJacobian(*ADJ.energy_gradient_inputs(w).as_args())
    b = host_solver(w)
    b.setJacobianBlocks(ADJ.energy_gradient_blocks(w))
    _same(_state(a), _state(b))
    assert a.kernelInfo(True) == b.kernelInfo(True)


def test_blocks_with_every_parameter_kind():
    w = W.chains(6, 3, 60, steps_per_sample=10)
    params = []
    for s in range(w.n_systems):
        ng = len(np.asarray(w.vecGamma[s].val).reshape(-1))
        params.append([("B", 3 + s % 3, 3), ("Gamma", s % ng), ("K", 4, 1), ("Lambda", 2)])
    a = host_solver(w)
    a.setJacobian(*ADJ.energy_gradient_inputs(w, params).as_args())
    b = host_solver(w)
    b.setJacobianBlocks(ADJ.energy_gradient_blocks(w, params))
    _same(_state(a), _state(b))
