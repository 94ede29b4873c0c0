"""Assembler of the adjoint ("Jacobian") system that ``PASIf.setJacobian``
takes, derived from the forward system.

The reference leaves this to the upstream modelling tool: the user supplies
M', B', K', Gamma', Lambda', Psi' whose tensors mix adjoint indices (first two
axes) and forward indices (remaining axes) (``PASIf.py:279-369, 459-472``;
examples at ``examples/FullInterfaceTesting.py:147-230``).  For the forward
RHS ``f(x) = A x + G(x, x) + L(x, x, x) + P u`` (A, G, L already scaled by
``-M^{-1}``) the adjoint state is ``[lambda (n) ; g (p)]`` and integrates, in
reversed time ``s = T - t``,

    lambda' = J(x)^T lambda,          g_p' = lambda^T df/dtheta_p

so that with ``lambda(T) = dL/dx(T)`` and ``g(T) = 0`` the result at t = 0 is
``[dL/dx(0) ; dL/dtheta]``.  Term by term:

* ``A[o, j]``            -> ``A'[j, o]``
* ``G[o, j, k]``         -> ``G'[j, o, k]`` and ``G'[k, o, j]``
* ``L[o, j, k, l]``      -> ``L'[j, o, k, l]``, ``L'[k, o, j, l]``, ``L'[l, o, j, k]``
* parameter = user entry of K or B at (r, c):  ``G'[g_p, r, c] += D_r``
* parameter = user Gamma entry (o, j, k):      ``L'[g_p, o, j, k] += D_o``
* parameter = user Lambda entry (o, j, k, l):  ``Psi'[g_p, o, j, k, l] += D_o``

with ``D = -1/m``.  The adjoint mass is ``M' = -I`` so that PASIf's ``-M'^{-1}``
scaling (``PASIf.py:620-633``) is the identity and the values pass as built.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np
from scipy.sparse import coo_matrix, dia_matrix

from .PASIf import coo_tensor

# a parameter: ("K", r, c) | ("B", r, c) | ("Gamma", entry) | ("Lambda", entry)
Param = Tuple


@dataclass
class JacobianInputs:
    vecM: list
    vecB: list
    vecK: list
    vecGamma: list
    vecLambda: list
    vecForce: list
    vecLamT: list
    vecPsi: list
    sizes: np.ndarray

    def as_args(self):
        return (self.vecM, self.vecB, self.vecK, self.vecGamma, self.vecLambda, self.vecForce,
                self.vecLamT, self.vecPsi)


def _entries(t, order):
    v = np.asarray(t.val, dtype=np.float64).reshape(-1)
    ix = (np.asarray(t.indices, dtype=np.int64).reshape(-1, order) if v.size
          else np.zeros((0, order), dtype=np.int64))
    return ix, v


def _tensor(dims, idx: List[Sequence[int]], val: List[float]):
    t = coo_tensor(list(dims))
    t.val = np.asarray(val, dtype=np.float64)
    t.indices = (np.asarray(idx, dtype=np.int64).reshape(-1) if len(idx)
                 else np.zeros(0, dtype=np.int64))
    return t


def system_jacobian(M, B, K, G, L, params: Sequence[Param] = (),
                    terminal: Optional[np.ndarray] = None):
    """Adjoint inputs of one system: ``(M', B', K', G', L', f', lambda_T, Psi')``.
    ``terminal`` is dL/dx(T) (length n); accumulators start at 0."""
    m = np.asarray(M.diagonal(), dtype=np.float64)
    n = m.size
    D = -1.0 / m
    npar = len(params)
    na = n + npar
    # A' = A^T from the scaled B and K entries (duplicates are summed later)
    rows, cols, vals = [], [], []
    for mat in (B, K):
        c = coo_matrix(mat)
        rows.append(c.col.astype(np.int64))
        cols.append(c.row.astype(np.int64))
        vals.append(D[c.row] * np.asarray(c.data, dtype=np.float64))
    Kp = coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                    shape=(na, na))
    Bp = coo_matrix((na, na))
    gi, gv = _entries(G, 3)
    li, lv = _entries(L, 4)
    G_idx, G_val, L_idx, L_val, P_idx, P_val = [], [], [], [], [], []
    for (o, j, k), v in zip(gi, gv):
        s = D[o] * v
        G_idx += [(j, o, k), (k, o, j)]
        G_val += [s, s]
    for (o, j, k, l), v in zip(li, lv):
        s = D[o] * v
        L_idx += [(j, o, k, l), (k, o, j, l), (l, o, j, k)]
        L_val += [s, s, s]
    for p, par in enumerate(params):
        g = n + p
        kind = par[0]
        if kind in ("K", "B"):
            r, c = int(par[1]), int(par[2])
            G_idx.append((g, r, c))
            G_val.append(D[r])
        elif kind == "Gamma":
            o, j, k = (int(x) for x in gi[int(par[1])])
            L_idx.append((g, o, j, k))
            L_val.append(D[o])
        elif kind == "Lambda":
            o, j, k, l = (int(x) for x in li[int(par[1])])
            P_idx.append((g, o, j, k, l))
            P_val.append(D[o])
        else:
            raise ValueError(f"unknown parameter kind {kind!r}")
    lamT = np.zeros(na)
    if terminal is not None:
        lamT[:n] = np.asarray(terminal, dtype=np.float64)
    Mp = dia_matrix((-np.ones(na), [0]), shape=(na, na))
    return (Mp, Bp, Kp, _tensor([na, na, n], G_idx, G_val),
            _tensor([na, na, n, n], L_idx, L_val), np.zeros(na), lamT,
            _tensor([na, na, n, n, n], P_idx, P_val))


def jacobian_inputs(vecM, vecB, vecK, vecGamma, vecLambda,
                    params: Sequence[Sequence[Param]], terminals) -> JacobianInputs:
    """Adjoint inputs for a list of systems (``params[s]``, ``terminals[s]``)."""
    out = [system_jacobian(vecM[s], vecB[s], vecK[s], vecGamma[s], vecLambda[s], params[s],
                           terminals[s]) for s in range(len(vecM))]
    cols = [list(c) for c in zip(*out)]
    sizes = np.array([m.shape[0] for m in cols[0]], dtype=np.int64)
    return JacobianInputs(*cols, sizes=sizes)


def energy_gradient_inputs(w, params_of_system=None) -> JacobianInputs:
    """Gradient of the energy probe E(T) (the last DOF of each system) with
    respect to the K diagonal of the displacement rows and every Lambda
    coefficient -- the differentiation benchmark.  ``w`` is a workload of
    chain systems ``[x(d), v(d), E]``."""
    params, terms = [], []
    for s in range(w.n_systems):
        n = w.vecM[s].shape[0]
        d = (n - 1) // 2
        if params_of_system is not None:
            ps = params_of_system[s]
        else:
            ps = [("K", d + i, i) for i in range(d)]
            nl = len(np.asarray(w.vecLambda[s].val).reshape(-1))
            ps += [("Lambda", t) for t in range(nl)]
        params.append(ps)
        term = np.zeros(n)
        term[-1] = 1.0
        terms.append(term)
    return jacobian_inputs(w.vecM, w.vecB, w.vecK, w.vecGamma, w.vecLambda, params, terms)
