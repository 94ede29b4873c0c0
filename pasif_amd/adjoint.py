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

from .PASIf import _coo_parts, _diagonal, coo_tensor

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


def _tensor(dims, idx, val):
    t = coo_tensor(list(dims))
    t.val = np.asarray(val, dtype=np.float64).reshape(-1)
    t.indices = (np.asarray(idx, dtype=np.int64).reshape(-1) if len(idx)
                 else np.zeros(0, dtype=np.int64))
    return t


# The adjoint mass (-I) and the empty B' depend only on the size: one shared,
# read-only matrix per size (building a scipy object per system dominates the
# assembly of large batches).
_SHARED = {}


def _minus_identity(na: int) -> dia_matrix:
    key = ("M", na)
    if key not in _SHARED:
        _SHARED[key] = dia_matrix((-np.ones(na), [0]), shape=(na, na))
    return _SHARED[key]


def _empty(na: int) -> coo_matrix:
    key = ("B", na)
    if key not in _SHARED:
        _SHARED[key] = coo_matrix((na, na))
    return _SHARED[key]


def _interleave(ix: np.ndarray, perms) -> np.ndarray:
    """Rows ``ix[e][perm]`` for every entry e and, within it, every perm in
    order: entry-major, as the term-by-term construction lists them."""
    if ix.shape[0] == 0:
        return np.zeros((0, ix.shape[1]), dtype=np.int64)
    return np.stack([ix[:, list(p)] for p in perms], axis=1).reshape(-1, ix.shape[1])


def system_jacobian(M, B, K, G, L, params: Sequence[Param] = (),
                    terminal: Optional[np.ndarray] = None):
    """Adjoint inputs of one system: ``(M', B', K', G', L', f', lambda_T, Psi')``.
    ``terminal`` is dL/dx(T) (length n); accumulators start at 0."""
    m = _diagonal(M)
    n = m.size
    D = -1.0 / m
    npar = len(params)
    na = n + npar
    # A' = A^T from the scaled B and K entries (duplicates are summed later)
    rows, cols, vals = [], [], []
    for mat in (B, K):
        r, c, v = _coo_parts(mat)
        r = np.asarray(r, dtype=np.int64)
        rows.append(np.asarray(c, dtype=np.int64))
        cols.append(r)
        vals.append(D[r] * np.asarray(v, dtype=np.float64))
    Kp = coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                    shape=(na, na))
    Bp = _empty(na)
    gi, gv = _entries(G, 3)
    li, lv = _entries(L, 4)
    # G[o, j, k] -> G'[j, o, k], G'[k, o, j]; L[o, j, k, l] -> L'[j, o, k, l],
    # L'[k, o, j, l], L'[l, o, j, k] (each entry's images consecutive)
    G_idx = [_interleave(gi, ((1, 0, 2), (2, 0, 1)))]
    G_val = [np.repeat(D[gi[:, 0]] * gv, 2)]
    L_idx = [_interleave(li, ((1, 0, 2, 3), (2, 0, 1, 3), (3, 0, 1, 2)))]
    L_val = [np.repeat(D[li[:, 0]] * lv, 3)]
    # parameters, each kind into its tensor in parameter order
    kinds = [par[0] for par in params]
    for k in kinds:
        if k not in ("K", "B", "Gamma", "Lambda"):
            raise ValueError(f"unknown parameter kind {k!r}")
    sel = [p for p, k in enumerate(kinds) if k in ("K", "B")]
    if sel:
        rc = np.array([(int(params[p][1]), int(params[p][2])) for p in sel], dtype=np.int64)
        G_idx.append(np.column_stack([n + np.asarray(sel, dtype=np.int64), rc]))
        G_val.append(D[rc[:, 0]])
    sel = [p for p, k in enumerate(kinds) if k == "Gamma"]
    if sel:
        e = gi[np.array([int(params[p][1]) for p in sel], dtype=np.int64)]
        L_idx.append(np.column_stack([n + np.asarray(sel, dtype=np.int64), e]))
        L_val.append(D[e[:, 0]])
    P_idx, P_val = [np.zeros((0, 5), dtype=np.int64)], [np.zeros(0)]
    sel = [p for p, k in enumerate(kinds) if k == "Lambda"]
    if sel:
        e = li[np.array([int(params[p][1]) for p in sel], dtype=np.int64)]
        P_idx.append(np.column_stack([n + np.asarray(sel, dtype=np.int64), e]))
        P_val.append(D[e[:, 0]])
    G_idx, G_val = np.concatenate(G_idx), np.concatenate(G_val)
    L_idx, L_val = np.concatenate(L_idx), np.concatenate(L_val)
    P_idx, P_val = np.concatenate(P_idx), np.concatenate(P_val)
    lamT = np.zeros(na)
    if terminal is not None:
        lamT[:n] = np.asarray(terminal, dtype=np.float64)
    Mp = _minus_identity(na)
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


@dataclass
class JacobianBlocks:
    """The adjoint systems of many systems already concatenated block-diagonal:
    exactly what ``PASIf.setJacobian`` unfolds from the per-system inputs of
    ``jacobian_inputs`` (same entries in the same order), built without a
    scipy object per system.  Pass it to ``PASIf.setJacobianBlocks``."""
    sizes: np.ndarray        # adjoint size of each system
    fwd_sizes: np.ndarray    # forward size of each system
    K: coo_matrix            # A' = A^T entries (the unfolded K'; B' is empty)
    gamma: coo_tensor        # [na, na, n]
    lam: coo_tensor          # [na, na, n, n]
    psi: coo_tensor          # [na, na, n, n, n]
    init: np.ndarray         # lambda(T) and zero accumulators


def jacobian_blocks(vecM, vecB, vecK, vecGamma, vecLambda,
                    params: Sequence[Sequence[Param]], terminals) -> JacobianBlocks:
    """``jacobian_inputs`` concatenated: systems with the same entry pattern
    (sizes, B / K / Gamma / Lambda indices and parameters) are computed
    together, then every system's entries are put back in system order."""
    S = len(vecM)
    raw, groups = [], {}
    for s in range(S):
        m = _diagonal(vecM[s])
        rB, cB, vB = (np.asarray(x) for x in _coo_parts(vecB[s]))
        rK, cK, vK = (np.asarray(x) for x in _coo_parts(vecK[s]))
        gi, gv = _entries(vecGamma[s], 3)
        li, lv = _entries(vecLambda[s], 4)
        ps = tuple(tuple(p) for p in params[s])
        key = (m.size, rB.astype(np.int64).tobytes(), cB.astype(np.int64).tobytes(),
               rK.astype(np.int64).tobytes(), cK.astype(np.int64).tobytes(), gi.tobytes(),
               li.tobytes(), ps)
        groups.setdefault(key, []).append(s)
        raw.append((m, vB, vK, gv, lv))
    n_f = np.array([raw[s][0].size for s in range(S)], dtype=np.int64)
    n_a = np.array([raw[s][0].size + len(params[s]) for s in range(S)], dtype=np.int64)
    off_f = np.concatenate([[0], np.cumsum(n_f)[:-1]])
    off_a = np.concatenate([[0], np.cumsum(n_a)[:-1]])
    parts = {k: [] for k in ("K", "G", "L", "P")}
    init = [None] * S
    for key, members in groups.items():
        s0 = members[0]
        sel = np.asarray(members, dtype=np.int64)
        n = int(key[0])
        npar = len(params[s0])
        rB, cB, vB = (np.asarray(x) for x in _coo_parts(vecB[s0]))
        rK, cK, vK = (np.asarray(x) for x in _coo_parts(vecK[s0]))
        gi, _ = _entries(vecGamma[s0], 3)
        li, _ = _entries(vecLambda[s0], 4)
        D = -1.0 / np.stack([raw[s][0] for s in members])                # [k, n]
        # A' = A^T: B entries, then K entries, transposed, scaled by D of the row
        rows = np.concatenate([cB, cK]).astype(np.int64)
        cols = np.concatenate([rB, rK]).astype(np.int64)
        rBi, rKi = rB.astype(np.int64), rK.astype(np.int64)
        kv = np.concatenate([D[:, rBi] * np.stack([raw[s][1] for s in members]),
                             D[:, rKi] * np.stack([raw[s][2] for s in members])], axis=1)
        parts["K"].append((sel, np.stack([rows, cols], axis=1), kv, 2))
        # tensor images of the forward terms, then the parameters' entries
        gvals = np.stack([raw[s][3] for s in members]) if gi.shape[0] else np.zeros((len(sel), 0))
        lvals = np.stack([raw[s][4] for s in members]) if li.shape[0] else np.zeros((len(sel), 0))
        G_idx = [_interleave(gi, ((1, 0, 2), (2, 0, 1)))]
        G_val = [np.repeat(D[:, gi[:, 0]] * gvals, 2, axis=1)]
        L_idx = [_interleave(li, ((1, 0, 2, 3), (2, 0, 1, 3), (3, 0, 1, 2)))]
        L_val = [np.repeat(D[:, li[:, 0]] * lvals, 3, axis=1)]
        P_idx, P_val = [np.zeros((0, 5), dtype=np.int64)], [np.zeros((len(sel), 0))]
        kinds = [p[0] for p in params[s0]]
        for k in kinds:
            if k not in ("K", "B", "Gamma", "Lambda"):
                raise ValueError(f"unknown parameter kind {k!r}")
        ps = [p for p, k in enumerate(kinds) if k in ("K", "B")]
        if ps:
            rc = np.array([(int(params[s0][p][1]), int(params[s0][p][2])) for p in ps],
                          dtype=np.int64)
            G_idx.append(np.column_stack([n + np.asarray(ps, dtype=np.int64), rc]))
            G_val.append(D[:, rc[:, 0]])
        ps = [p for p, k in enumerate(kinds) if k == "Gamma"]
        if ps:
            e = gi[np.array([int(params[s0][p][1]) for p in ps], dtype=np.int64)]
            L_idx.append(np.column_stack([n + np.asarray(ps, dtype=np.int64), e]))
            L_val.append(D[:, e[:, 0]])
        ps = [p for p, k in enumerate(kinds) if k == "Lambda"]
        if ps:
            e = li[np.array([int(params[s0][p][1]) for p in ps], dtype=np.int64)]
            P_idx.append(np.column_stack([n + np.asarray(ps, dtype=np.int64), e]))
            P_val.append(D[:, e[:, 0]])
        parts["G"].append((sel, np.concatenate(G_idx), np.concatenate(G_val, axis=1), 2))
        parts["L"].append((sel, np.concatenate(L_idx), np.concatenate(L_val, axis=1), 2))
        parts["P"].append((sel, np.concatenate(P_idx), np.concatenate(P_val, axis=1), 2))
        for s in members:
            lam_t = np.zeros(n + npar)
            if terminals[s] is not None:
                lam_t[:n] = np.asarray(terminals[s], dtype=np.float64)
            init[s] = lam_t

    def assemble(kind, order):
        """every group's entries offset into the concatenation (the first two
        axes by adjoint sizes, the rest by forward sizes), in system order"""
        sys_ids, idx, val = [], [], []
        for sel, li_, v, na_axes in parts[kind]:
            k, e = v.shape
            if e == 0:
                continue
            off = np.concatenate([np.repeat(off_a[sel][:, None], na_axes, axis=1),
                                  np.repeat(off_f[sel][:, None], order - na_axes, axis=1)], axis=1)
            idx.append((li_[None, :, :] + off[:, None, :]).reshape(-1, order))
            val.append(v.reshape(-1))
            sys_ids.append(np.repeat(sel, e))
        if not idx:
            return np.zeros((0, order), dtype=np.int64), np.zeros(0)
        sid = np.concatenate(sys_ids)
        o = np.argsort(sid, kind="stable")
        return np.concatenate(idx)[o], np.concatenate(val)[o]

    Na, Nf = int(n_a.sum()), int(n_f.sum())
    ki, kv = assemble("K", 2)
    K = coo_matrix((kv, (ki[:, 0], ki[:, 1])), shape=(Na, Na))
    gi, gv = assemble("G", 3)
    li, lv = assemble("L", 4)
    pi, pv = assemble("P", 5)
    return JacobianBlocks(sizes=n_a, fwd_sizes=n_f, K=K,
                          gamma=_tensor([Na, Na, Nf], gi, gv),
                          lam=_tensor([Na, Na, Nf, Nf], li, lv),
                          psi=_tensor([Na, Na, Nf, Nf, Nf], pi, pv),
                          init=np.concatenate(init) if S else np.zeros(0))


def energy_gradient_blocks(w, params_of_system=None) -> JacobianBlocks:
    """``energy_gradient_inputs`` as ``JacobianBlocks``."""
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
    return jacobian_blocks(w.vecM, w.vecB, w.vecK, w.vecGamma, w.vecLambda, params, terms)
