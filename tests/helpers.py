"""Shared test helpers: tiny user-level systems and a pure-Python loop
transcription of the RK4 / adjoint recursion used to cross-check the NumPy
oracle on small cases."""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import scipy.sparse as sp


def tensor(dims, val, idx):
    return SimpleNamespace(dimensions=list(dims), val=list(val), indices=list(idx))


def random_system(rng, n, nnzA=None, nG=4, nL=3, nP=0, n_set=None, diag_range=(0.5, 2.0),
                  scale=0.3):
    """A random user-level system: diagonal M, COO B and K, tensors with
    possible duplicate entries, dense force and initial conditions."""
    n_set = n_set or n
    nnzA = nnzA or 2 * n
    m = rng.uniform(*diag_range, n)
    M = sp.dia_matrix((m, [0]), shape=(n, n))
    r = rng.integers(0, n, nnzA)
    c = rng.integers(0, n, nnzA)
    B = sp.coo_matrix((rng.normal(0, scale, nnzA), (r, c)), shape=(n, n))
    r2 = rng.integers(0, n, nnzA)
    c2 = rng.integers(0, n, nnzA)
    K = sp.coo_matrix((rng.normal(0, scale, nnzA), (r2, c2)), shape=(n, n))

    def rt(order, nnz):
        dims = [n, n] + [n_set] * (order - 2) if n_set != n else [n] * order
        if order == 3 and n_set == n:
            dims = [n] * 3
        idx = []
        for _ in range(nnz):
            e = [rng.integers(0, n), rng.integers(0, n)] + [rng.integers(0, n_set) for _ in range(order - 2)]
            idx += [int(v) for v in e]
        return tensor(dims, rng.normal(0, scale, nnz), idx)

    G = rt(3, nG)
    L = rt(4, nL)
    Ps = rt(5, nP)
    f = rng.normal(0, 1, n) * (rng.uniform(0, 1, n) < 0.5)
    q0 = rng.normal(0, 0.3, n)
    return M, B, K, G, L, Ps, f, q0


def dense_tensors(asm):
    """Dense Gamma/Lambda/Psi of an assembled system (duplicates summed)."""
    n, ns = asm.n, asm.n_set
    G = np.zeros((n, n, ns))
    np.add.at(G, tuple(asm.gamma_idx.T), asm.gamma_val)
    L = np.zeros((n, n, ns, ns))
    np.add.at(L, tuple(asm.lam_idx.T), asm.lam_val)
    P = np.zeros((n, n, ns, ns, ns))
    if len(asm.psi_val):
        np.add.at(P, tuple(asm.psi_idx.T), asm.psi_val)
    return G, L, P


def loop_rhs(asm, S, X, u_dof):
    """Pure-Python RHS in the oracle's documented order: CSR row order of A,
    then Gamma / Lambda / Psi in COO order (each summed from 0), then
    ((((A + G) + L) + Psi) + force)."""
    n = asm.n
    A = asm.A
    out = []
    gsum = [0.0] * n
    lsum = [0.0] * n
    psum = [0.0] * n
    for t in range(len(asm.gamma_val)):
        o, j, k = (int(v) for v in asm.gamma_idx[t])
        gsum[o] += float(asm.gamma_val[t]) * S[j] * X[k]
    for t in range(len(asm.lam_val)):
        o, j, k, l = (int(v) for v in asm.lam_idx[t])
        lsum[o] += float(asm.lam_val[t]) * S[j] * X[k] * X[l]
    for t in range(len(asm.psi_val)):
        o, j, k, l, m = (int(v) for v in asm.psi_idx[t])
        psum[o] += float(asm.psi_val[t]) * S[j] * X[k] * X[l] * X[m]
    for i in range(n):
        a = 0.0
        for jj in range(A.indptr[i], A.indptr[i + 1]):
            a += float(A.data[jj]) * S[int(A.indices[jj])]
        v = (((a + gsum[i]) + lsum[i]) + psum[i])
        if u_dof is not None and asm.force[i] != 0.0:
            v = v + float(asm.force[i]) * u_dof[i]
        out.append(v)
    return out
