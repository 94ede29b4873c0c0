"""CPU checks of the adjoint-system assembler (pasif_amd.adjoint):

* on the Duffing oscillator it reproduces the hand-derived adjoint (dense
  comparison after the reference's preprocessing);
* on random systems, the oracle gradient of an assembled adjoint matches
  central finite differences of the oracle forward for every parameter kind
  (K, B, Gamma, Lambda entries and the initial state) within the scheme's
  discretisation error.
"""
import numpy as np
import scipy.sparse as sp

import oracle as O
from pasif_amd import adjoint as ADJ
from pasif_amd import coo_tensor
from tests.helpers import dense_tensors
from tests.test_oracle_adjoint import duffing_adjoint_hand


def _coo(dims, val, idx):
    t = coo_tensor(dims)
    t.val = list(val)
    t.indices = list(idx)
    return t


def _duffing_user(k=1.2, b=0.1, lam=0.5):
    n = 3
    M = sp.dia_matrix((np.ones(n), [0]), shape=(n, n))
    B = sp.coo_matrix(([-1.0, b], ([0, 1], [1, 1])), shape=(n, n))
    K = sp.coo_matrix(([k], ([1], [0])), shape=(n, n))
    G = _coo([n] * 3, [-b], [2, 1, 1])
    L = _coo([n] * 4, [lam], [1, 0, 0, 0])
    return M, B, K, G, L


def _dense_adj(asm):
    G, L, P = dense_tensors(asm)
    return asm.A.toarray(), G, L, P


def test_duffing_assembler_matches_hand_derived():
    M, B, K, G, L = _duffing_user()
    Mp, Bp, Kp, Gp, Lp, fp, lamT, Pp = ADJ.system_jacobian(M, B, K, G, L, params=[("K", 1, 0)],
                                                           terminal=np.array([0, 0, 1.0]))
    got = O.assemble_adjoint([Mp], [Bp], [Kp], [Gp], [Lp], [fp], [lamT], [Pp], n_fwd=3)
    want = duffing_adjoint_hand(1.2)
    for a, b in zip(_dense_adj(got), _dense_adj(want)):
        np.testing.assert_allclose(a, b, rtol=1e-15, atol=1e-15)
    np.testing.assert_array_equal(got.q0, want.q0)


def test_random_system_gradients_match_finite_differences():
    rng = np.random.default_rng(21)
    n = 4
    m = np.array([1.0, 1.5, 1.0, 0.8])
    M = sp.dia_matrix((m, [0]), shape=(n, n))
    # x' = v ; v' = -(B v + K x) - G(x,x) - L(x,x,x) + f u  on 2 DOF
    B = sp.coo_matrix(([-1.0, -1.0, 0.1, 0.02, 0.12], ([0, 1, 2, 2, 3], [2, 3, 2, 3, 3])),
                      shape=(n, n))
    K = sp.coo_matrix(([2.0, -0.8, -0.8, 1.7], ([2, 2, 3, 3], [0, 1, 0, 1])), shape=(n, n))
    G = _coo([n] * 3, [0.3, -0.2], [2, 0, 1, 3, 1, 1])
    L = _coo([n] * 4, [0.6, 0.4], [2, 0, 0, 0, 3, 1, 1, 1])
    f = np.array([0.0, 0.0, 1.0, 0.0])
    q0 = np.array([0.1, -0.05, 0.0, 0.02])
    sr, T = 400, 600
    exc = np.sin(1.7 * np.arange(T + 1) / sr)[None, :]
    F = O.Forcing(exc)
    lay = O.cartesian_layout(n, 1, T)
    wobj = np.array([0.3, -0.7, 0.2, 0.5])   # L = w . x(T)
    params = [("K", 2, 0), ("K", 3, 1), ("B", 2, 3), ("Gamma", 0), ("Lambda", 1)]
    Mp, Bp, Kp, Gp, Lp, fp, lamT, Pp = ADJ.system_jacobian(M, B, K, G, L, params, terminal=wobj)
    adj = O.assemble_adjoint([Mp], [Bp], [Kp], [Gp], [Lp], [fp], [lamT], [Pp], n_fwd=n)

    def objective(Mx, Bx, Kx, Gx, Lx, q0x):
        fw = O.assemble_forward([Mx], [Bx], [Kx], [Gx], [Lx], [f], [q0x])
        return float(wobj @ O.rk4_forward(fw, F, sr, lay).q[0])

    fwd = O.assemble_forward([M], [B], [K], [G], [L], [f], [q0])
    lam0 = O.adjoint_gradient(fwd, adj, F, sr, lay, np.full((1, adj.n), T), checkpoint_every=50)[0]

    def bump(par, h):
        Bx, Kx = B.copy().tolil(), K.copy().tolil()
        Gx = _coo(G.dimensions, list(G.val), list(G.indices))
        Lx = _coo(L.dimensions, list(L.val), list(L.indices))
        if par[0] == "K":
            Kx[par[1], par[2]] += h
        elif par[0] == "B":
            Bx[par[1], par[2]] += h
        elif par[0] == "Gamma":
            Gx.val[par[1]] += h
        else:
            Lx.val[par[1]] += h
        return objective(M, sp.coo_matrix(Bx), sp.coo_matrix(Kx), Gx, Lx, q0)

    h = 1e-6
    for p, par in enumerate(params):
        fd = (bump(par, h) - bump(par, -h)) / (2 * h)
        assert abs(lam0[n + p] - fd) <= 2e-4 * abs(fd) + 1e-9, (par, lam0[n + p], fd)
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        fd = (objective(M, B, K, G, L, q0 + e) - objective(M, B, K, G, L, q0 - e)) / (2 * h)
        assert abs(lam0[i] - fd) <= 2e-4 * abs(fd) + 1e-9, (i, lam0[i], fd)
