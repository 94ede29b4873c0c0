"""Pins the oracle's assembly / preprocessing (PASIf.py:202-277, 475-633)
against closed forms: ``A = -(diag(1/m)) (B + K)`` per block, tensors scaled
by ``-1/m[first index]``, force ``f/m``, block-diagonal index offsets."""
import numpy as np
import scipy.sparse as sp

import oracle as O
from tests.helpers import dense_tensors, random_system, tensor


def _example_1dof():
    # the 4-DOF probed oscillator of examples/FullInterfaceTesting_1dof.py:43-82
    n = 4
    M = sp.dia_matrix(([1., 1., 1., 1.], [0]), shape=(n, n))
    B = sp.coo_matrix(([-1., -1.], ([0, 1], [2, 3])), shape=(n, n))
    K = sp.coo_matrix(([1., 1.], ([2, 2], [0, 2])), shape=(n, n))
    G = tensor([n] * 3, [-1.], [3, 0, 0])
    L = tensor([n] * 4, [0.], [0, 0, 0, 0])
    return M, B, K, G, L, np.array([0., 0., 1., 0.]), np.zeros(n)


def _example_8dof():
    # the two-oscillator system of examples/FullInterfaceTesting.py:48-76
    n = 8
    M = sp.dia_matrix(([1., 1., 1., 1., 1., 10., 1., 1.], [0]), shape=(n, n))
    B = sp.coo_matrix(([-1, -1, -1, -1, 1, 10], ([0, 1, 2, 3, 4, 5], [4, 5, 6, 7, 4, 5])), shape=(n, n))
    K = sp.coo_matrix(([6, 10], ([4, 5], [0, 1])), shape=(n, n))
    G = tensor([n] * 3, [-10, -10, -1, -1], [4, 0, 1, 5, 0, 0, 6, 0, 0, 7, 1, 1])
    L = tensor([n] * 4, [40000], [5, 1, 1, 1])
    return M, B, K, G, L, np.array([0, 0, 0, 0, 0.5, 0, 0, 0]), np.zeros(n)


def _closed_form(systems):
    As, Gs, Ls, fs = [], [], [], []
    for (M, B, K, G, L, f, q0) in systems:
        m = M.diagonal()
        D = np.diag(-1.0 / m)
        As.append(D @ (B.toarray() + K.toarray()))
        n = len(m)
        Gd = np.zeros((n, n, n))
        for t in range(len(G.val)):
            o, j, k = G.indices[3 * t:3 * t + 3]
            Gd[o, j, k] += -G.val[t] / m[o]
        Ld = np.zeros((n, n, n, n))
        for t in range(len(L.val)):
            o, j, k, l = L.indices[4 * t:4 * t + 4]
            Ld[o, j, k, l] += -L.val[t] / m[o]
        Gs.append(Gd)
        Ls.append(Ld)
        fs.append(np.asarray(f, float) / m)
    N = sum(a.shape[0] for a in As)
    A = sp.block_diag(As).toarray()
    G = np.zeros((N, N, N))
    L = np.zeros((N, N, N, N))
    o = 0
    for Gd, Ld in zip(Gs, Ls):
        n = Gd.shape[0]
        G[o:o + n, o:o + n, o:o + n] = Gd
        L[o:o + n, o:o + n, o:o + n, o:o + n] = Ld
        o += n
    return A, G, L, np.concatenate(fs)


def _assemble(systems):
    cols = list(zip(*systems))
    return O.assemble_forward(*cols)


def test_examples_match_closed_form():
    systems = [_example_1dof(), _example_8dof(), _example_8dof(), _example_1dof()]
    asm = _assemble(systems)
    A, G, L, f = _closed_form(systems)
    assert asm.n == 24 and list(asm.sizes) == [4, 8, 8, 4]
    np.testing.assert_allclose(asm.A.toarray(), A, rtol=1e-15, atol=0)
    Gd, Ld, _ = dense_tensors(asm)
    np.testing.assert_allclose(Gd, G, rtol=1e-15, atol=0)
    np.testing.assert_allclose(Ld, L, rtol=1e-15, atol=0)
    np.testing.assert_allclose(asm.force, f, rtol=1e-15, atol=0)
    # CSR with sorted column indices (PASIf.py:250-251)
    for i in range(asm.n):
        cols = asm.A.indices[asm.A.indptr[i]:asm.A.indptr[i + 1]]
        assert np.all(np.diff(cols) > 0)


def test_random_systems_match_closed_form():
    rng = np.random.default_rng(3)
    systems = []
    for n in (1, 3, 5, 2):
        M, B, K, G, L, _, f, q0 = random_system(rng, n)
        systems.append((M, B, K, G, L, f, q0))
    asm = _assemble(systems)
    A, G, L, f = _closed_form(systems)
    np.testing.assert_allclose(asm.A.toarray(), A, rtol=1e-14, atol=1e-300)
    Gd, Ld, _ = dense_tensors(asm)
    np.testing.assert_allclose(Gd, G, rtol=1e-14, atol=1e-300)
    np.testing.assert_allclose(Ld, L, rtol=1e-14, atol=1e-300)
    np.testing.assert_allclose(asm.force, f, rtol=1e-15, atol=0)
    np.testing.assert_array_equal(asm.q0, np.concatenate([s[6] for s in systems]))


def test_concatenation_offsets_each_axis():
    # coo_tensor.concatenateTensor (PASIf.py:86-100): axis d of the second
    # tensor is offset by the first tensor's size along d -- including the
    # mixed adjoint dims [n_adj, n_adj, n_fwd].
    n_adj, n_fwd = 3, 2
    t1 = tensor([n_adj, n_adj, n_fwd], [1.0], [2, 1, 1])
    t2 = tensor([n_adj, n_adj, n_fwd], [2.0], [0, 2, 1])
    M = sp.dia_matrix((np.full(n_adj, -1.0), [0]), shape=(n_adj, n_adj))
    Z = sp.coo_matrix((n_adj, n_adj))
    e4 = tensor([n_adj, n_adj, n_fwd, n_fwd], [], [])
    e5 = tensor([n_adj, n_adj, n_fwd, n_fwd, n_fwd], [], [])
    asm = O.assemble_adjoint([M, M], [Z, Z], [Z, Z], [t1, t2], [e4, e4], [np.zeros(3)] * 2,
                             [np.zeros(3)] * 2, [e5, e5], n_fwd=2 * n_fwd)
    np.testing.assert_array_equal(asm.gamma_idx, [[2, 1, 1], [3, 5, 3]])
    # M' = -I  ->  D' = +1: values unchanged; adjoint force left unscaled
    np.testing.assert_array_equal(asm.gamma_val, [1.0, 2.0])
    assert asm.n == 6 and asm.n_set == 4


def test_force_epsilon_drop():
    # COOVector drops |v| <= eps of the working type (helpers.cu:676)
    f = np.array([1e-17, -2e-16, 3e-16, 1e-9, 0.0, -1.0])
    d = O.drop_small_forces(f, np.float64)
    np.testing.assert_array_equal(d, [0, 0, 3e-16, 1e-9, 0, -1.0])
    d32 = O.drop_small_forces(f, np.float32)
    np.testing.assert_array_equal(d32, [0, 0, 0, 0, 0, -1.0])


def test_step_constants():
    h, h2, h6 = O.step_constants(2000)
    assert h == 1.0 / 2000 and h2 == h / 2.0 and h6 == h / 6.0
    h, h2, h6 = O.step_constants(3, np.float32)
    assert h == np.float32(1.0 / 3) and h6 == np.float32(float(np.float32(1.0 / 3)) / 6.0)
