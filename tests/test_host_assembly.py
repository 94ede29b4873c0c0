"""Host assembly fast paths: the (row, col, value) arrays and diagonals read
straight from sparse storage are exactly what scipy's own conversions give
(same entries, same order, explicit zeros and duplicates kept), and the
vectorised adjoint assembler lists the same entries as the term-by-term
construction it replaces."""
import numpy as np
import pytest
from scipy.sparse import coo_matrix, csc_matrix, csr_matrix, dia_matrix

from pasif_amd import adjoint as ADJ
from pasif_amd.PASIf import _coo_parts, _diagonal, coo_tensor


def _random_coo(rng, n, nnz, dup=True):
    r = rng.integers(0, n, nnz)
    c = rng.integers(0, n, nnz)
    if dup:
        r[-1], c[-1] = r[0], c[0]   # a duplicate entry
    v = rng.normal(size=nnz)
    v[1] = 0.0                      # an explicit zero
    return coo_matrix((v, (r, c)), shape=(n, n))


@pytest.mark.parametrize("fmt", ["coo", "csr", "csc", "dense", "dia"])
def test_coo_parts_match_scipy(fmt):
    rng = np.random.default_rng(5)
    m = _random_coo(rng, 9, 20)
    if fmt == "csr":
        m = csr_matrix(m)
    elif fmt == "csc":
        m = csc_matrix(m)
    elif fmt == "dense":
        m = m.toarray()
    elif fmt == "dia":
        m = dia_matrix((rng.normal(size=(2, 9)), [0, 2]), shape=(9, 9))
    ref = coo_matrix(m)
    r, c, v = _coo_parts(m)
    assert np.array_equal(r, ref.row) and np.array_equal(c, ref.col)
    assert np.array_equal(v, ref.data)


def test_diagonal_fast_path():
    rng = np.random.default_rng(6)
    d = rng.normal(size=7)
    for m in (dia_matrix((d, [0]), shape=(7, 7)),
              dia_matrix((np.stack([d, d]), [1, 0]), shape=(7, 7)),
              dia_matrix((d[None, :5], [0]), shape=(7, 7)),
              csr_matrix(np.diag(d)), np.diag(d)):
        ref = m.diagonal() if hasattr(m, "diagonal") else np.diagonal(m)
        assert np.array_equal(_diagonal(m), np.asarray(ref, dtype=np.float64))


def _term_by_term(G, L, D, n, params):
    """The construction the assembler vectorises, entry by entry."""
    gi, gv = ADJ._entries(G, 3)
    li, lv = ADJ._entries(L, 4)
    Gx, Gv, Lx, Lv, Px, Pv = [], [], [], [], [], []
    for (o, j, k), v in zip(gi, gv):
        Gx += [(j, o, k), (k, o, j)]
        Gv += [D[o] * v] * 2
    for (o, j, k, l), v in zip(li, lv):
        Lx += [(j, o, k, l), (k, o, j, l), (l, o, j, k)]
        Lv += [D[o] * v] * 3
    for p, par in enumerate(params):
        g = n + p
        if par[0] in ("K", "B"):
            Gx.append((g, par[1], par[2]))
            Gv.append(D[par[1]])
        elif par[0] == "Gamma":
            e = gi[par[1]]
            Lx.append((g, *e))
            Lv.append(D[e[0]])
        else:
            e = li[par[1]]
            Px.append((g, *e))
            Pv.append(D[e[0]])
    return Gx, Gv, Lx, Lv, Px, Pv


def test_adjoint_assembler_entries_in_order():
    rng = np.random.default_rng(7)
    n = 5
    m = rng.uniform(0.5, 2.0, n)
    M = dia_matrix((m, [0]), shape=(n, n))
    B = _random_coo(rng, n, 6, dup=False)
    K = csr_matrix(_random_coo(rng, n, 7))
    G = coo_tensor([n] * 3)
    G.val = rng.normal(size=4)
    G.indices = rng.integers(0, n, 12)
    L = coo_tensor([n] * 4)
    L.val = rng.normal(size=3)
    L.indices = rng.integers(0, n, 12)
    params = [("K", 3, 1), ("Lambda", 2), ("B", 0, 4), ("Gamma", 1), ("Lambda", 0)]
    out = ADJ.system_jacobian(M, B, K, G, L, params, terminal=np.arange(n, dtype=float))
    Mp, Bp, Kp, Gp, Lp, fp, lamT, Pp = out
    D = -1.0 / m
    Gx, Gv, Lx, Lv, Px, Pv = _term_by_term(G, L, D, n, params)
    for t, ix, v, order in ((Gp, Gx, Gv, 3), (Lp, Lx, Lv, 4), (Pp, Px, Pv, 5)):
        got_ix, got_v = ADJ._entries(t, order)
        assert got_ix.tolist() == [list(map(int, e)) for e in ix]
        assert np.array_equal(got_v, np.asarray(v, dtype=np.float64))
    # A' = A^T: B and K entries transposed, scaled by D of their row
    rB, cB, vB = _coo_parts(B)
    rK, cK, vK = _coo_parts(K)
    assert np.array_equal(Kp.row, np.concatenate([cB, cK]))
    assert np.array_equal(Kp.col, np.concatenate([rB, rK]))
    assert np.array_equal(Kp.data, np.concatenate([D[rB] * vB, D[rK] * vK]))
    na = n + len(params)
    assert Kp.shape == (na, na) and Bp.shape == (na, na) and Bp.nnz == 0
    assert np.array_equal(Mp.diagonal(), -np.ones(na))
    assert np.array_equal(lamT[:n], np.arange(n)) and not lamT[n:].any()
    with pytest.raises(ValueError):
        ADJ.system_jacobian(M, B, K, G, L, [("Psi", 0)])


@pytest.mark.parametrize("name", ["two_dof", "duffing", "chains", "heterogeneous"])
def test_block_diagonal_assembly_matches_scipy_and_a_plain_loop(name):
    """setSystems' concatenation (offsets added once per batch) equals
    scipy's block_diag for B and K and a system-by-system loop for the
    tensors, after the -M^-1 scaling."""
    import os
    import scipy.sparse as sp
    from pasif_amd import workloads as W
    w = {"two_dof": lambda: W.two_dof(50, 20), "duffing": lambda: W.duffing(16, 20),
         "chains": lambda: W.chains(12, 5, 40),
         "heterogeneous": lambda: W.heterogeneous(60, 30, 300, steps_per_sample=5,
                                                  n_signals=3)}[name]()
    os.environ["LOCAL_RANK"] = "-1"
    try:
        p = W.make_solver(w)
    finally:
        del os.environ["LOCAL_RANK"]
    D = -1.0 / np.concatenate([np.asarray(m.diagonal()) for m in w.vecM])
    for got, mats in ((p.system_B, w.vecB), (p.system_K, w.vecK)):
        ref = sp.diags(D) @ sp.block_diag([sp.coo_matrix(m) for m in mats])
        assert np.array_equal(got.toarray(), ref.toarray())
    for got, ts in ((p.system_Gamma, w.vecGamma), (p.system_Lambda, w.vecLambda)):
        nd = len(ts[0].dimensions)
        off = np.zeros(nd, dtype=np.int64)
        idx, val = [], []
        for t in ts:
            v = np.asarray(t.val, dtype=np.float64).reshape(-1)
            ix = np.asarray(t.indices, dtype=np.int64).reshape(-1, nd)
            idx.append(ix + off)
            val.append(v)
            off = off + np.asarray(t.dimensions, dtype=np.int64)
        idx, val = np.concatenate(idx), np.concatenate(val)
        assert np.array_equal(np.asarray(got.indices).reshape(-1, nd), idx)
        assert np.array_equal(np.asarray(got.val), val * D[idx[:, 0]])
