"""Pins the oracle's adjoint recursion (src/__GpuDriver.cu:715-783,
src/kernels.cu:162-195) and its checkpointed schedule:

* linear closed form: with an adjoint that does not depend on the forward
  state, every step multiplies by T(hA') = I + hA' + (hA')^2/2 + (hA')^3/6 +
  (hA')^4/24, so lambda_0 = T(hA')^T lambda_T;
* a pure-Python loop transcription with the full stored trajectory agrees
  bitwise with the NumPy oracle;
* the checkpoint interval does not change the result (bitwise);
* on a forced Duffing oscillator with a hand-derived adjoint (state
  [x, v, E], gradient of E(T) w.r.t. k and x0), the gradient matches central
  finite differences of the forward oracle to the scheme's discretisation
  order (relative 2e-4 at dt = 1e-3), and the error falls ~4x when dt halves.

The exact discrete adjoint (``discrete_adjoint_gradient``) is pinned by the
same linear closed form, by central differences of the forward oracle that
it matches to finite-difference precision (1e-9 and below, where the
reference recursion keeps its O(dt^2) gap), by checkpoint invariance, and on
a forced chain whose adjoint system comes from ``pasif_amd.adjoint`` (K_ii and
Lambda_i gradients against central differences).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from tests.helpers import random_system, tensor


def _T(Z):
    I = np.eye(Z.shape[0])
    return I + Z + Z @ Z / 2 + Z @ Z @ Z / 6 + Z @ Z @ Z @ Z / 24


def _fwd_linear(rng, n):
    m = np.ones(n)
    M = sp.dia_matrix((m, [0]), shape=(n, n))
    B = sp.coo_matrix(rng.normal(0, 0.3, (n, n)))
    K = sp.coo_matrix((n, n))
    return O.assemble_forward([M], [B], [K], [tensor([n] * 3, [], [])], [tensor([n] * 4, [], [])],
                              [np.zeros(n)], [rng.normal(0, 1, n)])


def _adj_from_arrays(Ap, n_fwd, G=None, L=None, P=None, lamT=None):
    na = Ap.shape[0]
    e3 = (np.zeros((0, 3), dtype=np.int64), np.zeros(0))
    e4 = (np.zeros((0, 4), dtype=np.int64), np.zeros(0))
    e5 = (np.zeros((0, 5), dtype=np.int64), np.zeros(0))
    # M' = -I -> D' = +1: values pass unscaled
    return O.assemble_from_arrays(-np.ones(na), sp.coo_matrix(Ap), sp.coo_matrix((na, na)),
                                  [G or e3, L or e4, P or e5], np.zeros(na),
                                  lamT if lamT is not None else np.zeros(na),
                                  scale_force=False, n_set=n_fwd)


def test_linear_adjoint_closed_form():
    rng = np.random.default_rng(10)
    fwd = _fwd_linear(rng, 3)
    Ap = rng.normal(0, 0.5, (4, 4))
    lamT = rng.normal(0, 1, 4)
    adj = _adj_from_arrays(Ap, 3, lamT=lamT)
    sr, T = 50, 23
    lay = O.cartesian_layout(3, 1, T)
    lam0 = O.adjoint_gradient(fwd, adj, None, sr, lay, np.full((1, 4), T), checkpoint_every=5)
    want = np.linalg.matrix_power(_T(Ap / sr), T) @ lamT
    np.testing.assert_allclose(lam0[0], want, rtol=1e-12)


def _loop_adjoint(fwd, adj, sr, T, lamT):
    """Pure-Python transcription: store x_0..x_T, then run the recursion."""
    from tests.helpers import loop_rhs
    h, h2, h6 = O.step_constants(sr)
    res = O.rk4_forward(fwd, None, sr, O.cartesian_layout(fwd.n, 1, T), keep_states=range(T + 1))
    xs = [list(res.checkpoints[i][0]) for i in range(T + 1)]
    lam = list(lamT)
    for n in range(T, 0, -1):
        xn, xp = xs[n], xs[n - 1]
        xm = [0.5 * (a + b) for a, b in zip(xn, xp)]
        m1 = loop_rhs(adj, lam, xn, None)
        y = [l + h2 * k for l, k in zip(lam, m1)]
        m2 = loop_rhs(adj, y, xm, None)
        y = [l + h2 * k for l, k in zip(lam, m2)]
        m3 = loop_rhs(adj, y, xm, None)
        y = [l + h * k for l, k in zip(lam, m3)]
        m4 = loop_rhs(adj, y, xp, None)
        lam = [l + h6 * (((a + 2.0 * b) + 2.0 * c) + d) for l, a, b, c, d in zip(lam, m1, m2, m3, m4)]
    return np.array(lam)


def test_nonlinear_adjoint_matches_loop_and_checkpoint_invariance():
    rng = np.random.default_rng(11)
    n_f, n_a = 4, 6
    M, B, K, G, L, _, f, q0 = random_system(rng, n_f, nG=4, nL=3, scale=0.2)
    fwd = O.assemble_forward([M], [B], [K], [G], [L], [np.zeros(n_f)], [q0])
    Ma, Ba, Ka, Ga, La, Pa, fa, la = random_system(rng, n_a, nG=6, nL=4, nP=3, n_set=n_f, scale=0.2)
    adj = O.assemble_adjoint([Ma], [Ba], [Ka], [Ga], [La], [fa], [la], [Pa], n_fwd=n_f)
    sr, T = 40, 37
    lay = O.cartesian_layout(n_f, 1, T)
    st = np.full((1, n_a), T)
    ref = _loop_adjoint(fwd, adj, sr, T, adj.q0)
    outs = [O.adjoint_gradient(fwd, adj, None, sr, lay, st, checkpoint_every=k)[0]
            for k in (1, 5, 16, 37, 100)]
    np.testing.assert_allclose(outs[0], ref, rtol=1e-13, atol=1e-15)
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


def duffing_forward(k, x0=0.0, b=0.1, lam=0.5, amp=1.0, w=1.3, sr=1000, T=2000):
    n = 3
    M = sp.dia_matrix((np.ones(n), [0]), shape=(n, n))
    B = sp.coo_matrix(([-1.0, b], ([0, 1], [1, 1])), shape=(n, n))
    K = sp.coo_matrix(([k], ([1], [0])), shape=(n, n))
    G = tensor([n] * 3, [-b], [2, 1, 1])
    L = tensor([n] * 4, [lam], [1, 0, 0, 0])
    fwd = O.assemble_forward([M], [B], [K], [G], [L], [np.array([0.0, amp, 0.0])],
                             [np.array([x0, 0.0, 0.0])])
    F = O.Forcing(np.sin(w * np.arange(T + 1) / sr)[None, :])
    return fwd, F


def duffing_adjoint_hand(k, b=0.1, lam=0.5):
    """Hand-derived adjoint of [x, v, E] with one accumulator g_k:
    lam_x' = lam_v (-k - 3 lam x^2);  lam_v' = lam_x - b lam_v + 2 b v lam_E;
    lam_E' = 0;  g_k' = -x lam_v.  Terminal lam = e_E (objective E(T))."""
    na, nf = 4, 3
    Ap = np.zeros((na, na))
    Ap[0, 1] = -k          # lam_x <- lam_v * (-k)
    Ap[1, 0] = 1.0         # lam_v <- lam_x
    Ap[1, 1] = -b
    G = (np.array([[1, 2, 1],      # lam_v <- 2 b v lam_E
                   [3, 1, 0]]),    # g_k   <- -x lam_v
         np.array([2 * b, -1.0]))
    L = (np.array([[0, 1, 0, 0]]), np.array([-3.0 * lam]))   # lam_x <- -3 lam x^2 lam_v
    lamT = np.array([0.0, 0.0, 1.0, 0.0])
    return _adj_from_arrays(Ap, nf, G=G, L=L, lamT=lamT)


@pytest.mark.parametrize("sr", [1000])
def test_duffing_gradient_against_finite_differences(sr):
    k, T = 1.2, 2 * sr
    fwd, F = duffing_forward(k, sr=sr, T=T)
    adj = duffing_adjoint_hand(k)
    lay = O.cartesian_layout(3, 1, T)
    lam0 = O.adjoint_gradient(fwd, adj, F, sr, lay, np.full((1, 4), T), checkpoint_every=64)[0]

    def E(kk, x0=0.0):
        f2, F2 = duffing_forward(kk, x0=x0, sr=sr, T=T)
        return O.rk4_forward(f2, F2, sr, lay).q[0, 2]

    hk = 1e-6
    dk = (E(k + hk) - E(k - hk)) / (2 * hk)
    dx0 = (E(k, 1e-6) - E(k, -1e-6)) / 2e-6
    assert abs(lam0[3] - dk) / abs(dk) < 2e-4
    assert abs(lam0[0] - dx0) / abs(dx0) < 2e-4


def test_duffing_gradient_error_is_second_order():
    errs = []
    for sr in (250, 500):
        k, T = 1.2, 2 * sr
        fwd, F = duffing_forward(k, sr=sr, T=T)
        adj = duffing_adjoint_hand(k)
        lay = O.cartesian_layout(3, 1, T)
        lam0 = O.adjoint_gradient(fwd, adj, F, sr, lay, np.full((1, 4), T), checkpoint_every=T)[0]

        def E(kk):
            f2, F2 = duffing_forward(kk, sr=sr, T=T)
            return O.rk4_forward(f2, F2, sr, lay).q[0, 2]
        dk = (E(k + 1e-6) - E(k - 1e-6)) / 2e-6
        errs.append(abs(lam0[3] - dk))
    rate = np.log2(errs[0] / errs[1])
    assert 1.5 < rate < 2.6, (errs, rate)


# ------------------------------------------------------------------ exact mode
def test_exact_discrete_adjoint_linear_closed_form():
    # for a state-independent adjoint the exact recursion is T(hA')^T too
    rng = np.random.default_rng(12)
    fwd = _fwd_linear(rng, 3)
    Ap = rng.normal(0, 0.5, (4, 4))
    lamT = rng.normal(0, 1, 4)
    adj = _adj_from_arrays(Ap, 3, lamT=lamT)
    sr, T = 50, 23
    lay = O.cartesian_layout(3, 1, T)
    lam0 = O.discrete_adjoint_gradient(fwd, adj, None, sr, lay, np.full((1, 4), T),
                                       checkpoint_every=5)
    want = np.linalg.matrix_power(_T(Ap / sr), T) @ lamT
    np.testing.assert_allclose(lam0[0], want, rtol=1e-12)


@pytest.mark.parametrize("sr", [250, 1000])
def test_exact_discrete_adjoint_is_the_derivative_of_the_computed_map(sr):
    # Duffing with forcing: central differences (h = 1e-5) of the forward
    # oracle agree with the exact mode to finite-difference precision
    # (~1e-11), while the reference recursion keeps its O(h^2) gap (9e-7 at
    # 250 Hz, 6e-8 at 1 kHz); the exact mode does not depend on the
    # checkpoint interval
    k, T = 1.2, 2 * sr
    fwd, F = duffing_forward(k, sr=sr, T=T)
    adj = duffing_adjoint_hand(k)
    lay = O.cartesian_layout(3, 1, T)
    st = np.full((1, 4), T)
    ex = [O.discrete_adjoint_gradient(fwd, adj, F, sr, lay, st, checkpoint_every=c)[0]
          for c in (1, 64, 300, T)]
    for e in ex[1:]:
        np.testing.assert_array_equal(e, ex[0])
    ref = O.adjoint_gradient(fwd, adj, F, sr, lay, st, checkpoint_every=64)[0]

    def E(kk, x0=0.0):
        f2, F2 = duffing_forward(kk, x0=x0, sr=sr, T=T)
        return O.rk4_forward(f2, F2, sr, lay).q[0, 2]

    h = 1e-5
    dk = (E(k + h) - E(k - h)) / (2 * h)
    dx0 = (E(k, h) - E(k, -h)) / (2 * h)
    assert abs(ex[0][3] - dk) / abs(dk) < 1e-9
    assert abs(ex[0][0] - dx0) / abs(dx0) < 1e-9
    assert abs(ref[3] - dk) / abs(dk) > 30 * abs(ex[0][3] - dk) / abs(dk)


def test_exact_mode_with_assembled_chain_jacobian_matches_fd():
    # the adjoint assembler's system is the transposed Jacobian: with it the
    # exact mode's K_ii and Lambda_i gradients match central differences of
    # the forward oracle on a forced 3-DOF chain
    from pasif_amd import adjoint as ADJ
    from pasif_amd import workloads as W
    w = W.chains(n_systems=1, d=3, steps=400, steps_per_sample=10)
    J = ADJ.energy_gradient_inputs(w)
    fwd = O.assemble_forward(w.vecM, w.vecB, w.vecK, w.vecGamma, w.vecLambda, w.vecForce,
                             w.vecQ0)
    adj = O.assemble_adjoint(*J.as_args(), n_fwd=fwd.n)
    F = O.Forcing(w.excitations, w.interp)
    lay = O.mapped_layout(fwd.sizes, [0], [w.numsteps])
    g = O.discrete_adjoint_gradient(fwd, adj, F, w.sample_rate, lay,
                                    np.full((1, adj.n), w.numsteps), checkpoint_every=32)[0]
    d = 3

    def E(dK=None, dL=None):
        K = w.vecK[0].copy().tocoo()
        L = ADJ.coo_tensor(list(w.vecLambda[0].dimensions))
        L.val = np.array(w.vecLambda[0].val, dtype=float).copy()
        L.indices = np.array(w.vecLambda[0].indices).copy()
        if dK is not None:
            i, eps = dK
            msk = (K.row == d + i) & (K.col == i)
            K.data = K.data.copy()
            K.data[msk] += eps
        if dL is not None:
            t, eps = dL
            L.val[t] += eps
        f2 = O.assemble_forward(w.vecM, w.vecB, [K], w.vecGamma, [L], w.vecForce, w.vecQ0)
        return O.rk4_forward(f2, F, w.sample_rate, lay).q[0, 2 * d]

    h = 1e-6
    for i in range(d):
        fd = (E(dK=(i, h)) - E(dK=(i, -h))) / (2 * h)
        assert abs(g[2 * d + 1 + i] - fd) <= 1e-6 * abs(fd) + 1e-12, (i, g[2 * d + 1 + i], fd)
    for t in range(d):
        fd = (E(dL=(t, h)) - E(dL=(t, -h))) / (2 * h)
        assert abs(g[3 * d + 1 + t] - fd) <= 1e-6 * abs(fd) + 1e-12, (t, g[3 * d + 1 + t], fd)
