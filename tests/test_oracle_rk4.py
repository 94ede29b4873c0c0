"""Pins the oracle's RHS, forcing rules and RK4 step:

* RHS against a dense einsum brute force and a pure-Python loop in the
  documented summation order (bitwise);
* the utest known-answer cases (src/utest.cu:220-296);
* forcing sample-index tables, guards included (src/kernels.cu:71-126);
* RK4 of a linear unforced system against the matrix polynomial
  T(hA) = I + hA + (hA)^2/2 + (hA)^3/6 + (hA)^4/24;
* a forced damped oscillator converging to its analytic solution at order 1
  (zero-order hold) and order 2 (linear interpolation);
* trajectory slot / window semantics, Cartesian-product and per-system
  step-count equivalences.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from tests.helpers import dense_tensors, loop_rhs, random_system


def _asm(rng, n, **kw):
    M, B, K, G, L, Ps, f, q0 = random_system(rng, n, **kw)
    return O.assemble_forward([M], [B], [K], [G], [L], [f], [q0])


def test_rhs_matches_dense_einsum_and_loop():
    rng = np.random.default_rng(0)
    asm = _asm(rng, 6, nG=9, nL=7)
    S = rng.normal(0, 1, asm.n)
    u = rng.normal(0, 1, asm.n)
    got = O.rhs(asm, S, None, u)
    G, L, _ = dense_tensors(asm)
    want = (asm.A.toarray() @ S + np.einsum("ojk,j,k->o", G, S, S)
            + np.einsum("ojkl,j,k,l->o", L, S, S, S) + asm.force * u)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
    np.testing.assert_array_equal(got, loop_rhs(asm, list(S), list(S), list(u)))


def test_rhs_adjoint_source_setpoint_split():
    # X_source / X_setpoint split of SpTdV (src/kernels.cu:52-61): first
    # factor from the integrated state, the others from the set point.
    rng = np.random.default_rng(1)
    n_adj, n_fwd = 5, 3
    M, B, K, G, L, Ps, f, q0 = random_system(rng, n_adj, nG=6, nL=5, nP=4, n_set=n_fwd)
    asm = O.assemble_adjoint([M], [B], [K], [G], [L], [f], [q0], [Ps], n_fwd=n_fwd)
    lam = rng.normal(0, 1, n_adj)
    x = rng.normal(0, 1, n_fwd)
    got = O.rhs(asm, lam, x, None)
    Gd, Ld, Pd = dense_tensors(asm)
    want = (asm.A.toarray() @ lam + np.einsum("ojk,j,k->o", Gd, lam, x)
            + np.einsum("ojkl,j,k,l->o", Ld, lam, x, x)
            + np.einsum("ojklm,j,k,l,m->o", Pd, lam, x, x, x))
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
    np.testing.assert_array_equal(got, loop_rhs(asm, list(lam), list(x), None))


def test_utest_known_answers():
    # src/utest.cu:220-253 and 255-296
    q = np.array([1., 1., 1.])
    rk = np.array([1., 2., 3.])
    np.testing.assert_array_equal(O.update_slope(q, rk, 2.0), [3., 5., 7.])
    np.testing.assert_array_equal(O.integrate(q, rk, rk, rk, rk, 2.0), [13., 25., 37.])


def test_zoh_index_table():
    L = 5
    # step 0: s = 0 is guarded off for stages 1-3; stage 4 reads e[1]
    assert [O.stage_sample_index_zoh(0, s, L) for s in range(4)] == [None, None, None, 1]
    assert [O.stage_sample_index_zoh(2, s, L) for s in range(4)] == [2, 2, 2, 3]
    # stage 4 of the last step would read e[L]: guarded off (s < L)
    assert [O.stage_sample_index_zoh(4, s, L) for s in range(4)] == [4, 4, 4, None]
    assert [O.stage_sample_index_zoh(9, s, L) for s in range(4)] == [None] * 4


def test_interp_index_table():
    L = 6
    # W = 4: hs = 2(step+off)+half; i = hs>>2, j = hs&3; used when 0 < i < L
    tab = {st: [O.stage_sample_index_interp(st, s, L, 4) for s in range(4)] for st in range(4)}
    assert tab[0] == [None, None, None, None]          # hs 0,1,1,2 -> i = 0
    assert tab[1] == [None, None, None, (1, 0)]        # hs 2,3,3,4
    assert tab[2] == [(1, 0), (1, 1), (1, 1), (1, 2)]  # hs 4,5,5,6
    assert tab[3] == [(1, 2), (1, 3), (1, 3), (2, 0)]
    # last usable sample index is L-1
    assert O.stage_sample_index_interp(11, 3, L, 4) is None     # hs 24 -> i 6
    assert O.stage_sample_index_interp(11, 0, L, 4) == (5, 2)
    # a non-power-of-two window (W = 20: 10 steps per sample) divides
    assert O.stage_sample_index_interp(25, 1, 100, 20) == (2, 11)


def test_interp_values_and_end_padding():
    exc = np.array([[0., 1., 2., 3.]])
    Mi = np.array([[0., .25, .5, .75], [1., .75, .5, .25]])   # linear, W = 4
    F = O.Forcing(exc, Mi)
    u = O.excitation_stage_values(F, 2, np.float64)   # hs 4,5,5,6 -> i=1
    np.testing.assert_allclose(u[:, 0], [1.0, 1.25, 1.25, 1.5])
    u = O.excitation_stage_values(F, 7, np.float64)   # hs 14,15,15,16 -> i=3,3,3,4
    np.testing.assert_allclose(u[:, 0], [3 * .5, 3 * .25, 3 * .25, 0.0])


def _linear_system(rng, n):
    m = rng.uniform(0.5, 2.0, n)
    B = sp.coo_matrix(rng.normal(0, 0.5, (n, n)) * (rng.uniform(0, 1, (n, n)) < 0.6))
    K = sp.coo_matrix(np.diag(rng.uniform(0.5, 1.5, n)))
    M = sp.dia_matrix((m, [0]), shape=(n, n))
    from tests.helpers import tensor
    G = tensor([n] * 3, [], [])
    L = tensor([n] * 4, [], [])
    q0 = rng.normal(0, 1, n)
    return O.assemble_forward([M], [B], [K], [G], [L], [np.zeros(n)], [q0])


def _T(hA):
    n = hA.shape[0]
    I = np.eye(n)
    return I + hA + hA @ hA / 2 + hA @ hA @ hA / 6 + hA @ hA @ hA @ hA / 24


def test_linear_step_is_matrix_polynomial():
    rng = np.random.default_rng(2)
    asm = _linear_system(rng, 5)
    sr, T = 100, 37
    lay = O.cartesian_layout(asm.n, 1, T)
    res = O.rk4_forward(asm, None, sr, lay, save_steps=1, window=5)
    Tm = _T(asm.A.toarray() / sr)
    states = [asm.q0]
    for _ in range(T):
        states.append(Tm @ states[-1])
    np.testing.assert_allclose(res.q[0], states[T], rtol=1e-12)
    # slot j holds the post-step state q_{j+1} (src/__GpuDriver.cu:669-687)
    for j in range(T):
        np.testing.assert_allclose(res.traj[j, 0], states[j + 1], rtol=1e-12, atol=1e-14)
    # window max over q_m, m in (T-5, T]
    want = np.max(np.abs(np.array(states[T - 4:T + 1])), axis=0)
    np.testing.assert_allclose(res.maxabs[0], want, rtol=1e-12)


def test_trajectory_slots_with_stride():
    rng = np.random.default_rng(4)
    asm = _linear_system(rng, 3)
    T, s = 20, 3
    res = O.rk4_forward(asm, None, 50, O.cartesian_layout(asm.n, 1, T), save_steps=s)
    assert res.traj.shape[0] == T // s          # PASIf.py:412
    full = O.rk4_forward(asm, None, 50, O.cartesian_layout(asm.n, 1, T), save_steps=1)
    for j in range(T // s):
        np.testing.assert_array_equal(res.traj[j], full.traj[j * s])


def _oscillator(wn=2.0, zeta=0.05):
    # x' = v ; v' = -wn^2 x - 2 zeta wn v + u ; user form M=I, B, K, f = e_v
    n = 2
    M = sp.dia_matrix(([1., 1.], [0]), shape=(n, n))
    B = sp.coo_matrix(([-1.0, 2 * zeta * wn], ([0, 1], [1, 1])), shape=(n, n))
    K = sp.coo_matrix(([wn ** 2], ([1], [0])), shape=(n, n))
    from tests.helpers import tensor
    return O.assemble_forward([M], [B], [K], [tensor([2] * 3, [], [])],
                              [tensor([2] * 4, [], [])], [np.array([0., 1.])], [np.zeros(2)])


def _analytic(t, wn, zeta, Om):
    # x'' + 2 zeta wn x' + wn^2 x = sin(Om t), x(0)=x'(0)=0
    from scipy.integrate import solve_ivp
    f = lambda tt, y: [y[1], -wn ** 2 * y[0] - 2 * zeta * wn * y[1] + np.sin(Om * tt)]
    sol = solve_ivp(f, (0, t), [0.0, 0.0], rtol=1e-13, atol=1e-14, method="DOP853")
    return sol.y[:, -1]


@pytest.mark.parametrize("mode,order", [("zoh", 1.0), ("interp", 2.0)])
def test_forced_oscillator_convergence_order(mode, order):
    wn, zeta, Om, tend = 2.0, 0.05, 1.3, 2.0
    asm = _oscillator(wn, zeta)
    exact = _analytic(tend, wn, zeta, Om)
    errs = []
    for sr in (200, 400, 800):
        T = int(tend * sr)
        if mode == "zoh":
            e = np.sin(Om * np.arange(T + 1) / sr)[None, :]
            F = O.Forcing(e)
        else:
            e = np.sin(Om * np.arange(T + 2) / sr)[None, :]
            Mi = np.array([[0.0, 0.5], [1.0, 0.5]])      # W = 2: one sample per step
            F = O.Forcing(e, Mi)
        res = O.rk4_forward(asm, F, sr, O.cartesian_layout(asm.n, 1, T))
        errs.append(np.abs(res.q[0] - exact).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(np.abs(rates - order) < 0.25), (errs, rates)


def test_cartesian_product_equals_individual_runs():
    rng = np.random.default_rng(5)
    M, B, K, G, L, _, f, q0 = random_system(rng, 4, nG=3, nL=2)
    f[:] = [0.0, 1.0, 0.0, 0.5]
    asm = O.assemble_forward([M, M], [B, B], [K, K], [G, G], [L, L], [f, f], [q0, q0])
    exc = rng.normal(0, 1, (3, 40))
    T = 30
    full = O.rk4_forward(asm, O.Forcing(exc), 100, O.cartesian_layout(asm.n, 3, T))
    for p in range(3):
        one = O.rk4_forward(asm, O.Forcing(exc[p:p + 1]), 100, O.cartesian_layout(asm.n, 1, T))
        np.testing.assert_array_equal(full.q[p], one.q[0])


def test_mapped_layout_per_system_steps_and_excitation():
    rng = np.random.default_rng(6)
    M, B, K, G, L, _, f, q0 = random_system(rng, 3, nG=3, nL=2)
    f[:] = [0.0, 1.0, 0.0]
    asm = O.assemble_forward([M, M], [B, B], [K, K], [G, G], [L, L], [f, f], [q0, q0])
    exc = rng.normal(0, 1, (2, 60))
    lay = O.mapped_layout([3, 3], [1, 0], [25, 40])
    res = O.rk4_forward(asm, O.Forcing(exc), 100, lay, window=4)
    one = O.assemble_forward([M], [B], [K], [G], [L], [f], [q0])
    a = O.rk4_forward(one, O.Forcing(exc[1:2]), 100, O.cartesian_layout(3, 1, 25), window=4)
    b = O.rk4_forward(one, O.Forcing(exc[0:1]), 100, O.cartesian_layout(3, 1, 40), window=4)
    np.testing.assert_array_equal(res.q[0], np.concatenate([a.q[0], b.q[0]]))
    np.testing.assert_array_equal(res.maxabs[0], np.concatenate([a.maxabs[0], b.maxabs[0]]))
