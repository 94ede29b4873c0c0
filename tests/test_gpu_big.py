"""The row-parallel path (``setKernel("big")``, and what ``"auto"`` runs for
systems too large for a tile) on the GPU against the oracle: FP64 within
1e-10 per output (``oracle.parity``) for the forward amplitudes, the
reference adjoint recursion and the exact discrete adjoint; bitwise
invariant under the checkpoint interval and the solver, with uneven step
counts, interpolated and zero-order-hold forcing; one chain of 2,001 DOF
(4,001 adjoint rows) alone and beside small chains in one solve.
"""
import numpy as np
import pytest

from oracle import parity as PAR
from pasif_amd import adjoint as ADJ
from pasif_amd import workloads as W
from tests.test_gpu_adjoint import oracle_gradient
from tests.test_gpu_exact_adjoint import oracle_exact
from tests.test_gpu_forward import oracle_run

pytestmark = pytest.mark.gpu


def _case(name):
    print(f"[gpu case] {name}", flush=True)


def _sizes(w):
    return [m.shape[0] for m in w.vecM]


def _check(g, ref, sizes, what):
    r = PAR.report(g, np.asarray(ref).reshape(-1), sizes)
    assert r.max_rel < PAR.TOL64, (what, r.as_dict())


def test_big_forward_and_adjoints_on_small_chains():
    _case("big: chains 10 x d4 x 500, uneven steps, forward / reference / exact")
    w = W.chains(n_systems=10, d=4, steps=500, steps_per_sample=10)
    w.steps_of_system = np.array([500, 430, 37, 500, 1, 260, 499, 120, 333, 64])
    J = ADJ.energy_gradient_inputs(w)
    p = W.make_solver(w)
    p.setKernel("big")
    assert p.kernelInfo()["kernel"] == "big"
    q = p.getAmplitudes()
    _check(q, oracle_run(w).q, _sizes(w), "forward")
    p.setJacobian(*J.as_args())
    ref = oracle_gradient(w, J, 64)
    outs = []
    for k in (1, 7, 64, 500):
        p.setTuning(checkpoint_every=k)
        g = p.getGradient()
        _check(g, ref, J.sizes, ("reference", k))
        outs.append(g)
    for g in outs[1:]:
        np.testing.assert_array_equal(g, outs[0])
    p.setSolver("plain")
    np.testing.assert_array_equal(p.getGradient(), outs[-1])
    assert p._lastTiming()["launches"] == 3 * 500
    p.setSolver("persistent")
    p.setAdjointMode("exact")
    assert p.kernelInfo(adjoint=True)["kernel"] == "big"
    ex = oracle_exact(w, J)
    g1 = p.getGradient()
    _check(g1, ex, J.sizes, "exact")
    p.setTuning(checkpoint_every=7)
    np.testing.assert_array_equal(p.getGradient(), g1)
    # the tile kernel on the same problem agrees to round-off
    p.setKernel("tile")
    _check(p.getGradient(), ex, J.sizes, "exact on tiles")


def test_big_zero_order_hold_cartesian_duffing():
    _case("big: duffing 8 x 700, cartesian excitations, zero-order hold")
    w = W.duffing(n_systems=8, steps=700)
    w.exc_of_system = None      # every system against every excitation
    p = W.make_solver(w)
    p.setKernel("big")
    q = p.getAmplitudes()
    _check(q, oracle_run(w).q, _sizes(w) * len(w.excitations), "duffing forward")
    with pytest.raises(Exception):
        p.getMaxAmplitudes(window=100)
    with pytest.raises(Exception):
        p.getTrajectory(saveSteps=10)


def test_big_chain_of_2001_dof_alone_and_beside_small_chains():
    _case("big: one 2001-DOF chain x 2000 steps, forward + reference + exact adjoint")
    w = W.chains(n_systems=1, d=1000, steps=2000, steps_per_sample=10)
    J = ADJ.energy_gradient_inputs(w)
    p = W.make_solver(w)
    info = p.kernelInfo()
    assert info["kernel"] == "big", info
    _check(p.getAmplitudes(), oracle_run(w).q, _sizes(w), "2001-DOF forward")
    p.setJacobian(*J.as_args())
    assert p.kernelInfo(adjoint=True)["kernel"] == "big"
    p.setTuning(checkpoint_every=256)
    _check(p.getGradient(), oracle_gradient(w, J, 256), J.sizes, "2001-DOF reference adjoint")
    p.setAdjointMode("exact")
    _check(p.getGradient(), oracle_exact(w, J, 256), J.sizes, "2001-DOF exact adjoint")

    _case("big: 2001-DOF chain beside 6 chains of 8 DOF, auto kernel (mixed)")
    small = W.chains(n_systems=6, d=8, steps=2000, steps_per_sample=10, seed=5)
    big = W.chains(n_systems=1, d=1000, steps=2000, steps_per_sample=10, seed=6)
    cols = []
    for a in ("vecM", "vecB", "vecK", "vecGamma", "vecLambda", "vecForce", "vecQ0"):
        cols.append(list(getattr(small, a)[:3]) + list(getattr(big, a))
                    + list(getattr(small, a)[3:]))
    exc = np.concatenate([small.excitations, big.excitations])
    wm = W.Workload("mixed", *cols, exc, small.sample_rate, 2000,
                    exc_of_system=np.array([0, 1, 2, 6, 3, 4, 5]),
                    steps_of_system=np.array([2000, 1500, 999, 1800, 2000, 3, 1234]),
                    interp=small.interp)
    Jm = ADJ.energy_gradient_inputs(wm)
    pm = W.make_solver(wm)
    assert pm.kernelInfo()["kernel"] == "mixed"
    _check(pm.getAmplitudes(), oracle_run(wm).q, _sizes(wm), "mixed forward")
    pm.setJacobian(*Jm.as_args())
    pm.setTuning(checkpoint_every=128)
    _check(pm.getGradient(), oracle_gradient(wm, Jm, 128), Jm.sizes, "mixed reference adjoint")
    pm.setAdjointMode("exact")
    assert pm.kernelInfo(adjoint=True)["kernel"] == "mixed"
    _check(pm.getGradient(), oracle_exact(wm, Jm, 128), Jm.sizes, "mixed exact adjoint")
