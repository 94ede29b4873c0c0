"""GPU parity of the lane kernel (one lane per instance, state in registers,
code generated per sparsity structure) against the FP64 oracle and the tile
kernel, through the PASIf API and the C ABI.

Tolerances as for the tile kernel: FP64 within 1e-10 relative per output
(floor 1e-8 * max|o|); FP32 within 1e-4 of the FP32 oracle and 5e-3 of the
FP64 oracle.  Launch boundaries do not change the arithmetic, so the plain
and graph-sliced solvers are bitwise equal to the persistent launch.
"""
import numpy as np
import pytest

from pasif_amd import adjoint as ADJ
from pasif_amd import workloads as W
from tests.test_gpu_adjoint import oracle_gradient
from tests.test_gpu_forward import TOL64, as_coo, oracle_run, rel_err

pytestmark = pytest.mark.gpu


def _case(name):
    print(f"[gpu case] lane {name}", flush=True)


def _lane(w, lanes_per_system=0):
    p = W.make_solver(w)
    p.setKernel("lane", lanes_per_system)
    return p


def _random_cartesian(seed=7):
    from tests.helpers import random_system
    rng = np.random.default_rng(seed)
    systems = [random_system(rng, n, nG=5, nL=4) for n in (3, 5, 2, 6)]
    vM, vB, vK, vG, vL, _, vF, vQ = (list(x) for x in zip(*systems))
    vG = [as_coo(t) for t in vG]
    vL = [as_coo(t) for t in vL]
    exc = rng.normal(0, 1, (3, 121))
    return W.Workload("random", vM, vB, vK, vG, vL, vF, vQ, exc, 100, 120)


def test_lane_forward_duffing_two_dof_chain_interp():
    _case("forward duffing / two_dof / chains d2 interp")
    w = W.duffing(n_systems=16, steps=1500)
    p = _lane(w)
    assert p.kernelInfo()["kernel"] == "lane"
    q, mx = p.getMaxAmplitudes(window=300)
    ref = oracle_run(w, window=300)
    assert rel_err(q, ref.q) < TOL64
    assert rel_err(mx, ref.maxabs) < TOL64

    w = W.two_dof(n_systems=100, steps=900)
    p = _lane(w)
    q = p.getAmplitudes()
    assert rel_err(q, oracle_run(w).q) < TOL64
    p.setKernel("tile")
    assert rel_err(p.getAmplitudes(), q) < TOL64

    w = W.chains(n_systems=70, d=2, steps=700, steps_per_sample=10)
    p = _lane(w)
    assert rel_err(p.getAmplitudes(), oracle_run(w).q) < TOL64


def test_lane_cartesian_classes_and_trajectory():
    _case("random cartesian, 4 classes, trajectory")
    w = _random_cartesian()
    p = _lane(w)
    info = p.kernelInfo()
    assert info["kernel"] == "lane" and info["classes"] == 4
    q = p.getAmplitudes()
    ref = oracle_run(w, save_steps=7)
    assert rel_err(q, ref.q) < TOL64
    tr = p.getTrajectory(saveSteps=7)
    assert rel_err(tr[1:].T, ref.traj.reshape(ref.traj.shape[0], -1)) < TOL64


def test_lane_uneven_steps_and_solver_modes_bitwise():
    _case("heterogeneous d1/d2/d4, plain + graph")
    w = W.heterogeneous(n_systems=150, tmin=30, tmax=500, steps_per_sample=5, n_signals=4,
                        dof_choices=(1, 2, 4), dof_probs=(0.4, 0.4, 0.2))
    p = _lane(w)
    assert p.kernelInfo()["classes"] == 3
    q = p.getAmplitudes()
    assert rel_err(q, oracle_run(w).q) < TOL64
    p.setSolver("plain")
    np.testing.assert_array_equal(p.getAmplitudes(), q)
    assert p._lastTiming()["launches"] == int(np.max(w.steps_of_system))
    p.setSolver("graph", steps_per_launch=37, launches_per_graph=3)
    np.testing.assert_array_equal(p.getAmplitudes(), q)


def test_lane_fp32():
    _case("two_dof fp32")
    w = W.two_dof(n_systems=32, steps=1000)
    p = _lane(w)
    p.setPrecision(32)
    q = p.getAmplitudes()
    assert rel_err(q, oracle_run(w, dtype=np.float32).q) < 1e-4
    assert rel_err(q, oracle_run(w).q) < 5e-3


def test_lane_adjoint_checkpoints_and_solvers():
    _case("adjoint chains d2")
    w = W.chains(n_systems=20, d=2, steps=400, steps_per_sample=10)
    J = ADJ.energy_gradient_inputs(w)
    p = _lane(w)
    p.setJacobian(*J.as_args())
    outs = []
    for k in (1, 7, 256):
        p.setTuning(checkpoint_every=k)
        info = p.kernelInfo(adjoint=True)
        assert info["kernel"] == "lane" and 1 <= info["checkpoint_every"] <= k
        g = p.getGradient()
        assert rel_err(g, oracle_gradient(w, J, 64)) < TOL64, k
        outs.append(g)
    for g in outs[1:]:
        np.testing.assert_array_equal(g, outs[0])
    p.setSolver("plain")
    np.testing.assert_array_equal(p.getGradient(), outs[-1])
    assert p._lastTiming()["launches"] == 3 * 400
    p.setSolver("graph", steps_per_launch=29, launches_per_graph=5)
    np.testing.assert_array_equal(p.getGradient(), outs[-1])


def test_lane_adjoint_duffing_and_psi():
    _case("adjoint duffing + random psi")
    w = W.duffing(n_systems=8, steps=700)
    params = [[("K", 1, 0), ("Lambda", 0)] for _ in range(w.n_systems)]
    terms = [np.array([0.0, 0.0, 1.0])] * w.n_systems
    J = ADJ.jacobian_inputs(w.vecM, w.vecB, w.vecK, w.vecGamma, w.vecLambda, params, terms)
    p = _lane(w)
    p.setJacobian(*J.as_args())
    assert rel_err(p.getGradient(), oracle_gradient(w, J, 256)) < TOL64

    from tests.helpers import random_system
    rng = np.random.default_rng(31)
    fw = [random_system(rng, n, nG=4, nL=3, scale=0.2) for n in (3, 4)]
    vM, vB, vK, vG, vL, _, vF, vQ = (list(x) for x in zip(*fw))
    exc = rng.normal(0, 1, (2, 101))
    w = W.Workload("random", vM, vB, vK, [as_coo(t) for t in vG], [as_coo(t) for t in vL],
                   vF, vQ, exc, 100, 100)
    aw = [random_system(rng, na, nG=5, nL=4, nP=3, n_set=nf, scale=0.2)
          for na, nf in ((5, 3), (6, 4))]
    aM, aB, aK, aG, aL, aP, aF, aQ = (list(x) for x in zip(*aw))
    J = ADJ.JacobianInputs(aM, aB, aK, [as_coo(t) for t in aG], [as_coo(t) for t in aL], aF, aQ,
                           [as_coo(t) for t in aP], sizes=np.array([5, 6]))
    p = _lane(w)
    p.setJacobian(*J.as_args())
    assert rel_err(p.getGradient(), oracle_gradient(w, J, 16)) < TOL64


def _waves(p, adjoint=False):
    head = p._laneSource(adjoint).splitlines()[0]
    return int(head.split(" wave(s)")[0].rsplit(" ", 1)[1])


def test_multiwave_forward_interp_and_solver_modes():
    _case("multi-wave forward chains d8")
    w = W.chains(n_systems=70, d=8, steps=500, steps_per_sample=10)
    p = _lane(w, 1)
    assert _waves(p) > 1
    q = p.getAmplitudes()
    assert rel_err(q, oracle_run(w).q) < TOL64
    p.setSolver("plain")
    np.testing.assert_array_equal(p.getAmplitudes(), q)
    p.setSolver("graph", steps_per_launch=23, launches_per_graph=4)
    np.testing.assert_array_equal(p.getAmplitudes(), q)
    p.setSolver("persistent")
    p.setKernel("tile")
    assert rel_err(p.getAmplitudes(), q) < TOL64


def test_multiwave_adjoint_checkpoints_and_solvers():
    _case("multi-wave adjoint chains d8")
    w = W.chains(n_systems=20, d=8, steps=300, steps_per_sample=10)
    J = ADJ.energy_gradient_inputs(w)
    p = _lane(w, 1)
    p.setJacobian(*J.as_args())
    assert _waves(p, True) > 1
    ref = oracle_gradient(w, J, 64)
    outs = []
    for k in (5, 64):
        p.setTuning(checkpoint_every=k)
        g = p.getGradient()
        assert rel_err(g, ref) < TOL64, k
        outs.append(g)
    np.testing.assert_array_equal(outs[0], outs[1])
    p.setSolver("plain")
    np.testing.assert_array_equal(p.getGradient(), outs[1])
    p.setSolver("graph", steps_per_launch=17, launches_per_graph=6)
    np.testing.assert_array_equal(p.getGradient(), outs[1])


def test_multiwave_mixed_sizes_and_cfg2_shape():
    _case("multi-wave heterogeneous d1..d16 + chains d16 adjoint")
    w = W.heterogeneous(n_systems=120, tmin=30, tmax=400, steps_per_sample=5, n_signals=4,
                        dof_choices=(1, 2, 4, 8, 16), dof_probs=(0.3, 0.25, 0.2, 0.15, 0.1))
    p = _lane(w, 1)
    assert p.kernelInfo()["classes"] == 5
    assert rel_err(p.getAmplitudes(), oracle_run(w).q) < TOL64
    w = W.chains(n_systems=40, d=16, steps=200, steps_per_sample=10)
    J = ADJ.energy_gradient_inputs(w)
    p = _lane(w, 1)
    p.setJacobian(*J.as_args())
    assert rel_err(p.getAmplitudes(), oracle_run(w).q) < TOL64
    assert rel_err(p.getGradient(), oracle_gradient(w, J, 64)) < TOL64


def test_mixed_lane_and_tile_in_one_solve():
    _case("mixed lane + tile, heterogeneous d1..d64, forward + adjoint + solvers")
    w = W.heterogeneous(n_systems=90, tmin=30, tmax=300, steps_per_sample=5, n_signals=4)
    J = ADJ.energy_gradient_inputs(w)
    p = W.make_solver(w)
    p.setKernel("auto", 1)  # one lane per instance: the 64-DOF systems go to tiles
    p.setJacobian(*J.as_args())
    assert p.kernelInfo()["kernel"] == "mixed"
    q = p.getAmplitudes()
    assert rel_err(q, oracle_run(w).q) < TOL64
    g = p.getGradient()
    assert rel_err(g, oracle_gradient(w, J, 256)) < TOL64
    p.setSolver("plain")
    np.testing.assert_array_equal(p.getAmplitudes(), q)
    p.setSolver("graph", steps_per_launch=41, launches_per_graph=3)
    np.testing.assert_array_equal(p.getGradient(), g)
    tr = p.getTrajectory(saveSteps=11)
    ref = oracle_run(w, save_steps=11)
    assert rel_err(tr[1:].T, ref.traj.reshape(ref.traj.shape[0], -1)) < TOL64
