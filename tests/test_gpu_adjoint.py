"""GPU parity of the checkpointed adjoint (getGradient) against the oracle's
adjoint recursion on the same inputs: FP64 within 1e-10 relative per
output (same floor as the forward tests), for several checkpoint intervals,
both solvers, Cartesian and mapped layouts, uneven step counts and adjoint
systems with Psi terms; plus central finite differences of the GPU forward
(the reference recursion is a discretised continuous adjoint, so agreement
with FD is to the scheme's order: 2e-4 relative at dt = 1e-3).
"""
import numpy as np
import pytest

import oracle as O
from pasif_amd import adjoint as ADJ
from pasif_amd import workloads as W
from tests.test_gpu_forward import TOL64, as_coo, rel_err

pytestmark = pytest.mark.gpu


def _layout(w, asm):
    if w.exc_of_system is None and w.steps_of_system is None:
        return O.cartesian_layout(asm.n, len(w.excitations), w.numsteps)
    exc = w.exc_of_system if w.exc_of_system is not None else np.zeros(w.n_systems, int)
    st = w.steps_of_system if w.steps_of_system is not None else np.full(w.n_systems, w.numsteps)
    return O.mapped_layout(asm.sizes, exc, st)


def oracle_gradient(w, J, k):
    fwd = O.assemble_forward(w.vecM, w.vecB, w.vecK, w.vecGamma, w.vecLambda, w.vecForce, w.vecQ0)
    adj = O.assemble_adjoint(*J.as_args(), n_fwd=fwd.n)
    lay = _layout(w, fwd)
    if w.steps_of_system is not None:
        ast = np.repeat(np.asarray(w.steps_of_system), J.sizes)[None, :]
    else:
        ast = np.full((lay.P, adj.n), w.numsteps)
    F = O.Forcing(np.asarray(w.excitations, dtype=np.float64), w.interp)
    return O.adjoint_gradient(fwd, adj, F, w.sample_rate, lay, ast, checkpoint_every=k)


def _print_case(name):
    print(f"[gpu case] {name}", flush=True)


def test_chain_gradient_checkpoint_intervals_and_plain():
    _print_case("chains 12 x d8 x 300 adjoint")
    w = W.chains(n_systems=12, d=8, steps=300, steps_per_sample=10)
    J = ADJ.energy_gradient_inputs(w)
    p = W.make_solver(w)
    p.setJacobian(*J.as_args())
    ref = oracle_gradient(w, J, 64)
    outs = []
    for k in (1, 7, 64, 300, 1000):
        p.setTuning(checkpoint_every=k)
        g = p.getGradient()
        assert rel_err(g, ref) < TOL64, k
        outs.append(g)
    for g in outs[1:]:
        np.testing.assert_array_equal(g, outs[0])
    p.setTuning(checkpoint_every=64)
    p.setSolver("plain")
    gp = p.getGradient()
    np.testing.assert_array_equal(gp, outs[0])
    assert p._lastTiming()["launches"] == 3 * 300


def test_duffing_gradient_and_forward_amplitudes_together():
    _print_case("duffing 8 adjoint")
    w = W.duffing(n_systems=8, steps=700)
    params = [[("K", 1, 0), ("Lambda", 0)] for _ in range(w.n_systems)]
    terms = [np.array([0.0, 0.0, 1.0])] * w.n_systems
    J = ADJ.jacobian_inputs(w.vecM, w.vecB, w.vecK, w.vecGamma, w.vecLambda, params, terms)
    p = W.make_solver(w)
    p.setKernel("tile")
    p.setJacobian(*J.as_args())
    g = p.getGradient()
    assert rel_err(g, oracle_gradient(w, J, 256)) < TOL64


def test_cartesian_random_adjoint_with_psi():
    _print_case("random cartesian adjoint with psi")
    from tests.helpers import random_system
    rng = np.random.default_rng(31)
    fw = [random_system(rng, n, nG=4, nL=3, scale=0.2) for n in (3, 4)]
    vM, vB, vK, vG, vL, _, vF, vQ = (list(x) for x in zip(*fw))
    vG = [as_coo(t) for t in vG]
    vL = [as_coo(t) for t in vL]
    exc = rng.normal(0, 1, (2, 101))
    w = W.Workload("random", vM, vB, vK, vG, vL, vF, vQ, exc, 100, 100)
    aw = [random_system(rng, na, nG=5, nL=4, nP=3, n_set=nf, scale=0.2)
          for na, nf in ((5, 3), (6, 4))]
    aM, aB, aK, aG, aL, aP, aF, aQ = (list(x) for x in zip(*aw))
    J = ADJ.JacobianInputs(aM, aB, aK, [as_coo(t) for t in aG], [as_coo(t) for t in aL], aF, aQ,
                           [as_coo(t) for t in aP], sizes=np.array([5, 6]))
    p = W.make_solver(w)
    p.setKernel("tile")
    p.setJacobian(*J.as_args())
    g = p.getGradient()
    ref = oracle_gradient(w, J, 16)
    assert g.shape == ref.reshape(-1).shape
    assert rel_err(g, ref) < TOL64


def test_heterogeneous_gradient():
    _print_case("heterogeneous 30 adjoint")
    w = W.heterogeneous(n_systems=30, tmin=40, tmax=300, steps_per_sample=5, n_signals=3)
    J = ADJ.energy_gradient_inputs(w)
    p = W.make_solver(w)
    p.setKernel("tile")
    p.setJacobian(*J.as_args())
    p.setTuning(checkpoint_every=32)
    g = p.getGradient()
    assert rel_err(g, oracle_gradient(w, J, 32)) < TOL64


def test_gradient_against_gpu_finite_differences():
    _print_case("chains FD")
    w = W.chains(n_systems=4, d=4, steps=1000, steps_per_sample=10)
    J = ADJ.energy_gradient_inputs(w)
    p = W.make_solver(w)
    p.setJacobian(*J.as_args())
    g = p.getGradient()
    d = 4
    n = 2 * d + 1
    na = n + 2 * d
    h = 1e-6
    rng = np.random.default_rng(5)
    for s, i in [(0, 0), (1, 2), (2, 3), (3, 1)]:
        # K_ii of system s
        def energy(delta):
            K = w.vecK[s].tolil()
            K[d + i, i] += delta
            vK = list(w.vecK)
            vK[s] = K.tocoo()
            q = W.make_solver(W.Workload(w.name, w.vecM, w.vecB, vK, w.vecGamma, w.vecLambda,
                                         w.vecForce, w.vecQ0, w.excitations, w.sample_rate,
                                         w.numsteps, w.exc_of_system, None, w.interp)).getAmplitudes()
            return q[s * n + n - 1]
        fd = (energy(h) - energy(-h)) / (2 * h)
        ad = g[s * na + n + i]
        assert abs(ad - fd) <= 2e-4 * abs(fd) + 1e-10, (s, i, ad, fd)


def test_concatenated_adjoint_inputs_give_the_same_gradients():
    """setJacobianBlocks (the adjoint systems built concatenated) and
    setJacobian (built per system) hand the driver the same arrays: the
    gradients are bitwise equal."""
    print("[gpu case] jacobian blocks vs per-system inputs", flush=True)
    w = W.heterogeneous(48, 50, 600, steps_per_sample=5, n_signals=3,
                        dof_choices=(1, 2, 4, 8), dof_probs=(1, 1, 1, 1))
    a = W.make_solver(w)
    a.setJacobian(*ADJ.energy_gradient_inputs(w).as_args())
    b = W.make_solver(w)
    b.setJacobianBlocks(ADJ.energy_gradient_blocks(w))
    ga = np.asarray(a.getGradient())
    gb = np.asarray(b.getGradient())
    assert ga.shape == gb.shape and np.array_equal(ga, gb)
