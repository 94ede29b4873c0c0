"""GPU parity of the forward path: the HIP library (through the PASIf API and
the C ABI) against the FP64 NumPy oracle on the same seeded inputs.

Tolerance: FP64 results agree within 1e-10 relative per output,
``|g - o| / max(|o|, 1e-8 * max|o|)`` (the floor keeps outputs that sit at
zero from dividing by zero).  FP32 results are compared with the FP32
oracle (same recursion in float32) within 1e-4 and with the FP64 oracle
within 5e-3 relative, same floor.
"""
import numpy as np
import pytest

import oracle as O
from pasif_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def rel_err(g, o, floor_frac=1e-8):
    g = np.asarray(g, dtype=np.float64).reshape(-1)
    o = np.asarray(o, dtype=np.float64).reshape(-1)
    floor = floor_frac * max(np.abs(o).max(), 1e-300)
    return float(np.max(np.abs(g - o) / np.maximum(np.abs(o), floor)))


def oracle_run(w, dtype=np.float64, save_steps=0, window=0, n_exc_cartesian=False):
    asm = O.assemble_forward(w.vecM, w.vecB, w.vecK, w.vecGamma, w.vecLambda, w.vecForce, w.vecQ0)
    F = O.Forcing(np.asarray(w.excitations, dtype=np.float64), w.interp)
    if w.exc_of_system is None and w.steps_of_system is None:
        lay = O.cartesian_layout(asm.n, len(w.excitations), w.numsteps)
    else:
        exc = w.exc_of_system if w.exc_of_system is not None else np.zeros(w.n_systems, int)
        st = w.steps_of_system if w.steps_of_system is not None else \
            np.full(w.n_systems, w.numsteps)
        lay = O.mapped_layout(asm.sizes, exc, st)
    return O.rk4_forward(asm, F, w.sample_rate, lay, dtype=dtype, save_steps=save_steps,
                         window=window)


def as_coo(t):
    from pasif_amd import coo_tensor
    c = coo_tensor(t.dimensions)
    c.val = list(t.val)
    c.indices = list(t.indices)
    return c


def _print_case(name):
    print(f"[gpu case] {name}", flush=True)


def test_duffing_small_fp64():
    _print_case("duffing 16 x 1500")
    w = W.duffing(n_systems=16, steps=1500)
    p = W.make_solver(w)
    p.setKernel("tile")
    q, mx = p.getMaxAmplitudes(window=300)
    ref = oracle_run(w, window=300)
    assert rel_err(q, ref.q) < TOL64
    assert rel_err(mx, ref.maxabs) < TOL64


def test_two_dof_fp64_and_plain_matches_persistent():
    _print_case("two_dof 64 x 800")
    w = W.two_dof(n_systems=64, steps=800)
    p = W.make_solver(w)
    p.setKernel("tile")
    q = p.getAmplitudes()
    ref = oracle_run(w)
    assert rel_err(q, ref.q) < TOL64
    p.setSolver("plain")
    q_plain = p.getAmplitudes()
    np.testing.assert_array_equal(q_plain, q)
    assert p._lastTiming()["launches"] == 800


def test_chain_interp_non_power_of_two_window():
    _print_case("chains 24 x d8 x 600 interp W=20")
    w = W.chains(n_systems=24, d=8, steps=600, steps_per_sample=10)
    p = W.make_solver(w)
    q = p.getAmplitudes()
    ref = oracle_run(w)
    assert rel_err(q, ref.q) < TOL64


def test_cartesian_product_random_systems_and_trajectory():
    _print_case("random cartesian 3 exc")
    from tests.helpers import random_system
    rng = np.random.default_rng(7)
    systems = [random_system(rng, n, nG=5, nL=4) for n in (3, 5, 2, 6)]
    vM, vB, vK, vG, vL, _, vF, vQ = (list(x) for x in zip(*systems))
    vG = [as_coo(t) for t in vG]
    vL = [as_coo(t) for t in vL]
    exc = rng.normal(0, 1, (3, 121))
    w = W.Workload("random", vM, vB, vK, vG, vL, vF, vQ, exc, 100, 120)
    p = W.make_solver(w)
    p.setKernel("tile")
    q = p.getAmplitudes()
    ref = oracle_run(w, save_steps=7)
    assert q.shape == ref.q.reshape(-1).shape
    assert rel_err(q, ref.q) < TOL64
    tr = p.getTrajectory(saveSteps=7)
    assert tr.shape == (ref.q.size + 1, 120 // 7)
    np.testing.assert_allclose(tr[0], np.arange(120 // 7) * 7 / 100)
    assert rel_err(tr[1:].T, ref.traj.reshape(ref.traj.shape[0], -1)) < TOL64


def test_split_rows_and_tile_shapes_identical():
    # chains with the energy probe row (many Gamma terms) force rows to span
    # several lanes; results must not depend on the tile shape
    _print_case("chains 40 x d16 tile shapes")
    w = W.chains(n_systems=40, d=16, steps=300, steps_per_sample=10)
    p = W.make_solver(w)
    ref = oracle_run(w)
    outs = []
    for lanes, slots in ((64, 1), (128, 1), (256, 2), (0, 0)):
        p.setTuning(tile_lanes=lanes, slot_config=slots)
        q = p.getAmplitudes()
        assert rel_err(q, ref.q) < TOL64, (lanes, slots)
        outs.append(q)
    # same slot configuration -> bitwise identical whatever the packing
    np.testing.assert_array_equal(outs[0], outs[1])


def test_heterogeneous_steps_and_sizes():
    _print_case("heterogeneous 60 systems")
    w = W.heterogeneous(n_systems=60, tmin=50, tmax=400, steps_per_sample=5, n_signals=4)
    p = W.make_solver(w)
    p.setKernel("tile")
    q = p.getAmplitudes()
    ref = oracle_run(w)
    assert rel_err(q, ref.q) < TOL64


def test_fp32_against_fp32_and_fp64_oracle():
    _print_case("two_dof fp32")
    w = W.two_dof(n_systems=32, steps=1000)
    p = W.make_solver(w)
    p.setKernel("tile")
    p.setPrecision(32)
    q = p.getAmplitudes()
    r32 = oracle_run(w, dtype=np.float32)
    r64 = oracle_run(w)
    assert rel_err(q, r32.q) < 1e-4
    assert rel_err(q, r64.q) < 5e-3
