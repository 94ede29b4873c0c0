"""FP64 parity at the benchmark configurations' full sizes, per output.

Both kernel families (the lane kernel the library picks and the tile kernel)
run the full workloads on the GPU and are compared with the FP64 oracle's
outputs committed in ``tests/golden/`` (written by
``tests/golden/make_fullsize.py``; recomputing them takes minutes on a host):

* 4,096 two-DOF systems x 50,000 steps: every final state;
* 8,192 16-DOF chains x 20,000 steps, forward + adjoint: the final states
  and energy gradients of 256 chains (including chain 3088, the largest
  lane-vs-tile gradient gap).

Metric (``oracle.parity``): per output, ``|g - o| / max(|o|, 1e-4 max_s |o|)``
with the floor taken over the outputs of the same system; tolerance 1e-10.
The floor is where the oracle's own outputs stop being determined to 1e-10:
the same oracle with every coefficient moved by one ULP differs from itself
by up to 7.5e-10 without the floor and 5e-11 with it (checked below from the
fixtures).  FP32: against the FP32 oracle within ``TOL32`` = 3e-2 under the
same metric, stated from the FP32 oracle's own spread: moving its
coefficients by one FP32 ULP changes its outputs by up to 1.9e-2 (and the
FP32 oracle is 6.1e-2 from the FP64 one after 50,000 steps).
"""
import os

import numpy as np
import pytest

from oracle import parity as PAR
from pasif_amd import adjoint as ADJ
from pasif_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL32 = 3e-2


def _case(name):
    print(f"[gpu case] {name}", flush=True)


@pytest.fixture(scope="module")
def two_dof_gold():
    return np.load(os.path.join(GOLD, "fullsize_two_dof.npz"))


@pytest.fixture(scope="module")
def chains_gold():
    return np.load(os.path.join(GOLD, "fullsize_chains.npz"))


@pytest.fixture(scope="module")
def two_dof_solver():
    w = W.two_dof()
    return W.make_solver(w)


def test_two_dof_fullsize_fp64_lane_and_tile(two_dof_gold, two_dof_solver):
    _case("two-DOF 4096 x 50000 FP64, lane and tile")
    p = two_dof_solver
    o = two_dof_gold["o"]
    sizes = np.full(4096, 4)
    sens = PAR.sensitivity(o, two_dof_gold["op"], sizes)
    assert sens["floored"] < PAR.TOL64 < sens["unfloored"]
    p.setPrecision(64)
    for kernel in ("auto", "tile"):
        p.setKernel(kernel)
        r = PAR.report(p.getAmplitudes(), o, sizes)
        print(f"  {kernel}: {r.as_dict()}", flush=True)
        assert r.max_rel < PAR.TOL64, (kernel, r.as_dict())
    p.setKernel("auto")


def test_two_dof_fullsize_fp32(two_dof_gold, two_dof_solver):
    _case("two-DOF 4096 x 50000 FP32, lane and tile")
    p = two_dof_solver
    o32 = two_dof_gold["o32"]
    sizes = np.full(4096, 4)
    # the FP32 oracle's own spread under a one-FP32-ULP perturbation sets
    # the bound: the GPU stays within TOL32 of the FP32 oracle
    assert PAR.sensitivity(o32, two_dof_gold["p32"], sizes)["floored"] < TOL32
    p.setPrecision(32)
    try:
        for kernel in ("auto", "tile"):
            p.setKernel(kernel)
            r = PAR.report(p.getAmplitudes(), o32, sizes)
            print(f"  {kernel}: {r.as_dict()}", flush=True)
            assert r.max_rel < TOL32, (kernel, r.as_dict())
    finally:
        p.setKernel("auto")
        p.setPrecision(64)


def test_chains_fullsize_fwd_adj_fp64_lane_and_tile(chains_gold):
    _case("chains 8192 x 16 DOF x 20000 FP64 forward + adjoint, lane and tile")
    w = W.chains()
    p = W.make_solver(w)
    p.setJacobian(*ADJ.energy_gradient_inputs(w).as_args())
    p.setTuning(checkpoint_every=256)
    ids = chains_gold["ids"]
    oq, og = chains_gold["oq"], chains_gold["og"]
    for kernel in ("auto", "tile"):
        p.setKernel(kernel)
        q = p.getAmplitudes().reshape(8192, 33)[ids]
        g = p.getGradient().reshape(8192, 65)[ids]
        rq = PAR.report(q, oq, np.full(len(ids), 33))
        rg = PAR.report(g, og, np.full(len(ids), 65))
        print(f"  {kernel}: states {rq.as_dict()}\n  {kernel}: gradients {rg.as_dict()}",
              flush=True)
        assert rq.max_rel < PAR.TOL64 and rg.max_rel < PAR.TOL64, (kernel, rq.max_rel,
                                                                  rg.max_rel)
