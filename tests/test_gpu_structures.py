"""Batches of many distinct sparsity structures on the GPU against the
oracle: the heterogeneous batch with each system's sparsity drawn from 12
variants of its size (66 structures: nearest-neighbour Gamma couplings
dropped at random, up to three extra next-nearest springs), which the lane
planner merges into union classes (zero coefficients where a system lacks a
term) of one generated kernel.  FP64 within 1e-10 per output
(``oracle.parity``), forward amplitudes and the reference adjoint recursion,
and the same gradients from the tile kernel.
"""
import numpy as np
import pytest

from oracle import parity as PAR
from pasif_amd import adjoint as ADJ
from pasif_amd import workloads as W
from tests.test_gpu_adjoint import oracle_gradient
from tests.test_gpu_forward import oracle_run

pytestmark = pytest.mark.gpu


def _case(name):
    print(f"[gpu case] {name}", flush=True)


def test_many_structures_forward_and_adjoint_match_the_oracle():
    _case("66 structures: heterogeneous 600 systems x 200-1500 steps, forward + adjoint")
    w = W.heterogeneous(600, 200, 1500, steps_per_sample=10, n_signals=8, variants_per_size=12,
                        seed=11)
    sizes = [m.shape[0] for m in w.vecM]
    p = W.make_solver(w)
    info = p.kernelInfo()
    print("  forward plan:", info, flush=True)
    assert info["kernel"] in ("lane", "mixed") and info["classes"] <= 8
    q = p.getAmplitudes()
    r = PAR.report(q, oracle_run(w).q.reshape(-1), sizes)
    assert r.max_rel < PAR.TOL64, r.as_dict()
    J = ADJ.energy_gradient_inputs(w)
    p.setJacobian(*J.as_args())
    p.setTuning(checkpoint_every=64)
    info = p.kernelInfo(adjoint=True)
    print("  adjoint plan:", info, flush=True)
    assert info["kernel"] == "lane" and info["classes"] <= 8
    g = p.getGradient()
    ref = oracle_gradient(w, J, 64)
    r = PAR.report(g, ref.reshape(-1), J.sizes)
    assert r.max_rel < PAR.TOL64, r.as_dict()
    # every structure on tiles agrees with the merged lane classes
    p.setKernel("tile")
    gt = p.getGradient()
    r = PAR.report(gt, ref.reshape(-1), J.sizes)
    assert r.max_rel < PAR.TOL64, r.as_dict()
