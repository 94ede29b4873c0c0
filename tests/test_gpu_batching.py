"""GPU checks of the heterogeneous batching form: systems of uneven size and
step count in one launch, the hipGraph-sliced solver and the autotuner.

* graph-sliced runs are bitwise identical to the single persistent launch
  (same kernel, same tiles; only the launch boundaries move), forward and
  adjoint;
* the autotuned configuration still matches the oracle within 1e-10.
"""
import numpy as np
import pytest

from pasif_amd import adjoint as ADJ
from pasif_amd import workloads as W
from tests.test_gpu_adjoint import oracle_gradient
from tests.test_gpu_forward import TOL64, oracle_run, rel_err

pytestmark = pytest.mark.gpu


def test_graph_sliced_equals_persistent_forward_and_adjoint():
    print("[gpu case] graph slicing heterogeneous 50", flush=True)
    w = W.heterogeneous(n_systems=50, tmin=100, tmax=900, steps_per_sample=5, n_signals=4)
    p = W.make_solver(w)
    J = ADJ.energy_gradient_inputs(w)
    p.setJacobian(*J.as_args())
    p.setTuning(checkpoint_every=64)
    q = p.getAmplitudes()
    g = p.getGradient()
    for slice_, per in ((1, 7), (37, 3), (256, 16)):
        p.setSolver("graph", steps_per_launch=slice_, launches_per_graph=per)
        np.testing.assert_array_equal(p.getAmplitudes(), q)
        np.testing.assert_array_equal(p.getGradient(), g)
    p.setSolver("persistent")
    assert rel_err(q, oracle_run(w).q) < TOL64
    assert rel_err(g, oracle_gradient(w, J, 64)) < TOL64


def test_autotuned_configuration_keeps_parity(tmp_path, monkeypatch):
    print("[gpu case] autotune chains", flush=True)
    monkeypatch.setenv("PASIF_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    import importlib

    from pasif_amd import autotune as AT
    importlib.reload(AT)
    w = W.chains(n_systems=64, d=16, steps=500, steps_per_sample=10)
    p = W.make_solver(w)
    # on the lane kernel (the default here) tile shapes change nothing
    # (only the group sizes are probed: the model's, one size narrower, wider)
    lane = AT.autotune(p, probe_steps=100, use_cache=False)
    assert lane["lane_only"] and 1 <= lane["trials"] <= 3 and lane["lane_bias"] in (-1, 0, 1)
    assert rel_err(p.getAmplitudes(), oracle_run(w).q) < TOL64
    p.setKernel("tile")
    best = AT.autotune(p, probe_steps=100)
    assert best["trials"] >= 3 and not best["lane_only"]
    q = p.getAmplitudes()
    assert rel_err(q, oracle_run(w).q) < TOL64
    again = AT.autotune(p, probe_steps=100)       # cached
    assert again["tile_lanes"] == best["tile_lanes"]
