"""The memory plan on the host (no GPU): the run-buffer footprint a solve
would allocate, checkpoint blocks sized by each tile's own step count, and
the autotuner's estimate of a full-size footprint from a probe."""
import numpy as np

from pasif_amd import adjoint as ADJ
from pasif_amd import autotune as AT
from pasif_amd import workloads as W


def _host(w, adjoint=True):
    import os
    os.environ["LOCAL_RANK"] = "-1"
    try:
        p = W.make_solver(w)
    finally:
        del os.environ["LOCAL_RANK"]
    if adjoint:
        p.setJacobian(*ADJ.energy_gradient_inputs(w).as_args())
    return p


def test_tile_checkpoints_scale_with_the_interval():
    w = W.chains(64, 4, 1000, steps_per_sample=10)
    p = _host(w)
    p.setKernel("tile")
    p.setTuning(tile_lanes=64, checkpoint_every=16)
    a = p._runFootprint(True)
    p.setTuning(tile_lanes=64, checkpoint_every=50)
    b = p._runFootprint(True)
    # every tile runs 1000 steps: ceil(1000/16) = 63 and ceil(1000/50) = 20
    # slots of one value per lane
    assert a["ckpt"] * 20 == b["ckpt"] * 63
    assert a["ckpt"] % (63 * 64 * 8) == 0
    assert a["free"] == 0 and a["run"] > a["ckpt"]


def test_tile_checkpoints_sized_by_each_tiles_steps():
    # half the systems run 1000 steps, half 100: packed longest first, the
    # short ones fill their own tiles, whose checkpoint blocks are 10x smaller
    w = W.chains(64, 4, 1000, steps_per_sample=10)
    p = _host(w)
    p.setKernel("tile")
    p.setTuning(tile_lanes=64, checkpoint_every=10)
    uniform = p._runFootprint(True)["ckpt"]
    steps = np.where(np.arange(64) < 32, 1000, 100)
    p.setSystemExcitations(np.arange(64), steps)
    uneven = p._runFootprint(True)["ckpt"]
    assert uneven * 20 == uniform * 11      # (100 + 10) / 2 slots per tile on average


def test_full_size_estimate_from_a_probe():
    w = W.heterogeneous(200, 100, 4000, steps_per_sample=5, n_signals=3,
                        dof_choices=(2, 8), dof_probs=(1, 1))
    p = _host(w)
    p.setKernel("tile")
    p.setTuning(tile_lanes=128, checkpoint_every=8)
    full = p._runFootprint(True)
    cap = AT.probe_caps(w.steps_of_system, 400)
    p.setSystemExcitations(p._sys_exc, cap)
    probe = p._runFootprint(True)
    est = AT.full_size_bytes(probe, int(cap.max()), int(w.steps_of_system.max()))
    # ceil() per tile makes the probe's slots round up: the estimate is an
    # upper bound within a few percent
    assert full["run"] <= est <= 1.10 * full["run"]
    assert AT.fits(probe, int(cap.max()), int(w.steps_of_system.max()))
    tight = dict(probe, free=int(est / 0.8) - 1)
    assert not AT.fits(tight, int(cap.max()), int(w.steps_of_system.max()))
