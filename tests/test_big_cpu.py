"""Host-side selection and memory plan of the row-parallel path (no GPU):
systems too large for a tile go to it under the automatic kernel choice,
beside the tile and lane kernels; its checkpoint and run buffers follow the
instances' own sizes and step counts."""
import numpy as np
import pytest

from pasif_amd import adjoint as ADJ
from pasif_amd import workloads as W
from tests.test_memory_plan import _host


def test_large_chain_goes_to_the_row_parallel_path():
    w = W.chains(n_systems=2, d=1000, steps=2000, steps_per_sample=10)
    p = _host(w)
    assert p.kernelInfo()["kernel"] == "big"
    info = p.kernelInfo(adjoint=True)
    assert info["kernel"] == "big" and info["classes"] == 0
    p.setTuning(checkpoint_every=256)
    assert p.kernelInfo(adjoint=True)["checkpoint_every"] == 256
    fp = p._runFootprint(True)
    # ceil(2000 / 256) = 8 checkpoint rows of 2001 values per instance
    assert fp["ckpt"] == 2 * 8 * 2001 * 8
    # the tile kernel refuses these systems
    p.setKernel("tile")
    with pytest.raises(Exception):
        p.kernelInfo()


def test_small_systems_stay_on_tiles_beside_a_large_one():
    small = W.chains(n_systems=3, d=4, steps=300, steps_per_sample=10)
    big = W.chains(n_systems=1, d=700, steps=300, steps_per_sample=10)
    cols = [list(getattr(small, a)) + list(getattr(big, a))
            for a in ("vecM", "vecB", "vecK", "vecGamma", "vecLambda", "vecForce", "vecQ0")]
    w = W.Workload("mixed", *cols, np.concatenate([small.excitations, big.excitations]),
                   small.sample_rate, 300, exc_of_system=np.arange(4),
                   steps_of_system=np.array([300, 300, 300, 100]), interp=small.interp)
    p = _host(w)
    p.setKernel("tile")
    with pytest.raises(Exception):
        p.kernelInfo()
    p.setKernel("auto")
    p.setAdjointMode("exact")    # the exact adjoint: tile kernel plus the big path
    assert p.kernelInfo(adjoint=True)["kernel"] == "mixed"
    p.setTuning(checkpoint_every=32)
    fp = p._runFootprint(True)
    # the large chain (1401 rows) runs 100 steps: ceil(100 / 32) = 4 rows of
    # checkpoints; the tiles hold their own beside them
    assert fp["ckpt"] > 4 * 1401 * 8
    p.setKernel("big")
    big_only = p._runFootprint(True)["ckpt"]
    assert big_only == (3 * 10 * 9 + 4 * 1401) * 8
    assert ADJ is not None
