"""CPU checks of the C ABI (no GPU): the library loads, exports every entry
point ``include/pasif.h`` declares with the signatures the binding uses, and
its host-side validation (a handle created with gpu_id = -1 packs and checks
tiles without a device) rejects malformed inputs with status codes and
messages instead of aborting -- the reference exits the process with status
0 on errors (src/helpers.cuh:43-51).
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from pasif_amd import _lib
from pasif_amd import workloads as W
from pasif_amd.PASIfgpu import __GpuDriver as Driver

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "pasif.h")


def declared():
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(pasif_\w+)\s*\(",
                                 open(HEADER).read(), re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    from pasif_amd import serving
    assert set(names) == set(_lib.exported_symbols()) | set(serving.pool_symbols())
    assert b"gfx950" in lib.pasif_version()


def host_driver(w, **kw):
    """A driver without a device (gpu_id = -1): setters, assembly, packing and
    validation run; solving is refused."""
    d = Driver(list(w.excitations), w.sample_rate, w.numsteps, False, False, False, -1)
    return d


def _load(d, w):
    from pasif_amd.PASIf import PASIf
    # reuse PASIf's assembly through a host-only instance
    os.environ["LOCAL_RANK"] = "-1"
    try:
        p = PASIf(list(w.excitations), w.sample_rate, w.numsteps)
        p.setSystems(w.vecM, w.vecB, w.vecK, w.vecGamma, w.vecLambda, w.vecForce, w.vecQ0)
        if w.exc_of_system is not None or w.steps_of_system is not None:
            p.setSystemExcitations(w.exc_of_system, w.steps_of_system)
        if w.interp is not None:
            p.setInterpolationMatrix(list(w.interp))
    finally:
        del os.environ["LOCAL_RANK"]
    return p


def _len(p, fn, *args):
    n = C.c_size_t(0)
    rc = fn(p._h, *args, None, 0, C.byref(n))
    return rc, n.value, p._pl.pasif_last_error(p._h).decode()


@pytest.mark.parametrize("maker", [lambda: W.duffing(16, 100), lambda: W.two_dof(64, 50),
                                   lambda: W.chains(12, 8, 60),
                                   lambda: W.heterogeneous(40, 50, 400, steps_per_sample=5,
                                                           n_signals=4)])
def test_host_packing_validates_every_tile_shape(maker):
    w = maker()
    p = _load(None, w)
    for lanes, slots in ((0, 0), (64, 1), (128, 3), (256, 2), (1024, 1)):
        p.setTuning(lanes, slots)
        rc, n, msg = _len(p, p._pl.pasif_get_amplitudes)
        if rc == 3 and slots in (2, 3):
            # wide / small slots exist only for tiles of <= 256 lanes; a system
            # that needs more is refused with a message, never packed wrong
            assert "slot configuration" in msg
            continue
        assert rc == 0, msg
        assert n == sum(m.shape[0] for m in w.vecM)
    # solving needs a device
    with pytest.raises(_lib.PasifError) as e:
        p.getAmplitudes()
    assert e.value.code == 5


def test_adjoint_packing_and_length():
    from pasif_amd import adjoint as ADJ
    w = W.chains(6, 8, 40)
    p = _load(None, w)
    J = ADJ.energy_gradient_inputs(w)
    p.setJacobian(*J.as_args())
    rc, n, msg = _len(p, p._pl.pasif_get_gradient, C.c_uint32(0))
    assert rc == 0, msg
    assert n == int(J.sizes.sum())


def test_error_paths_return_codes():
    w = W.duffing(4, 20)
    p = _load(None, w)
    # ragged excitations: the reference returns 1 and keeps going
    assert p._loadExcitationsSet([np.zeros(5), np.zeros(6)], 100) == 1
    # interpolation window must be >= 2 and the matrix needs two rows
    with pytest.raises(ValueError):
        p.setInterpolationMatrix([np.zeros(3)])
    with pytest.raises(ValueError):
        p.setInterpolationMatrix([np.zeros(4)])
    with pytest.raises(_lib.PasifError):
        p._setInterpolationMatrix(np.zeros(3), 3)
    # modulation buffer bounds (PASIf.py:389-392)
    with pytest.raises(ValueError):
        p.setModulationBuffer(np.zeros(0))
    with pytest.raises(ValueError):
        p.setModulationBuffer(np.zeros(32001))
    p.setModulationBuffer(np.ones(4))
    # bad precision / solver / tuning
    with pytest.raises(_lib.PasifError):
        p.setPrecision(16)
    with pytest.raises(KeyError):
        p.setSolver("nope")
    with pytest.raises(_lib.PasifError):
        p.setTuning(tile_lanes=100)
    # partial gradients are not supported; no jacobian set -> state error
    with pytest.raises(ValueError):
        p.getGradient()
    # trajectory stride 0
    with pytest.raises(_lib.PasifError):
        p._getTrajectory(0)


def test_rejects_systems_coupled_across_blocks():
    w = W.duffing(3, 10)
    p = _load(None, w)
    # a CSR entry coupling system 0 to system 1 with explicit sizes
    A, G, L, f, q0 = p.getMatrices()
    A = A.tocsr().tolil()
    A[0, 4] = 1.0
    A = A.tocsr()
    p._setFwdK(A.shape, A.data, A.indices, A.indptr)
    with pytest.raises(_lib.PasifError) as e:
        p._allocateSystemOnDevice()
    assert "outside its block" in str(e.value)


def test_autodetected_blocks_without_layout():
    # without sizes the driver finds the block-diagonal structure itself
    w = W.two_dof(10, 20)
    p = _load(None, w)
    p._clearSystemsLayout()
    p._allocateSystemOnDevice()
    rc, n, msg = _len(p, p._pl.pasif_get_amplitudes)
    assert rc == 0, msg
    assert n == 40 * len(w.excitations)   # Cartesian product without the map
