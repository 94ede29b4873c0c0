"""pasif_solve_noalloc (the no-allocation solve) through the C ABI against
the allocating entry points it mirrors: pasif_get_amplitudes (forward) and
pasif_get_gradient (forward + adjoint).  Both run the same kernels, so the
results must be bitwise equal.  Two problems in a row share the library's
static scratch array, the second smaller than the first; a problem whose run
buffers exceed the array is refused before anything runs; a call without
the set-up is refused."""
import ctypes as C

import numpy as np
import pytest

from pasif_amd import _lib
from pasif_amd import adjoint as ADJ
from pasif_amd import workloads as W

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def _case(name):
    print(f"[gpu case] {name}", flush=True)


def _lengths(p, adjoint):
    """The output-length queries: they pack the problem and upload it."""
    lib, n = p._pl, C.c_size_t(0)
    _lib.check(lib, p._h, lib.pasif_get_amplitudes(p._h, None, 0, C.byref(n)))
    nq = n.value
    ng = 0
    if adjoint:
        _lib.check(lib, p._h, lib.pasif_get_gradient(p._h, 0, None, 0, C.byref(n)))
        ng = n.value
    return nq, ng


def _hip():
    """The HIP runtime the library itself is linked against (already loaded
    with it): device buffers for the results come from the caller."""
    import os
    _lib.load()
    hip = C.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD)
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


class _DevBuf:
    """A device buffer of n values, filled with NaN (host <-> device copies
    through hipMemcpy)."""
    H2D, D2H = 1, 2

    def __init__(self, hip, n, dtype):
        self.hip, self.dtype, self.n = hip, np.dtype(dtype), max(int(n), 1)
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), self.n * self.dtype.itemsize) == 0
        nan = np.full(self.n, np.nan, dtype=self.dtype)
        assert hip.hipMemcpy(self.ptr, nan.ctypes.data, nan.nbytes, self.H2D) == 0

    def read(self, n):
        out = np.empty(self.n, dtype=self.dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, self.D2H) == 0
        return out[:n].astype(np.float64)

    def free(self):
        self.hip.hipFree(self.ptr)


def _noalloc(p, adjoint, nq, ng, bits=64):
    hip = _hip()
    dt = np.float64 if bits == 64 else np.float32
    q, g = _DevBuf(hip, nq, dt), _DevBuf(hip, ng, dt)
    try:
        rc = p._pl.pasif_solve_noalloc(p._h, int(adjoint), q.ptr, g.ptr if adjoint else None)
        _lib.check(p._pl, p._h, rc)
        return q.read(nq), (g.read(ng) if adjoint else None)
    finally:
        q.free()
        g.free()


def _chains(n, d, steps):
    w = W.chains(n_systems=n, d=d, steps=steps, steps_per_sample=10)
    p = W.make_solver(w)
    p.setJacobian(*ADJ.energy_gradient_inputs(w).as_args())
    p.setTuning(checkpoint_every=64)
    return p


@pytest.mark.parametrize("kernel", ["auto", "tile"])
def test_noalloc_equals_allocating_two_calls_second_smaller(kernel):
    for n, d, steps in ((48, 4, 700), (8, 2, 300)):
        _case(f"noalloc {kernel}: chains {n} x {d} DOF x {steps} steps")
        p = _chains(n, d, steps)
        p.setKernel(kernel)
        q_ref = p.getAmplitudes()
        g_ref = p.getGradient()
        # forward: set-up by the forward length query, then the no-alloc solve
        nq, _ = _lengths(p, False)
        q, _ = _noalloc(p, False, nq, 0)
        assert np.array_equal(q, q_ref)
        # forward + adjoint, twice (the second call reuses the scratch)
        nq, ng = _lengths(p, True)
        for _ in range(2):
            q, g = _noalloc(p, True, nq, ng)
            assert np.array_equal(g, g_ref)
            assert np.array_equal(q, q_ref)


def test_noalloc_fp32_equals_allocating():
    _case("noalloc FP32 two-DOF 64 x 2000")
    p = W.make_solver(W.two_dof(n_systems=64, steps=2000))
    p.setPrecision(32)
    q_ref = p.getAmplitudes()
    nq, _ = _lengths(p, False)
    q, _ = _noalloc(p, False, nq, 0, bits=32)
    assert np.array_equal(q, q_ref)


def test_noalloc_refuses_without_setup_and_beyond_the_scratch():
    _case("noalloc refusals")
    p = W.make_solver(W.two_dof(n_systems=16, steps=200))
    with pytest.raises(_lib.PasifError) as e:
        _noalloc(p, False, 64, 0)
    assert e.value.code == 5
    # 256 chains of 8 DOF on tiles, checkpoint every 16 of 50,000 steps:
    # the checkpoints alone exceed the 64 MiB array; refused before any launch
    p = _chains(256, 8, 50_000)
    p.setKernel("tile")
    p.setTuning(checkpoint_every=16)
    nq, ng = _lengths(p, True)
    with pytest.raises(_lib.PasifError) as e:
        _noalloc(p, True, nq, ng)
    assert e.value.code == 4 and "run buffers" in str(e.value)
