"""The parity metric of ``oracle.parity`` on hand-made cases: the per-system
near-zero floor, the argmax bookkeeping, and the one-ULP perturbation used
to measure the oracle's own sensitivity."""
import numpy as np
import scipy.sparse as sp

import oracle as O
from oracle import parity as PAR


def test_floor_is_per_system():
    # two systems of 2 outputs; system 1's second output sits near zero
    o = np.array([1.0, -2.0, 3.0, 1e-9])
    g = o.copy()
    g[3] += 1e-12                      # 1e-3 relative to the output itself
    sizes = [2, 2]
    e = PAR.rel_err(g, o, sizes)
    # floor of system 1: 1e-4 * 3 = 3e-4 -> error 1e-12 / 3e-4
    assert np.isclose(e[3], 1e-12 / 3e-4)
    assert np.all(e[:3] == 0)
    assert np.isclose(PAR.rel_err(g, o, sizes, floor=0.0)[3], 1e-3)
    r = PAR.report(g, o, sizes)
    assert (r.argmax_system, r.argmax_output) == (1, 1)
    assert (r.unfloored_system, r.unfloored_output) == (1, 1)
    assert np.isclose(r.unfloored_max, 1e-3) and r.system_scale == 3.0
    assert r.per_system.tolist() == [0.0, e[3]]
    # the floor of one system does not lift another's
    o2 = np.array([1e-6, 0.0, 5.0, 5.0])
    g2 = o2 + np.array([1e-16, 0.0, 0.0, 0.0])
    assert np.isclose(PAR.rel_err(g2, o2, [2, 2])[0], 1e-16 / 1e-6)


def test_shape_checks():
    import pytest
    with pytest.raises(ValueError):
        PAR.rel_err(np.zeros(3), np.zeros(4), [4])
    with pytest.raises(ValueError):
        PAR.rel_err(np.zeros(4), np.zeros(4), [3])


def test_ulp_perturbation_moves_every_coefficient_one_ulp():
    A = sp.csr_matrix(np.array([[1.0, 0.5], [0.25, -3.0]]))
    asm = O.Assembled(n=2, A=A, gamma_idx=np.array([[0, 1, 1]]), gamma_val=np.array([0.1]),
                      lam_idx=np.zeros((0, 4), dtype=np.int64), lam_val=np.zeros(0),
                      psi_idx=np.zeros((0, 5), dtype=np.int64), psi_val=np.zeros(0),
                      force=np.zeros(2), q0=np.zeros(2))
    for dt in (np.float64, np.float32):
        a = asm.astype(dt)
        b = PAR.ulp_perturbed(a, seed=3)
        for x, y in ((a.A.data, b.A.data), (a.gamma_val, b.gamma_val)):
            assert y.dtype == x.dtype
            assert np.all(np.abs(y - x) == np.spacing(np.abs(x)) * (np.abs(y) > np.abs(x)) +
                          np.spacing(np.abs(y)) * (np.abs(y) < np.abs(x)))
        assert np.array_equal(a.A.data, asm.astype(dt).A.data)   # input untouched
