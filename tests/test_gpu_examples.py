"""The reference's own example systems through the drop-in ``PASIf`` on the
GPU, against the FP64 oracle (``tests/golden/reference_examples.npz``, from
``tests/golden/make_examples.py``):

* the 4-DOF probed oscillator and its 6-DOF adjoint
  (``examples/FullInterfaceTesting_1dof.py:43-82, 144-205``), three copies
  passed as ``[X] * 3`` (the aliasing idiom of the examples);
* the 8-DOF two-oscillator system and its 12-DOF adjoint with a Psi term
  (``examples/FullInterfaceTesting.py:48-76, 147-230``), four copies against
  three excitations (the reference's Cartesian product of systems and
  excitations).

Both run the examples' own 78,001 steps at 16 kHz with excitations of ones.
Final states, the trajectory every 1,000 steps and lambda_0 are compared per
output (``oracle.parity``, tolerance 1e-10); every copy and every
excitation must give the one-system result.  The inputs are not modified by
the calls."""
import copy
import os

import numpy as np
import pytest

from oracle import parity as PAR
from pasif_amd import PASIf
from tests import reference_examples as RX

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                    "reference_examples.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


@pytest.mark.parametrize("example,n,n_exc", [("one_dof", 3, 1), ("two_oscillators", 4, 3)])
def test_reference_example_through_pasif(gold, example, n, n_exc):
    print(f"[gpu case] reference example {example}: {n} aliased copies x {n_exc} "
          f"excitations x {RX.EXC_LEN} steps", flush=True)
    fwd, adj = getattr(RX, example)(n)
    before = copy.deepcopy((fwd, adj))
    p = PASIf(RX.excitations(n_exc), RX.SAMPLE_RATE, 0)
    assert p.numsteps == RX.EXC_LEN
    p.setSystems(*fwd)
    q = p.getAmplitudes()
    oq = gold[f"{example}_q"]
    ns = oq.size
    rep = q.reshape(n_exc * n, ns)
    ref = np.tile(oq, (n_exc * n, 1))
    r = PAR.report(rep, ref, np.full(n_exc * n, ns))
    print(f"  states {r.as_dict()}", flush=True)
    assert r.max_rel < PAR.TOL64
    save = int(gold["save_steps"])
    traj = p.getTrajectory(saveSteps=save)
    ot = gold[f"{example}_traj"]                       # [slots][dof]
    assert traj.shape == (n_exc * n * ns + 1, ot.shape[0])
    np.testing.assert_array_equal(traj[0], np.arange(ot.shape[0]) * save / RX.SAMPLE_RATE)
    states = traj[1:].T.reshape(ot.shape[0], n_exc * n, ns)
    rt = PAR.report(states, np.repeat(ot[:, None, :], n_exc * n, axis=1),
                    np.full(ot.shape[0] * n_exc * n, ns))
    assert rt.max_rel < PAR.TOL64, rt.as_dict()
    p.setJacobian(*adj)
    g = p.getGradient()
    og = gold[f"{example}_g"]
    na = og.size
    rg = PAR.report(g.reshape(n_exc * n, na), np.tile(og, (n_exc * n, 1)),
                    np.full(n_exc * n, na))
    print(f"  gradient {rg.as_dict()}", flush=True)
    assert rg.max_rel < PAR.TOL64
    # a second gradient needs no new setJacobian (the reference overwrites its
    # terminal condition, src/__GpuDriver.cu:743) and the inputs are unchanged
    assert np.array_equal(p.getGradient(), g)
    for a, b in zip(before[0] + before[1], fwd + adj):
        for x, y in zip(a, b):
            if hasattr(x, "val"):
                assert list(np.asarray(x.val).ravel()) == list(np.asarray(y.val).ravel())
                assert list(np.asarray(x.indices).ravel()) == list(np.asarray(y.indices).ravel())
            elif hasattr(x, "toarray"):
                assert np.array_equal(x.toarray(), y.toarray())
            else:
                assert np.array_equal(x, y)
