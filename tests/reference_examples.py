"""The reference's own example systems, restated as data (inputs only; the
expected outputs come from the oracle, ``tests/golden/make_examples.py``).

* ``one_dof``: the 4-DOF probed oscillator and its 6-DOF adjoint system of
  ``examples/FullInterfaceTesting_1dof.py:43-82, 144-205`` (k = b = 1,
  excitation ones(78001) at 16 kHz, numsteps 0 -> 78,001 steps);
* ``two_oscillators``: the 8-DOF two-oscillator system and its 12-DOF
  adjoint of ``examples/FullInterfaceTesting.py:48-76, 147-230`` (the second
  Gamma list of lines 189-198 is the one the script keeps).

Lists are built the way the examples build them, ``[X] * n``: every entry is
the same object (the aliasing the reference's in-place concatenation turns
into self-concatenation, ``PASIf.py:584-586``)."""
from __future__ import annotations

import numpy as np
from scipy.sparse import coo_matrix, dia_matrix

from pasif_amd import coo_tensor

EXC_LEN = 78001
SAMPLE_RATE = 16000


def _tensor(dims, val, idx):
    t = coo_tensor(list(dims))
    t.val = list(val)
    t.indices = list(idx)
    return t


def one_dof(n: int):
    """(forward inputs, adjoint inputs), each a tuple of aliased lists."""
    s = 4
    M = dia_matrix(([1., 1., 1., 1.], [0]), shape=(s, s))
    B = coo_matrix(([-1., -1.], ([0, 1], [2, 3])), shape=(s, s))
    K = coo_matrix(([1., 1.], ([2, 2], [0, 2])), shape=(s, s))
    G = _tensor([s] * 3, [-1.], [3, 0, 0])
    L = _tensor([s] * 4, [0], [0, 0, 0, 0])
    f = np.array([0., 0., 1., 0.])
    q0 = np.zeros(s)
    fwd = ([M] * n, [B] * n, [K] * n, [G] * n, [L] * n, [f] * n, [q0] * n)
    a = 6
    Ma = dia_matrix((np.ones(a), [0]), shape=(a, a))
    Ba = coo_matrix(([-1, -1], ([1, 2], [4, 5])), shape=(a, a))
    Ka = coo_matrix(([], ([], [])), shape=(a, a))
    Ga = _tensor([a, a, s], [], [])
    La = _tensor([a, a, s, s], [-1, -1], [4, 0, 0, 0, 5, 0, 1, 1])
    fa = np.zeros(a)
    lT = np.zeros(a)
    lT[0] = 1.
    Pa = _tensor([a, a, s, s, s], [0], [1, 1, 1, 1, 1])
    adj = ([Ma] * n, [Ba] * n, [Ka] * n, [Ga] * n, [La] * n, [fa] * n, [lT] * n, [Pa] * n)
    return fwd, adj


def two_oscillators(n: int):
    s = 8
    M = dia_matrix(([1., 1., 1., 1., 1., 10., 1., 1.], [0]), shape=(s, s))
    B = coo_matrix(([-1, -1, -1, -1, 1, 10], ([0, 1, 2, 3, 4, 5], [4, 5, 6, 7, 4, 5])),
                   shape=(s, s))
    K = coo_matrix(([6, 10], ([4, 5], [0, 1])), shape=(s, s))
    G = _tensor([s] * 3, [-10, -10, -1, -1], [4, 0, 1, 5, 0, 0, 6, 0, 0, 7, 1, 1])
    L = _tensor([s] * 4, [40000], [5, 1, 1, 1])
    f = np.array([0, 0, 0, 0, 0.5, 0, 0, 0])
    q0 = np.zeros(s)
    fwd = ([M] * n, [B] * n, [K] * n, [G] * n, [L] * n, [f] * n, [q0] * n)
    a = 12
    Ma = dia_matrix((np.ones(a), [0]), shape=(a, a))
    Bi = [4, 0, 5, 1, 6, 2, 7, 3]
    Ba = coo_matrix(([1, 1, 1, 1], (Bi[::2], Bi[1::2])), shape=(a, a))
    Ki = [0, 4, 4, 4, 1, 5, 5, 5]
    Ka = coo_matrix(([-12.0, -2.0, -1, -1.0], (Ki[::2], Ki[1::2])), shape=(a, a))
    Ga = _tensor([a, a, s], [20.0, 2.0, 1, 20, 1, -2.0, -1.0, -0.2, -0.1],
                 [0, 4, 1, 0, 5, 0, 0, 6, 0, 1, 4, 0, 1, 7, 1, 8, 4, 0, 9, 4, 4, 10, 5, 1,
                  11, 5, 5])
    La = _tensor([a, a, s, s], [-120000.0], [1, 5, 1, 1])
    fa = np.zeros(a)
    lT = np.zeros(a)
    lT[2] = 1.
    lT[3] = 1.
    Pa = _tensor([a, a, s, s, s], [1], [1, 1, 1, 1, 1])
    adj = ([Ma] * n, [Ba] * n, [Ka] * n, [Ga] * n, [La] * n, [fa] * n, [lT] * n, [Pa] * n)
    return fwd, adj


def excitations(n_exc: int):
    e = np.ones(EXC_LEN)
    return [e] * n_exc
