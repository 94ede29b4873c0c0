"""The committed full-size fixtures are what their generators compute: a few
systems of each recomputed from the seeded workloads with the oracle and
compared bit for bit (the oracle is deterministic and its per-system
arithmetic does not depend on which other systems share the call)."""
import os

import numpy as np
import pytest

from tests.golden import make_fullsize as MF
from tests.golden import make_structures as MS

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not generated")
    return np.load(path)


def test_two_dof_fixture_first_systems():
    G = _load("fullsize_two_dof.npz")
    q = MF.two_dof_shard((0, 2, np.float64, False))
    np.testing.assert_array_equal(q, G["o"][: q.size])


def test_chains_fixture_first_chain():
    G = _load("fullsize_chains.npz")
    ids = np.asarray(G["ids"])
    oq, og = MF.chains_shard((ids[:1], False))
    np.testing.assert_array_equal(oq, G["oq"][: oq.size])
    np.testing.assert_array_equal(og, G["og"][: og.size])


def _first_shortest(w, ids):
    steps = np.asarray(w.steps_of_system)[ids]
    return int(np.argmin(steps))


def test_heterogeneous_fixture_shortest_system():
    G = _load("fullsize_heterogeneous.npz")
    ids = np.asarray(G["ids"])
    w = MF._heterogeneous()
    k = _first_shortest(w, ids)
    q, g = MF.heterogeneous_shard((int(ids[k]), np.float64))
    nf = np.array([w.vecM[i].shape[0] for i in ids])
    off = np.concatenate([[0], np.cumsum(nf)])
    np.testing.assert_array_equal(q, G["q64"][off[k]:off[k + 1]])
    assert g.size > q.size


def test_structures_fixture_shortest_system():
    G = _load("fullsize_structures.npz")
    ids = np.asarray(G["ids"])
    w = MS.batch()
    np.testing.assert_array_equal(np.asarray(w.meta["variants"])[ids], G["variants"])
    k = _first_shortest(w, ids)
    q, g = MS.oracle_system(w, int(ids[k]))
    nf = np.array([w.vecM[i].shape[0] for i in ids])
    off = np.concatenate([[0], np.cumsum(nf)])
    np.testing.assert_array_equal(q, G["q64"][off[k]:off[k + 1]])
