"""Writes ``reference_examples.npz``: the FP64 oracle's outputs for the
reference's example systems (``tests/reference_examples.py``), one system
and one excitation each (every copy of an aliased list, and every excitation
of the examples, which are all ones(78001), has the same result):

    python tests/golden/make_examples.py          (about a minute)

Keys per example (``one_dof``, ``two_oscillators``): ``q`` final state,
``traj`` trajectory with saveSteps = 1000, ``g`` lambda_0 of the adjoint
system (checkpoint interval 256)."""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.dirname(HERE)):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle as O  # noqa: E402
import reference_examples as RX  # noqa: E402

OUT = os.path.join(HERE, "reference_examples.npz")
SAVE_STEPS = 1000


def run(example):
    fwd_in, adj_in = getattr(RX, example)(1)
    fwd = O.assemble_forward(*fwd_in)
    adj = O.assemble_adjoint(*adj_in, n_fwd=fwd.n)
    F = O.Forcing(np.stack(RX.excitations(1)))
    lay = O.cartesian_layout(fwd.n, 1, RX.EXC_LEN)
    r = O.rk4_forward(fwd, F, RX.SAMPLE_RATE, lay, save_steps=SAVE_STEPS)
    g = O.adjoint_gradient(fwd, adj, F, RX.SAMPLE_RATE, lay,
                           np.full((1, adj.n), RX.EXC_LEN), checkpoint_every=256)
    return {f"{example}_q": r.q[0], f"{example}_traj": r.traj[:, 0, :],
            f"{example}_g": g[0]}


def main():
    out = {}
    for ex in ("one_dof", "two_oscillators"):
        out.update(run(ex))
        print(ex, "done", flush=True)
    np.savez_compressed(OUT, save_steps=SAVE_STEPS, **out)


if __name__ == "__main__":
    main()
