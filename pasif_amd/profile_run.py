"""Profiling driver: one solve of a benchmark workload (plus, optionally, a
device copy of known size used to calibrate HBM byte counters).

    python -m pasif_amd.profile_run --workload two_dof --systems 4096 --steps 50000
    python -m pasif_amd.profile_run --calibrate-copy-mib 1024

Run it under ``rocprofv3 --kernel-trace --stats`` for durations or
``rocprofv3 --pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` for bytes.  The copy
moves exactly N MiB (read N, write N) so the counters' reading for it gives
the scale factor applied to the tile kernel's reading.
"""
from __future__ import annotations

import argparse
import json
import sys
import time


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="two_dof",
                    choices=["two_dof", "duffing", "chains", "heterogeneous"])
    ap.add_argument("--systems", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=50_000)
    ap.add_argument("--dof", type=int, default=16)
    ap.add_argument("--adjoint", action="store_true")
    ap.add_argument("--precision", type=int, default=64)
    ap.add_argument("--solver", default="persistent")
    ap.add_argument("--kernel", default="auto", choices=["auto", "tile", "lane", "big"])
    ap.add_argument("--adjoint-mode", default="reference", choices=["reference", "exact"])
    ap.add_argument("--variants", type=int, default=1,
                    help="heterogeneous: sparsity variants per DOF size")
    ap.add_argument("--lanes", type=int, default=0,
                    help="lanes per system in the lane kernel (0: chosen per structure)")
    ap.add_argument("--checkpoint-every", type=int, default=0,
                    help="adjoint checkpoint interval (0: the planner's choice)")
    ap.add_argument("--repeat", type=int, default=1)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0)
    a = ap.parse_args(argv)

    if a.calibrate_copy_mib:
        import torch
        n = a.calibrate_copy_mib * (1 << 20) // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8}), flush=True)
        if a.steps == 0:
            return

    from . import workloads as W
    if a.workload == "two_dof":
        w = W.two_dof(n_systems=a.systems, steps=a.steps)
    elif a.workload == "duffing":
        w = W.duffing(n_systems=a.systems, steps=a.steps)
    elif a.workload == "chains":
        w = W.chains(n_systems=a.systems, d=a.dof, steps=a.steps)
    else:
        w = W.heterogeneous(n_systems=a.systems, tmax=a.steps, variants_per_size=a.variants)
    p = W.make_solver(w)
    p.setPrecision(a.precision)
    p.setSolver(a.solver)
    p.setKernel(a.kernel, a.lanes)
    if a.adjoint:
        from .adjoint import energy_gradient_blocks
        p.setJacobianBlocks(energy_gradient_blocks(w))
        p.setAdjointMode(a.adjoint_mode)
    if a.checkpoint_every:
        p.setTuning(checkpoint_every=a.checkpoint_every)
    for _ in range(a.repeat):
        t0 = time.perf_counter()
        if a.adjoint:
            p.getGradient()
        else:
            p.getAmplitudes()
        wall = time.perf_counter() - t0
        t = p._lastTiming()
        print(json.dumps({"workload": w.name, "systems": w.n_systems, "steps": w.numsteps,
                          "adjoint": a.adjoint, "wall_s": wall, **t,
                          **p.kernelInfo(adjoint=a.adjoint)}), flush=True)


if __name__ == "__main__":
    sys.exit(main())
