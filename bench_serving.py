"""Serving benchmark: Poisson streams of solve requests against the
continuous-batching pool on one MI355X.

Each request is a batch of 1-256 systems drawn from the heterogeneous
distributions (DOF {1,..,64}, steps log-uniform on [1e3, 1e5], dt 1e-3,
chain structure, interpolated band-limited forcing); 30 % forward-only, 70 %
forward+adjoint (energy gradient), FP64, admitted against a 200 GB HBM
budget and a resident-tile (slot) cap; library systems are registered with
the pool once and requests travel as system ids.  With ``--fresh-frac f`` a
fraction f of the requests instead carry their own arrays: systems with
coefficients drawn afresh and one of 66 sparsity structures (12 variants per
DOF size), none registered, run by the pool's tile kernel beside the lane
modules.  For every offered load it prints one JSON line: p50 / p99 / mean
latency (arrival to results on the host) of completed requests, the backlog
left at the end with p99 over all arrived requests (waiting time so far for
the unfinished ones, a lower bound), sustained system-steps/s and GPU busy
fraction; and once, the plain one-launch-per-step kernel run per request on
a sample of requests for comparison.

    python bench_serving.py --loads 1000 5000 20000 --duration 60
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--loads", type=float, nargs="+", default=[50, 100, 1000, 5000, 20000])
    ap.add_argument("--duration", type=float, default=10.0)
    ap.add_argument("--max-wall", type=float, default=120.0)
    ap.add_argument("--tmin", type=float, default=1e3)
    ap.add_argument("--tmax", type=float, default=1e5)
    ap.add_argument("--max-systems", type=int, default=256)
    ap.add_argument("--steps-per-tick", type=int, default=5000)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--lanes", type=int, default=512)
    ap.add_argument("--slots", type=int, default=32768,
                    help="resident tile cap (0: HBM budget only)")
    ap.add_argument("--plain-sample", type=int, default=3)
    ap.add_argument("--fresh-frac", type=float, default=0.0,
                    help="fraction of requests carrying fresh, unregistered systems")
    ap.add_argument("--fresh-systems", type=int, default=2048)
    ap.add_argument("--fresh-structures", choices=["new", "registered"], default="new",
                    help="fresh requests' sparsity: structures the pool never registered "
                         "(tiles) or the registered ones with coefficients drawn afresh "
                         "(lane modules)")
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="hardware queues HIP gives this process (GPU_MAX_HW_QUEUES, at most "
                         "32; 0: HIP's default of 4): each pool lane module is one launch "
                         "per tick on a stream of its own, and launches beyond the queues "
                         "run one after another")
    ap.add_argument("--group-systems", type=int, default=4096,
                    help="most systems in one admitted group")
    ap.add_argument("--compile-new", action="store_true",
                    help="compile structures no lane module runs into further modules in "
                         "the background (pasif_pool_set_compile)")
    ap.add_argument("--warm-rerun", action="store_true",
                    help="with --compile-new: a second window on the same pool")
    ap.add_argument("--no-group", action="store_true",
                    help="admit one pool job per request instead of one per admission "
                         "batch and adjoint flag")
    a = ap.parse_args()
    if a.hw_queues:
        # read when HIP initialises: set before the library is loaded
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, max(1, a.hw_queues)))

    from pasif_amd import serving as S
    from pasif_amd.workloads import linear_interp_matrix
    rng = np.random.default_rng(a.seed)
    lib = S.SystemLibrary(variants=4)
    n_signals = 32
    bank = S.signal_bank(rng, n_signals, a.tmax, steps_per_sample=10)
    interp = linear_interp_matrix(10)
    fresh = (S.FreshSystems(a.fresh_systems, like=lib if a.fresh_structures == "registered"
                            else None) if a.fresh_frac > 0 else None)

    def new_pool(steps_per_tick):
        pool = S.Pool(lanes=a.lanes, slot_config=1, checkpoint_every=256,
                      steps_per_tick=steps_per_tick, hbm_budget=a.budget_gb * 1e9)
        pool.set_bank(bank, 1000, interp)
        pool.register_library(lib)
        pool.set_slots(a.slots)
        if a.compile_new:
            pool.set_compile(True)
        return pool

    def window(pool, load, li, state):
        arr = S.poisson_arrivals(rng, load, a.duration)
        reqs = S.RequestStream(lib, len(arr), a.seed * 1000 + li, n_signals, tmin=a.tmin,
                               tmax=a.tmax, max_systems=a.max_systems)
        if fresh is not None:
            reqs = S.MixedRequests(reqs, S.FreshRequestStream(
                fresh, len(arr), a.seed * 1000 + li + 500, n_signals, tmin=a.tmin, tmax=a.tmax,
                max_systems=a.max_systems), a.fresh_frac, seed=a.seed + li)
        rep = S.SolveServer(pool, group=not a.no_group, group_systems=a.group_systems).run(
            reqs, horizon_s=a.duration, drain=False, max_wall_s=a.max_wall, arrivals=arr)
        end_stats = pool.stats()
        if a.compile_new:
            t0 = time.perf_counter()
            while True:
                built, failed, queued = pool.compile_wait(1)
                print(f"[serving] modules compiled {built}, failed {failed}, "
                      f"structures queued {queued}", file=sys.stderr, flush=True)
                if not queued:
                    break
            end_stats["modules_built"], end_stats["builds_failed"] = built, failed
            end_stats["compile_wait_s"] = time.perf_counter() - t0
        refusals = dict(pool.refusals)
        pool.refusals.clear()
        print(json.dumps({"metric": "serving latency", "offered_rps": load,
                          "duration_s": a.duration, "arrived": rep.arrived,
                          "admitted": rep.admitted, "completed": rep.completed,
                          "backlog": rep.backlog,
                          "p50_ms": rep.p50_ms, "p99_ms": rep.p99_ms, "mean_ms": rep.mean_ms,
                          "p99_all_ms": rep.p99_all_ms, "admit_ms_mean": rep.admit_ms_mean,
                          "system_steps_per_s": rep.system_steps_per_s, "ticks": rep.ticks,
                          "gpu_busy_frac": rep.gpu_busy_frac, "wall_s": rep.duration_s,
                          "fwd_adj_fraction": 0.7, "fresh_fraction": a.fresh_frac,
                          "fresh_structures": a.fresh_structures if a.fresh_frac > 0 else None,
                          "compile_new": state,
                          "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
                          "grouped": not a.no_group, "group_systems": a.group_systems,
                          "pool_at_stop": end_stats,
                          "refused_groups": refusals,
                          "dtype": "f64"}), flush=True)

    for li, load in enumerate(a.loads):
        pool = new_pool(a.steps_per_tick)
        window(pool, load, li, "cold start" if a.compile_new else None)
        if a.compile_new and a.warm_rerun:
            # the same pool again: the structures of the first window compiled
            window(pool, load, li + 100, "after the first window")
        pool.close()

    # the plain kernel run per request: one launch per step, each request alone
    rates = []
    for k in range(a.plain_sample):
        r = S.draw_request(lib, rng, 10**6 + k, n_signals, tmin=a.tmin, tmax=a.tmin * 3,
                           max_systems=a.max_systems)
        pool = new_pool(1)
        job = pool.submit_ids(r)
        t0 = time.perf_counter()
        while not pool.finished(job):
            pool.tick(wait=True)
        wall = time.perf_counter() - t0
        pool.collect(job, r)
        pool.close()
        rates.append(r.system_steps / wall)
    if rates:
        print(json.dumps({"metric": "plain kernel per request", "requests": a.plain_sample,
                          "system_steps_per_s_mean": float(np.mean(rates)),
                          "note": "one launch per step (steps_per_tick=1), request alone"}),
              flush=True)


if __name__ == "__main__":
    main()
