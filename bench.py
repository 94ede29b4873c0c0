"""Benchmark of the batched RK4 hot path on one MI355X.

``python bench.py [--gpus 1] [--steps K] [--warmup W]`` prints ONE JSON line.

Workload (one bench "step" = one full solve): 4,096 two-DOF systems
``[x0, x1, v0, v1]`` with ``B = 0.05 I``, ``K = [[2,-1],[-1,2]] s``,
3 Gamma and 4 Lambda couplings, sinusoidal forcing on DOF 0 (one excitation
per system), ``dt = 5e-4``, 50,000 RK4 steps, FP64 -- 2.048e8 system-steps
per solve, inputs resident in HBM before the timed region.

``value`` is system-steps/s of the persistent solver (all steps in one
launch).  Also reported: the plain per-step kernel on the same GPU, the FP32
rate, the roofline of the tile kernel (algorithmic FP64 flops per launch over
the HIP-event duration of the launch), the FP64 error against the NumPy
oracle on the CPU sample, the oracle's own rate on the host cores, and
(``fwd_adj``) forward + checkpointed adjoint on 8,192 chains of 16 DOF x
20,000 steps, persistent vs plain.
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

PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 (vector and matrix)
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def flops_per_step(p) -> float:
    """Algorithmic FP64 flops of one RK4 step of every system (FMA = 2):
    ``4 (2 nnzA + 3 nnzG + 4 nnzL + 5 nnzP + 2 nnzF) + 12 n`` summed over
    systems; ``nnzF`` counts forced DOFs."""
    A, G, L, f, q0 = p.getMatrices()
    nnzA = A.tocsr().nnz
    nnzG = len(np.asarray(G.val).reshape(-1))
    nnzL = len(np.asarray(L.val).reshape(-1))
    nnzF = int(np.count_nonzero(np.abs(f) > np.finfo(np.float64).eps))
    n = len(q0)
    return 4.0 * (2 * nnzA + 3 * nnzG + 4 * nnzL + 2 * nnzF) + 12.0 * n


def bytes_per_launch(p, w) -> float:
    """Algorithmic HBM bytes of one persistent launch: coefficients and
    indices (8+4 B per term), initial state, final state, and the excitation
    samples read (each sample of a forced system once: steps + 1)."""
    A, G, L, f, q0 = p.getMatrices()
    nnz = A.tocsr().nnz + len(np.asarray(G.val).reshape(-1)) + len(np.asarray(L.val).reshape(-1))
    n = len(q0)
    forced = int(np.count_nonzero(np.abs(f) > np.finfo(np.float64).eps))
    return 12.0 * nnz + 8.0 * n + 8.0 * n + 8.0 * forced * (w.numsteps + 1)


def read_traffic(name: str):
    """HBM bytes per launch measured with rocprofv3 (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration), if the profile summary is
    committed under profiles/."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].split()[0])
            except ValueError:
                return None
    return None


def cpu_baseline(w, n_sample: int):
    """The FP64 NumPy oracle on the host: the first ``n_sample`` systems over
    all steps (vectorised over systems, one thread)."""
    import oracle as O
    from pasif_amd.workloads import Workload
    sub = Workload(w.name, w.vecM[:n_sample], w.vecB[:n_sample], w.vecK[:n_sample],
                   w.vecGamma[:n_sample], w.vecLambda[:n_sample], w.vecForce[:n_sample],
                   w.vecQ0[:n_sample], w.excitations[:n_sample], w.sample_rate, w.numsteps,
                   exc_of_system=np.arange(n_sample), interp=w.interp)
    asm = O.assemble_forward(sub.vecM, sub.vecB, sub.vecK, sub.vecGamma, sub.vecLambda,
                             sub.vecForce, sub.vecQ0)
    lay = O.mapped_layout(asm.sizes, sub.exc_of_system, np.full(n_sample, w.numsteps))
    t0 = time.perf_counter()
    res = O.rk4_forward(asm, O.Forcing(sub.excitations), w.sample_rate, lay)
    dt = time.perf_counter() - t0
    return res.q.reshape(-1), n_sample * w.numsteps / dt, dt


def adjoint_flops_per_step(p) -> float:
    """Algorithmic flops of one adjoint step of every system:
    ``4 (2 nnzA' + 3 nnzG' + 4 nnzL' + 5 nnzPsi') + 12 n_adj + 2 n_fwd``."""
    import scipy.sparse as sp
    Ja = (p.jacobian_B + p.jacobian_K).tocsr()
    nG = len(np.asarray(p.jacobian_Gamma.val).reshape(-1))
    nL = len(np.asarray(p.jacobian_Lambda.val).reshape(-1))
    nP = len(np.asarray(p.jacobian_Psi.val).reshape(-1))
    return 4.0 * (2 * Ja.nnz + 3 * nG + 4 * nL + 5 * nP) + 12.0 * p.globalAdjointSize + \
        2.0 * p.globalSystemSize


def _progress(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _timed(p, fn):
    t0 = time.perf_counter()
    out = fn()
    return out, time.perf_counter() - t0, p._lastTiming()


def fwd_adj_section(args) -> dict:
    """Forward + checkpointed adjoint on chains of 16 DOF (energy gradient
    w.r.t. the K diagonal and the Lambda coefficients), FP64, checkpoint
    every 256 steps requested: the persistent launch of the kernel the
    library picks, against the plain one-launch-per-step solver of the same
    kernel and of the tile kernel.  A system-step is one forward step plus
    one adjoint step of one system."""
    from pasif_amd import adjoint as ADJ
    from pasif_amd import workloads as W
    w = W.chains(n_systems=args.adj_systems, d=16, steps=args.adj_steps, steps_per_sample=10)
    p = W.make_solver(w)
    p.setJacobian(*ADJ.energy_gradient_inputs(w).as_args())
    p.setTuning(checkpoint_every=256)
    info = p.kernelInfo(adjoint=True)
    p.getGradient()                                   # warm-up (packing, compilation, upload)
    g, wall, t = _timed(p, p.getGradient)
    sys_steps = w.n_systems * w.numsteps
    # work in the launch: forward twice (checkpoint pass + recompute) and the adjoint
    flops = (2 * flops_per_step(p) + adjoint_flops_per_step(p)) * w.numsteps
    achieved = flops / (t["kernel_ms"] * 1e-3) / 1e12
    sec = {"workload": f"chains {w.n_systems} x 16 DOF x {w.numsteps} steps, dt 1e-3, "
                       f"interpolated band-limited forcing, gradient of E(T) w.r.t. 16 K_ii + "
                       f"16 Lambda per system, FP64",
           "kernel": info["kernel"], "checkpoint_every": info["checkpoint_every"],
           "lanes_per_system": info.get("lanes_per_instance"),
           "value": sys_steps / wall, "unit": "system-steps/s (fwd+adj)",
           "kernel_ms": t["kernel_ms"], "workgroups": t["tiles"],
           "roofline": {"achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP64_TFLOPS,
                        "flops_per_launch": flops}}
    if not args.no_plain:
        p.setSolver("plain")
        gp, pw, pt = _timed(p, p.getGradient)
        sec["plain"] = {"kernel": info["kernel"], "value": sys_steps / pw,
                        "kernel_ms": pt["kernel_ms"], "launches": pt["launches"],
                        "bitwise_equal_to_persistent": bool(np.array_equal(gp, g))}
        sec["persistent_vs_plain"] = pw / wall
        if info["kernel"] != "tile":
            p.setKernel("tile")
            p.setSolver("persistent")
            gt, tw, tt = _timed(p, p.getGradient)
            sec["tile_persistent"] = {"value": sys_steps / tw, "kernel_ms": tt["kernel_ms"],
                                      "max_rel_diff": float(np.max(np.abs(gt - g) / np.maximum(
                                          np.abs(g), 1e-8 * np.abs(g).max())))}
            p.setSolver("plain")
            _, tpw, tpt = _timed(p, p.getGradient)
            sec["plain_tile"] = {"value": sys_steps / tpw, "kernel_ms": tpt["kernel_ms"],
                                 "launches": tpt["launches"]}
            sec["persistent_vs_plain_tile"] = tpw / wall
            p.setKernel("auto")
        p.setSolver("persistent")
    if not args.no_cpu:
        sec["parity"] = fwd_adj_parity(w.numsteps)
    return sec


def fwd_adj_parity(steps: int, n: int = 8) -> dict:
    """FP64 error of the GPU forward states and gradients against the NumPy
    oracle on ``n`` chains of the same shape over the full step count."""
    import oracle as O
    from pasif_amd import adjoint as ADJ
    from pasif_amd import workloads as W
    w = W.chains(n_systems=n, d=16, steps=steps, steps_per_sample=10)
    J = ADJ.energy_gradient_inputs(w)
    p = W.make_solver(w)
    p.setJacobian(*J.as_args())
    p.setTuning(checkpoint_every=256)
    q, g = p.getAmplitudes(), p.getGradient()
    t0 = time.perf_counter()
    fwd = O.assemble_forward(w.vecM, w.vecB, w.vecK, w.vecGamma, w.vecLambda, w.vecForce,
                             w.vecQ0)
    adj = O.assemble_adjoint(*J.as_args(), n_fwd=fwd.n)
    lay = O.mapped_layout(fwd.sizes, w.exc_of_system, np.full(n, steps))
    F = O.Forcing(w.excitations, w.interp)
    oq = O.rk4_forward(fwd, F, w.sample_rate, lay).q[0]
    og = O.adjoint_gradient(fwd, adj, F, w.sample_rate, lay, np.full((1, adj.n), steps),
                            checkpoint_every=256)[0]

    def rel(a, o):
        return float(np.max(np.abs(a - o) / np.maximum(np.abs(o), 1e-8 * np.abs(o).max())))
    return {"sample": f"{n} chains x 16 DOF x {steps} steps (oracle {time.perf_counter() - t0:.1f} s)",
            "max_rel_err_q": rel(q, oq), "max_rel_err_grad": rel(g, og),
            "tolerance_fp64": 1e-10}


def batching_section(args) -> dict:
    """Heterogeneous batch, persistent: 100,000 systems with DOF
    drawn from {1,..,64} ({0.3,0.25,0.15,0.1,0.1,0.06,0.04}), steps
    log-uniform on [1e3, 1e5], dt 1e-3, chain structure at each size,
    interpolated band-limited forcing.  Systems that fit the lane kernel run
    on it, the rest on the tile kernel, each family in one persistent launch
    on its own stream; tile shape, lane group sizes (and checkpoint interval)
    autotuned per precision on probes with every system's steps scaled so
    the longest runs 5,000.  FP64 and FP32 forward, FP64
    forward+adjoint (energy gradient).  Host-side workload generation is not
    timed."""
    from pasif_amd import adjoint as ADJ
    from pasif_amd import autotune as AT
    from pasif_amd import workloads as W
    t0 = time.perf_counter()
    w = W.heterogeneous(n_systems=args.batch_systems, tmin=1e3, tmax=1e5, steps_per_sample=10)
    p = W.make_solver(w)
    setup_s = time.perf_counter() - t0
    _progress(f"batching: workload built in {setup_s:.0f} s")
    sys_steps = int(np.sum(w.steps_of_system))
    info = p.kernelInfo()
    tune = AT.autotune(p, probe_steps=5000, lanes=(128, 256), slots=(3, 1), use_cache=False)
    _progress(f"batching: forward tuned {tune}")
    p.getAmplitudes()
    _, wall, t = _timed(p, p.getAmplitudes)
    flops = flops_per_step(p) * (sys_steps / w.n_systems)  # mean steps per system
    sec = {"workload": f"heterogeneous {w.n_systems} systems, DOF 1..64, steps log-uniform "
                       f"[1e3, 1e5] (total {sys_steps} system-steps), dt 1e-3",
           "kernel": info["kernel"], "tuning": tune, "setup_s": setup_s,
           "fwd_fp64": {"value": sys_steps / wall, "unit": "system-steps/s",
                        "kernel_ms": t["kernel_ms"], "tiles": t["tiles"],
                        "tflops": flops / (t["kernel_ms"] * 1e-3) / 1e12}}
    p.setPrecision(32)
    tune32 = AT.autotune(p, probe_steps=5000, lanes=(128, 256), slots=(3, 1), use_cache=False)
    _progress(f"batching: fp32 forward tuned {tune32}")
    p.getAmplitudes()
    _, w32, t32 = _timed(p, p.getAmplitudes)
    sec["fwd_fp32"] = {"value": sys_steps / w32, "kernel_ms": t32["kernel_ms"], "tuning": tune32}
    _progress(f"batching: forward fp64 {t['kernel_ms']:.0f} ms, fp32 {t32['kernel_ms']:.0f} ms")
    p.setPrecision(64)
    t1 = time.perf_counter()
    p.setJacobianBlocks(ADJ.energy_gradient_blocks(w))
    sec["setup_s"] += time.perf_counter() - t1
    _progress("batching: adjoint inputs built")
    tune_a = AT.autotune(p, adjoint=True, probe_steps=5000, lanes=(128, 256), slots=(3, 1),
                         checkpoints=(128, 256), use_cache=False)
    _progress(f"batching: adjoint tuned {tune_a}")
    p.getGradient()
    _, wa, ta = _timed(p, p.getGradient)
    sec["fwd_adj_fp64"] = {"value": sys_steps / wa, "unit": "system-steps/s (fwd+adj)",
                           "kernel_ms": ta["kernel_ms"], "tuning": tune_a}
    return sec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--systems", type=int, default=4096)
    ap.add_argument("--rk-steps", type=int, default=50_000)
    ap.add_argument("--cpu-sample", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adj-systems", type=int, default=8192)
    ap.add_argument("--adj-steps", type=int, default=20_000)
    ap.add_argument("--batching", action="store_true",
                    help="run only the heterogeneous-batch measurement (100,000 systems)")
    ap.add_argument("--batch-systems", type=int, default=100_000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: systems are independent and the path needs no "
              f"communication; --gpus {args.gpus} is not supported (each GPU would run its "
              f"own share of systems).", file=sys.stderr)
        sys.exit(2)

    from pasif_amd import workloads as W

    if args.batching:
        print(json.dumps({"metric": "heterogeneous batch, persistent (one launch per kernel "
                                    "family, side by side)",
                          **batching_section(args)}), flush=True)
        return

    w = W.two_dof(n_systems=args.systems, steps=args.rk_steps)
    p = W.make_solver(w)
    sys_steps = w.n_systems * w.numsteps

    for _ in range(args.warmup):
        p.getAmplitudes()
    t0 = time.perf_counter()
    kernel_ms = []
    for _ in range(args.steps):
        q = p.getAmplitudes()
        kernel_ms.append(p._lastTiming()["kernel_ms"])
    wall = (time.perf_counter() - t0) / args.steps
    timing = p._lastTiming()
    value = sys_steps / wall
    kms = float(np.mean(kernel_ms))
    kinfo = p.kernelInfo()
    kname = f"{kinfo['kernel']}_kernel"

    flops = flops_per_step(p) * w.numsteps
    achieved_tf = flops / (kms * 1e-3) / 1e12
    roofline = {
        "bound": "mfma",
        "pipe": "FP64 VALU (no matrix op on this path; MI355X FP64 vector peak)",
        "kernel": kname,
        "achieved": round(achieved_tf, 4),
        "peak": PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved_tf / PEAK_FP64_TFLOPS, 5),
        "traffic": read_traffic(f"{kname}_two_dof_4096x50000_traffic.txt"),
        "algorithmic_bytes": bytes_per_launch(p, w),
        "flops_per_launch": flops,
        "kernel_ms": round(kms, 4),
        "tiles": timing["tiles"],
        "lanes_per_tile": timing["lanes"],
    }
    # the workload's own parallelism: SIMDs its waves occupy (1024 on the chip)
    waves = timing["tiles"] * max(1, timing["lanes"] // 64)
    used = min(1024, waves)
    roofline["simds_used"] = used
    roofline["frac_of_used_simds"] = round(achieved_tf / (PEAK_FP64_TFLOPS * used / 1024), 4)

    out = {
        "metric": "system-steps/s fwd & fwd+adjoint vs plain at rel err<=1e-10; serving p99 ms",
        "value": value,
        "unit": "system-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"two_dof {w.n_systems} systems x {w.numsteps} RK4 steps, "
                               f"dt 5e-4, sinusoidal forcing, forward, FP64",
                   "systems": w.n_systems, "rk_steps": w.numsteps, "solver": "persistent",
                   "kernel": kinfo["kernel"], "structure_classes": kinfo["classes"],
                   "lanes_per_system": kinfo.get("lanes_per_instance")},
        "roofline": roofline,
    }

    # plain per-step kernel (same GPU, same kernel, one launch per RK4 step)
    if not args.no_plain:
        p.setSolver("plain")
        t0 = time.perf_counter()
        qp = p.getAmplitudes()
        pw = time.perf_counter() - t0
        pt = p._lastTiming()
        p.setSolver("persistent")
        out["plain"] = {"kernel": kinfo["kernel"], "value": sys_steps / pw,
                        "kernel_ms": pt["kernel_ms"],
                        "launches": pt["launches"],
                        "bitwise_equal_to_persistent": bool(np.array_equal(qp, q))}
        out["persistent_vs_plain"] = (sys_steps / wall) / (sys_steps / pw)

    # the tile kernel (rows across lanes, LDS exchange) on the same problem
    if kinfo["kernel"] != "tile":
        p.setKernel("tile")
        p.getAmplitudes()
        qt, tw, tt = _timed(p, p.getAmplitudes)
        out["tile_kernel"] = {"value": sys_steps / tw, "kernel_ms": tt["kernel_ms"],
                              "max_rel_diff_vs_lane": float(np.max(
                                  np.abs(qt - q) / np.maximum(np.abs(q), 1e-8 * np.abs(q).max())))}
        if not args.no_plain:
            p.setSolver("plain")
            _, tpw, tpt = _timed(p, p.getAmplitudes)
            out["plain_tile"] = {"value": sys_steps / tpw, "kernel_ms": tpt["kernel_ms"],
                                 "launches": tpt["launches"]}
            out["persistent_vs_plain_tile"] = tpw / wall
            p.setSolver("persistent")
        p.setKernel("auto")

    # FP32 rate
    p.setPrecision(32)
    p.getAmplitudes()
    t0 = time.perf_counter()
    q32 = p.getAmplitudes()
    w32 = time.perf_counter() - t0
    out["fp32"] = {"value": sys_steps / w32, "kernel_ms": p._lastTiming()["kernel_ms"]}
    p.setPrecision(64)

    if not args.no_adjoint:
        out["fwd_adj"] = fwd_adj_section(args)

    if not args.no_cpu:
        ref, cpu_rate, cpu_s = cpu_baseline(w, args.cpu_sample)
        n_out = ref.size
        g = q[:n_out]
        floor = 1e-8 * np.abs(ref).max()
        err = float(np.max(np.abs(g - ref) / np.maximum(np.abs(ref), floor)))
        g32 = q32[:n_out]
        err32 = float(np.max(np.abs(g32 - ref) / np.maximum(np.abs(ref), floor)))
        out["cpu_baseline"] = {
            "value": cpu_rate, "unit": "system-steps/s", "cores": 1, "kind": "port",
            "sample": f"FP64 NumPy oracle, first {args.cpu_sample} of {w.n_systems} systems "
                      f"x all {w.numsteps} steps ({cpu_s:.1f} s), host cpu_count="
                      f"{os.cpu_count()}"}
        out["parity"] = {"max_rel_err_fp64_vs_oracle": err, "max_rel_err_fp32_vs_fp64_oracle": err32,
                         "sample_systems": args.cpu_sample, "tolerance_fp64": 1e-10}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
