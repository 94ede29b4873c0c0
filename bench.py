"""Benchmark of the batched RK4 hot path on one MI355X.

``python bench.py [--gpus 1] [--steps K] [--warmup W]`` prints ONE JSON line.

Headline workload (one bench "step" = one full solve): 4,096 two-DOF systems
``[x0, x1, v0, v1]`` with ``B = 0.05 I``, ``K = [[2,-1],[-1,2]] s``,
3 Gamma and 4 Lambda couplings, sinusoidal forcing on DOF 0 (one excitation
per system), ``dt = 5e-4``, 50,000 RK4 steps, FP64 -- 2.048e8 system-steps
per solve, inputs resident in HBM before the timed region.  ``value`` is
system-steps/s of the persistent solver.

Beside it, in the same line:

* ``plain`` / ``tile_kernel`` / ``plain_tile``: the plain one-launch-per-step
  solver of the same kernel, the tile kernel, and its plain solver;
* ``fp32``; ``roofline`` (algorithmic FP64 flops per launch over the
  HIP-event duration of the launch; the bound is the FP64 VALU peak);
* ``parity``: every output of all 4,096 systems against the FP64 oracle for
  the lane and the tile kernel, FP32 against the FP32 oracle;
* ``cpu_baseline``: the FP64 NumPy oracle on the host, every core (all 4,096
  systems, sharded) and one core (256 systems);
* ``fwd_adj``: forward + checkpointed adjoint on 8,192 chains of 16 DOF x
  20,000 steps, persistent vs plain, roofline and parity on 256 chains, and
  the differentiable form (torch.autograd.Function) on the same chains;
* ``batching``: the 100,000-system heterogeneous batch, FP64 and FP32,
  forward and forward + adjoint, persistent vs plain, parity on a stratified
  sample;
* ``serving``: Poisson requests at 1,000 per second for 60 s against the
  continuous-batching pool, p50 / p99 latency, backlog, system-steps/s, and
  the overload rows (5,000 and 20,000 per second over 20 s each).

The parity metric is ``oracle.parity.rel_err``: per output
``|g - o| / max(|o|, 1e-4 max_s |o|)``, the floor over the same system's
outputs.  Oracle values the host cannot recompute within the run (the FP32
oracle of the two-DOF systems, the chains' gradients, the batch sample) come
from ``tests/golden/`` (written by ``tests/golden/make_fullsize.py``).
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
GOLD = os.path.join(ROOT, "tests", "golden")

PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 vector
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
METRIC = "system-steps/s fwd & fwd+adjoint vs plain at rel err<=1e-10; serving p99 ms"


def _progress(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ------------------------------------------------------------------ work model
def _nnz(t) -> int:
    return len(np.asarray(t.val).reshape(-1))


def flops_per_step(p) -> float:
    """Algorithmic flops of one forward RK4 step of every system (FMA = 2):
    ``4 (2 nnzA + 3 nnzG + 4 nnzL + 2 nnzF) + 12 n`` summed over systems,
    ``nnzF`` the forced DOFs; interpolated forcing adds 3 flops per forced
    DOF and stage."""
    A, G, L, f, q0 = p.getMatrices()
    nnzF = int(np.count_nonzero(np.abs(f) > np.finfo(np.float64).eps))
    interp = 4 * 3 * nnzF if getattr(p, "interpolationMatrix", None) is not None else 0
    return 4.0 * (2 * A.tocsr().nnz + 3 * _nnz(G) + 4 * _nnz(L) + 2 * nnzF) + \
        12.0 * len(q0) + interp


def adjoint_flops_per_step(p) -> float:
    """Algorithmic flops of one adjoint step of every system:
    ``4 (2 nnzA' + 3 nnzG' + 4 nnzL' + 5 nnzPsi') + 12 n_adj + 2 n_fwd``."""
    Ja = (p.jacobian_B + p.jacobian_K).tocsr()
    return 4.0 * (2 * Ja.nnz + 3 * _nnz(p.jacobian_Gamma) + 4 * _nnz(p.jacobian_Lambda) +
                  5 * _nnz(p.jacobian_Psi)) + 12.0 * p.globalAdjointSize + \
        2.0 * p.globalSystemSize


def bytes_forward(p, samples_read: float) -> float:
    """Algorithmic HBM bytes of one forward launch: coefficients and indices
    (8 + 4 B per term), initial and final state, excitation samples read once."""
    A, G, L, f, q0 = p.getMatrices()
    nnz = A.tocsr().nnz + _nnz(G) + _nnz(L)
    return 12.0 * nnz + 16.0 * len(q0) + 8.0 * samples_read


def bytes_fwd_adj(p, samples_read: float, T: int, k: int) -> float:
    """Forward bytes plus the adjoint system's terms, its terminal condition
    and outputs, and the checkpoints (whole forward state every k steps,
    written once and read once)."""
    Ja = (p.jacobian_B + p.jacobian_K).tocsr()
    nnz_a = Ja.nnz + _nnz(p.jacobian_Gamma) + _nnz(p.jacobian_Lambda) + _nnz(p.jacobian_Psi)
    n = p.globalSystemSize
    return bytes_forward(p, samples_read) + 12.0 * nnz_a + 16.0 * p.globalAdjointSize + \
        2.0 * 8.0 * n * (-(-T // max(1, k)))


def read_traffic(name: str):
    """HBM bytes per launch measured with rocprofv3 (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration), from the committed summary."""
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


def _timed(p, fn):
    t0 = time.perf_counter()
    out = fn()
    return out, time.perf_counter() - t0, p._lastTiming()


# ------------------------------------------------------------------ CPU oracle
def _two_dof_shard(args):
    """FP64 oracle final states of systems lo..hi-1 of the headline workload."""
    lo, hi, n_systems, steps = args
    import oracle as O
    from pasif_amd import workloads as W
    w = _BENCH_W.get("two_dof") or W.two_dof(n_systems=n_systems, steps=steps)
    sl = slice(lo, hi)
    asm = O.assemble_forward(w.vecM[sl], w.vecB[sl], w.vecK[sl], w.vecGamma[sl],
                             w.vecLambda[sl], w.vecForce[sl], w.vecQ0[sl])
    n = hi - lo
    lay = O.mapped_layout(asm.sizes, np.arange(n), np.full(n, w.numsteps))
    t0 = time.perf_counter()
    r = O.rk4_forward(asm, O.Forcing(w.excitations[sl]), w.sample_rate, lay)
    return r.q.reshape(-1), time.perf_counter() - t0


_BENCH_W: dict = {}


def cpu_oracle(w, one_core_sample: int) -> dict:
    """The FP64 NumPy oracle ("port") on the host, before any GPU work in this
    process (the shard workers are forked): all systems sharded over every
    core, and the first ``one_core_sample`` systems on one core."""
    from oracle import parity as PAR
    _BENCH_W["two_dof"] = w
    # the cores this process may run on, at most 16 (one GPU's share of a box)
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    S = w.n_systems
    b = np.linspace(0, S, min(S, 4 * cores) + 1).astype(int)
    shards = [(int(b[i]), int(b[i + 1]), S, w.numsteps) for i in range(len(b) - 1)]
    t0 = time.perf_counter()
    res = PAR.run_sharded(_two_dof_shard, shards, processes=cores)
    all_s = time.perf_counter() - t0
    o = np.concatenate([r[0] for r in res])
    q1, one_s = _two_dof_shard((0, one_core_sample, S, w.numsteps))
    return {"o": o, "all_core": {"value": S * w.numsteps / all_s, "cores": cores,
                                 "seconds": all_s, "systems": S},
            "one_core": {"value": one_core_sample * w.numsteps / one_s, "cores": 1,
                         "seconds": one_s, "systems": one_core_sample},
            "one_core_matches_all_core": bool(np.array_equal(q1, o[:q1.size]))}


def duffing_cpu() -> dict:
    """The CPU-only configuration in full on the host: 64 Duffing oscillators
    x'' + 0.1 x' + k x + 0.5 x^3 = sin(w t) (k ~ U[0.5, 2], w = linspace(0.5,
    2, 64), seed 0), x0 = v0 = 0, dt = 1e-3, 10,000 steps, FP64 oracle: max
    |x| over the last 2,000 steps and E(T) = int b v^2 dt per system."""
    import oracle as O
    from pasif_amd import workloads as W
    w = W.duffing()
    asm = O.assemble_forward(w.vecM, w.vecB, w.vecK, w.vecGamma, w.vecLambda, w.vecForce,
                             w.vecQ0)
    lay = O.mapped_layout(asm.sizes, w.exc_of_system, np.full(w.n_systems, w.numsteps))
    t0 = time.perf_counter()
    r = O.rk4_forward(asm, O.Forcing(w.excitations), w.sample_rate, lay, window=2000)
    sec = time.perf_counter() - t0
    return {"w": w, "q": r.q.reshape(-1), "mx": r.maxabs.reshape(-1),
            "value": w.n_systems * w.numsteps / sec, "seconds": sec}


def duffing_section(cpu: dict) -> dict:
    """The same 64 Duffing oscillators on the GPU (fused window maximum)
    against the full CPU run."""
    from oracle import parity as PAR
    from pasif_amd import workloads as W
    w = cpu["w"]
    p = W.make_solver(w)
    q, mx = p.getMaxAmplitudes(window=2000)
    sizes = w.sizes()
    return {"workload": "64 Duffing oscillators x 10,000 steps, dt 1e-3, FP64: max |x| over "
                        "the last 2,000 steps and E(T)",
            "cpu_oracle": {"value": cpu["value"], "unit": "system-steps/s", "cores": 1,
                           "seconds": cpu["seconds"]},
            "gpu_kernel_ms": p._lastTiming()["kernel_ms"],
            "max_abs_x": {"gpu": mx[0::3].tolist()[:4], "oracle": cpu["mx"][0::3].tolist()[:4]},
            "parity_states": PAR.report(q, cpu["q"], sizes).max_rel,
            "parity_window_max": PAR.report(mx, cpu["mx"], sizes).max_rel,
            "tolerance_fp64": PAR.TOL64}


# ------------------------------------------------------------------ sections
def fwd_adj_section(args) -> dict:
    """Forward + checkpointed adjoint on chains of 16 DOF (energy gradient
    w.r.t. the K diagonal and the Lambda coefficients), FP64, checkpoint
    every 256 steps requested: the persistent launch of the kernel the
    library picks, against the plain one-launch-per-step solver of the same
    kernel and of the tile kernel.  A system-step is one forward step plus
    one adjoint step of one system; the checkpoint recompute is issued work,
    not counted as algorithmic work."""
    from oracle import parity as PAR
    from pasif_amd import adjoint as ADJ
    from pasif_amd import workloads as W
    w = W.chains(n_systems=args.adj_systems, d=16, steps=args.adj_steps, steps_per_sample=10)
    p = W.make_solver(w)
    p.setJacobian(*ADJ.energy_gradient_inputs(w).as_args())
    p.setTuning(checkpoint_every=256)
    info = p.kernelInfo(adjoint=True)
    p.getGradient()                                   # warm-up (packing, compilation, upload)
    g, wall, t = _timed(p, p.getGradient)
    T = w.numsteps
    sys_steps = w.n_systems * T
    f_fwd, f_adj = flops_per_step(p), adjoint_flops_per_step(p)
    flops = (f_fwd + f_adj) * T
    achieved = flops / (t["kernel_ms"] * 1e-3) / 1e12
    k = int(info["checkpoint_every"])
    samples = w.n_systems * (T // 10 + 1)
    roof = {"bound": "valu", "pipe": "FP64 VALU (no matrix op on this path)",
            "kernel": f"{info['kernel']}_kernel", "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
            "flops_per_launch": flops,
            "work": "F_fwd + F_adj per system-step (forward and adjoint once each)",
            "issued_flops_per_launch": (2 * f_fwd + f_adj) * T,
            "issued_frac": (2 * f_fwd + f_adj) * T / (t["kernel_ms"] * 1e-3) / 1e12 /
            PEAK_FP64_TFLOPS,
            "algorithmic_bytes": bytes_fwd_adj(p, samples, T, k),
            "traffic": read_traffic("lane_kernel_chains16_8192x20000_adjoint.md"),
            "checkpoint_every": k}
    sec = {"workload": f"chains {w.n_systems} x 16 DOF x {T} steps, dt 1e-3, interpolated "
                       f"band-limited forcing, gradient of E(T) w.r.t. 16 K_ii + 16 Lambda "
                       f"per system, FP64",
           "kernel": info["kernel"], "checkpoint_every": k,
           "lanes_per_system": info.get("lanes_per_instance"),
           "value": sys_steps / wall, "unit": "system-steps/s (fwd+adj)",
           "kernel_ms": t["kernel_ms"], "workgroups": t["tiles"], "roofline": roof}
    gold = os.path.join(GOLD, "fullsize_chains.npz")
    if not args.no_plain:
        p.setSolver("plain")
        gp, pw, pt = _timed(p, p.getGradient)
        sec["plain"] = {"kernel": info["kernel"], "value": sys_steps / pw,
                        "kernel_ms": pt["kernel_ms"], "launches": pt["launches"],
                        "bitwise_equal_to_persistent": bool(np.array_equal(gp, g))}
        sec["persistent_vs_plain"] = pw / wall
        p.setSolver("persistent")
    q = p.getAmplitudes()                             # final states for the parity check
    par = {}
    if os.path.exists(gold) and w.n_systems == 8192 and T == 20000:
        G = np.load(gold)
        ids = G["ids"]
        nq, ng = 33, 65

        def rep(qq, gg):
            rq = PAR.report(qq.reshape(-1, nq)[ids], G["oq"], np.full(len(ids), nq))
            rg = PAR.report(gg.reshape(-1, ng)[ids], G["og"], np.full(len(ids), ng))
            return {"states": rq.as_dict(), "gradients": rg.as_dict()}
        par = {"sample": f"{len(ids)} of {w.n_systems} chains (incl. 3088), full 20,000 "
                         f"steps, FP64 oracle from tests/golden/fullsize_chains.npz",
               "lane": rep(q, g), "tolerance_fp64": PAR.TOL64,
               "oracle_one_ulp_spread": {
                   "states": PAR.sensitivity(G["oq"], G["pq"], np.full(len(ids), nq)),
                   "gradients": PAR.sensitivity(G["og"], G["pg"], np.full(len(ids), ng))}}
    if info["kernel"] != "tile":
        p.setKernel("tile")
        p.getGradient()                               # warm-up (packing, upload)
        gt, tw, tt = _timed(p, p.getGradient)
        sec["tile_persistent"] = {"value": sys_steps / tw, "kernel_ms": tt["kernel_ms"]}
        if not args.no_plain:
            p.setSolver("plain")
            _, tpw, tpt = _timed(p, p.getGradient)
            sec["plain_tile"] = {"value": sys_steps / tpw, "kernel_ms": tpt["kernel_ms"],
                                 "launches": tpt["launches"]}
            sec["persistent_vs_plain_tile"] = tpw / wall
            p.setSolver("persistent")
        if par:
            par["tile"] = rep(p.getAmplitudes(), gt)
        p.setKernel("auto")
    if par:
        sec["parity"] = par
    return sec


def differentiable_section(args) -> dict:
    """The differentiable form (``autograd.rk4_solve``, a torch.autograd
    Function) on the 16-DOF chains: loss = sum(q(T)), gradients w.r.t. q0 and
    the 16 K_ii + 16 Lambda_i of every chain; one forward + backward, timed
    after a warm-up, against the raw driver calls it makes on the same
    problem (getAmplitudes + getGradient).  The host time is what the
    Function spends re-assembling the systems and their adjoints for the
    current coefficients."""
    import torch
    from pasif_amd import workloads as W
    from pasif_amd.autograd import DifferentiableProblem, rk4_solve
    w = W.chains(n_systems=args.adj_systems, d=16, steps=args.adj_steps, steps_per_sample=10)
    d = 16
    params = [[("K", d + i, i) for i in range(d)] + [("Lambda", t) for t in range(d)]
              for _ in range(w.n_systems)]
    out = {"workload": f"chains {w.n_systems} x 16 DOF x {w.numsteps} steps, loss sum(q(T)), "
                       f"gradient w.r.t. q0 and {32 * w.n_systems} coefficients"}
    for mode in ("reference", "exact"):
        prob = DifferentiableProblem(w, params, checkpoint_every=256, mode=mode)
        q0 = torch.tensor(prob.q00(), dtype=torch.float64, requires_grad=True)
        th = torch.tensor(prob.theta0(), dtype=torch.float64, requires_grad=True)

        def step():
            qT = rk4_solve(prob, q0, th)
            qT.sum().backward()
        step()                                       # warm-up (packing, compilation)
        prob.timing = {"host_s": 0.0, "solves": 0}
        t0 = time.perf_counter()
        step()
        wall = time.perf_counter() - t0
        p = prob.solver
        t1 = time.perf_counter()
        p.getAmplitudes()
        kq = p._lastTiming()["kernel_ms"]
        p.getGradient()
        kg = p._lastTiming()["kernel_ms"]
        raw = time.perf_counter() - t1
        out[mode] = {"kernel": p.kernelInfo(adjoint=True)["kernel"], "function_s": wall,
                     "host_assembly_s": prob.timing["host_s"], "raw_calls_s": raw,
                     "raw_kernel_ms": {"forward": kq, "forward_adjoint": kg},
                     "function_vs_raw": wall / raw}
        _progress(f"differentiable {mode}: {wall:.2f} s (raw {raw:.2f} s)")
    return out


def batching_section(args) -> dict:
    """Heterogeneous batch, persistent: 100,000 systems with DOF drawn from
    {1,..,64} ({0.3,0.25,0.15,0.1,0.1,0.06,0.04}), steps log-uniform on
    [1e3, 1e5], dt 1e-3, chain structure at each size, interpolated
    band-limited forcing.  Systems that fit the lane kernel run on it, the
    rest on the tile kernel, each family in one persistent launch.  Lane
    group sizes and checkpoint interval autotuned per precision on probes
    with every system's steps scaled so the longest runs 5,000.  FP64 and
    FP32, forward and forward + adjoint, each against the plain solver of the
    same kernel (one launch per step over the whole batch); errors against
    the FP64 (and FP32) oracle on a stratified sample.  Host-side workload
    generation is not timed."""
    from oracle import parity as PAR
    from pasif_amd import adjoint as ADJ
    from pasif_amd import autotune as AT
    from pasif_amd import workloads as W
    t0 = time.perf_counter()
    w = W.heterogeneous(n_systems=args.batch_systems, tmin=1e3, tmax=1e5, steps_per_sample=10)
    p = W.make_solver(w)
    setup_s = time.perf_counter() - t0
    _progress(f"batching: workload built in {setup_s:.0f} s")
    sys_steps = int(np.sum(w.steps_of_system))
    sizes = w.sizes()
    gold = os.path.join(GOLD, "fullsize_heterogeneous.npz")
    G = np.load(gold) if (os.path.exists(gold) and args.batch_systems == 100_000) else None
    sec = {"workload": f"heterogeneous {w.n_systems} systems, DOF 1..64, steps log-uniform "
                       f"[1e3, 1e5] (total {sys_steps} system-steps), dt 1e-3",
           "kernel": p.kernelInfo()["kernel"], "setup_s": setup_s}
    ids = G["ids"] if G is not None else np.zeros(0, dtype=np.int64)
    off_f = np.concatenate([[0], np.cumsum(sizes)])

    def sample_f(q):
        return np.concatenate([q[off_f[i]:off_f[i + 1]] for i in ids])

    def run(label, adjoint, bits):
        p.setPrecision(bits)
        tune = AT.autotune(p, adjoint=adjoint, probe_steps=5000, lanes=(128, 256),
                           slots=(3, 1), checkpoints=(128, 256) if adjoint else None,
                           use_cache=False)
        fn = p.getGradient if adjoint else p.getAmplitudes
        fn()
        out, wall, t = _timed(p, fn)
        r = {"value": sys_steps / wall, "unit": "system-steps/s" + (" (fwd+adj)" if adjoint
                                                                   else ""),
             "kernel_ms": t["kernel_ms"], "tuning": tune,
             "checkpoint_every": p.kernelInfo(adjoint)["checkpoint_every"] if adjoint else None}
        if not args.no_plain:
            p.setSolver("plain")
            try:
                _, pw, pt = _timed(p, fn)
                r["plain"] = {"value": sys_steps / pw, "kernel_ms": pt["kernel_ms"],
                              "launches": pt["launches"]}
                r["persistent_vs_plain"] = pw / wall
            finally:
                p.setSolver("persistent")
        _progress(f"batching: {label} {t['kernel_ms']:.0f} ms"
                  + (f", plain {r['plain']['kernel_ms']:.0f} ms" if "plain" in r else ""))
        return out, r

    q64, sec["fwd_fp64"] = run("forward FP64", False, 64)
    sec["fwd_fp64"]["tflops"] = flops_per_step(p) * (sys_steps / w.n_systems) / \
        (sec["fwd_fp64"]["kernel_ms"] * 1e-3) / 1e12
    q32, sec["fwd_fp32"] = run("forward FP32", False, 32)
    t1 = time.perf_counter()
    blocks = ADJ.energy_gradient_blocks(w)
    p.setJacobianBlocks(blocks)
    sec["setup_s"] += time.perf_counter() - t1
    g64, sec["fwd_adj_fp64"] = run("forward+adjoint FP64", True, 64)
    g32, sec["fwd_adj_fp32"] = run("forward+adjoint FP32", True, 32)
    p.setPrecision(64)
    if G is not None:
        off_a = np.concatenate([[0], np.cumsum(blocks.sizes)])

        def sample_a(g):
            return np.concatenate([g[off_a[i]:off_a[i + 1]] for i in ids])
        szf, sza = sizes[ids], blocks.sizes[ids]
        sec["parity"] = {
            "sample": f"{len(ids)} systems: per DOF size the two with the most steps and one "
                      f"at random (tests/golden/fullsize_heterogeneous.npz)",
            "fwd_fp64": PAR.report(sample_f(q64), G["q64"], szf).as_dict(),
            "fwd_adj_fp64": PAR.report(sample_a(g64), G["g64"], sza).as_dict(),
            "fwd_fp32_vs_fp64_oracle": PAR.report(sample_f(q32), G["q64"], szf).max_rel,
            "fwd_fp32_vs_fp32_oracle": PAR.report(sample_f(q32), G["q32"], szf).max_rel,
            "fwd_adj_fp32_vs_fp64_oracle": PAR.report(sample_a(g32), G["g64"], sza).max_rel,
            "fwd_adj_fp32_vs_fp32_oracle": PAR.report(sample_a(g32), G["g32"], sza).max_rel,
            "tolerance_fp64": PAR.TOL64}
    if not args.no_structures:
        sec["structures"] = structures_leg(args, sec)
    return sec


def structures_leg(args, batch: dict) -> dict:
    """The same batch with 66 sparsity structures: each system's drawn from
    12 variants of its size (nearest-neighbour Gamma couplings dropped at
    random, up to three extra next-nearest springs).  The lane planner merges
    them into union classes; FP64 forward and forward + adjoint, persistent,
    default tuning, against the 7-structure batch's autotuned times above;
    errors against the FP64 oracle on a stratified sample."""
    from oracle import parity as PAR
    from pasif_amd import adjoint as ADJ
    from pasif_amd import workloads as W
    t0 = time.perf_counter()
    w = W.heterogeneous(n_systems=args.batch_systems, tmin=1e3, tmax=1e5, steps_per_sample=10,
                        variants_per_size=12)
    p = W.make_solver(w)
    sys_steps = int(np.sum(w.steps_of_system))
    info = p.kernelInfo()
    sec = {"workload": "heterogeneous batch above, 66 sparsity structures (12 variants per "
                       "DOF size)", "setup_s": time.perf_counter() - t0,
           "fwd_classes": info["classes"], "fwd_kernel": info["kernel"]}
    _progress(f"structures: workload built, plan {info}")
    p.getAmplitudes()
    q, wall, t = _timed(p, p.getAmplitudes)
    sec["fwd_fp64"] = {"kernel_ms": t["kernel_ms"], "value": sys_steps / wall,
                       "vs_7_structures": t["kernel_ms"] / batch["fwd_fp64"]["kernel_ms"]}
    blocks = ADJ.energy_gradient_blocks(w)
    p.setJacobianBlocks(blocks)
    info = p.kernelInfo(True)
    sec["adj_classes"], sec["adj_kernel"] = info["classes"], info["kernel"]
    p.getGradient()
    g, wall, t = _timed(p, p.getGradient)
    sec["fwd_adj_fp64"] = {"kernel_ms": t["kernel_ms"], "value": sys_steps / wall,
                           "vs_7_structures": t["kernel_ms"] / batch["fwd_adj_fp64"]["kernel_ms"]}
    _progress(f"structures: forward {sec['fwd_fp64']['kernel_ms']:.0f} ms, forward+adjoint "
              f"{sec['fwd_adj_fp64']['kernel_ms']:.0f} ms")
    gold = os.path.join(GOLD, "fullsize_structures.npz")
    if os.path.exists(gold) and args.batch_systems == 100_000:
        G = np.load(gold)
        ids = G["ids"]
        sizes = w.sizes()
        off_f = np.concatenate([[0], np.cumsum(sizes)])
        off_a = np.concatenate([[0], np.cumsum(blocks.sizes)])
        qs = np.concatenate([q[off_f[i]:off_f[i + 1]] for i in ids])
        gs = np.concatenate([g[off_a[i]:off_a[i + 1]] for i in ids])
        sec["parity"] = {
            "sample": f"{len(ids)} systems: per DOF size the two with the most steps and two at "
                      f"random (tests/golden/fullsize_structures.npz)",
            "fwd_fp64": PAR.report(qs, G["q64"], sizes[ids]).as_dict(),
            "fwd_adj_fp64": PAR.report(gs, G["g64"], blocks.sizes[ids]).as_dict(),
            "tolerance_fp64": PAR.TOL64}
    return sec


def serving_section(args) -> dict:
    """Continuous batching: Poisson requests at ``--serving-rps`` for
    ``--serving-s`` seconds, each 1-256 systems from the heterogeneous size
    and step distributions, 70 % forward + adjoint, FP64, 200 GB HBM budget;
    p50 / p99 latency of completed requests (arrival to results on the host),
    the backlog at the end and p99 over every arrived request (a lower
    bound), sustained system-steps/s of completed requests; and the plain
    kernel run per request (one launch per step, request alone) on a
    sample."""
    from pasif_amd import serving as S
    from pasif_amd.workloads import linear_interp_matrix
    rng = np.random.default_rng(4)
    lib = S.SystemLibrary(variants=4)
    n_signals = 32
    bank = S.signal_bank(rng, n_signals, 1e5, steps_per_sample=10)
    interp = linear_interp_matrix(10)

    def new_pool(steps_per_tick):
        pool = S.Pool(lanes=512, slot_config=1, checkpoint_every=256,
                      steps_per_tick=steps_per_tick, hbm_budget=200e9)
        pool.set_bank(bank, 1000, interp)
        pool.register_library(lib)
        pool.set_slots(32768)
        return pool

    arr = S.poisson_arrivals(rng, args.serving_rps, args.serving_s)
    reqs = S.RequestStream(lib, len(arr), 4000, n_signals, tmin=1e3, tmax=1e5, max_systems=256)
    pool = new_pool(5000)
    rep = S.SolveServer(pool).run(reqs, horizon_s=args.serving_s, drain=False,
                                  max_wall_s=args.serving_s + 15, arrivals=arr)
    pool.close()
    sec = {"workload": f"Poisson {args.serving_rps:g} req/s for {args.serving_s:g} s, 1-256 "
                       f"systems per request (DOF 1..64, steps log-uniform [1e3, 1e5]), 70 % "
                       f"forward+adjoint, FP64, 200 GB HBM budget",
           "admission": "queued requests in groups of at most 4,096 systems per adjoint flag, "
                        "one pool job per group",
           "offered_rps": args.serving_rps, "arrived": rep.arrived, "admitted": rep.admitted,
           "completed": rep.completed, "completed_frac": rep.completed / max(rep.arrived, 1),
           "backlog": rep.backlog, "p50_ms": rep.p50_ms, "p99_ms": rep.p99_ms,
           "p99_all_arrived_ms": rep.p99_all_ms, "system_steps_per_s": rep.system_steps_per_s,
           "gpu_busy_frac": rep.gpu_busy_frac, "admit_ms_mean": rep.admit_ms_mean,
           "wall_s": rep.duration_s}
    # overload rows: the same stream at higher rates over shorter windows
    # (the pool saturates; latency is then the FIFO's wait)
    over = []
    for li, load in enumerate(args.serving_overload):
        arr = S.poisson_arrivals(rng, load, args.serving_overload_s)
        reqs = S.RequestStream(lib, len(arr), 4001 + li, n_signals, tmin=1e3, tmax=1e5,
                               max_systems=256)
        pool = new_pool(5000)
        rep = S.SolveServer(pool).run(reqs, horizon_s=args.serving_overload_s, drain=False,
                                      max_wall_s=args.serving_overload_s + 15, arrivals=arr)
        pool.close()
        over.append({"offered_rps": load, "window_s": args.serving_overload_s,
                     "arrived": rep.arrived, "admitted": rep.admitted,
                     "completed": rep.completed, "backlog": rep.backlog,
                     "p50_ms": rep.p50_ms, "p99_ms": rep.p99_ms,
                     "p99_all_arrived_ms": rep.p99_all_ms,
                     "system_steps_per_s": rep.system_steps_per_s,
                     "gpu_busy_frac": rep.gpu_busy_frac})
    sec["overload"] = over
    rates, nsys = [], []
    prng = np.random.default_rng(7)     # the same sample whatever ran before
    for k in range(3):
        r = S.draw_request(lib, prng, 10**6 + k, n_signals, tmin=1e3, tmax=3e3, max_systems=256)
        nsys.append(len(r.steps))
        pool = new_pool(1)
        job = pool.submit_ids(r)
        t0 = time.perf_counter()
        while not pool.finished(job):
            pool.tick(wait=True)
        wall = time.perf_counter() - t0
        pool.collect(job, r)
        pool.close()
        rates.append(r.system_steps / wall)
    sec["plain_per_request"] = {"system_steps_per_s": float(np.mean(rates)),
                                "systems_per_request": nsys,
                                "note": "one launch per step, request alone, 3 requests of "
                                        "steps in [1e3, 3e3]; its rate grows with the "
                                        "request's systems (the launch cost is per step)"}
    return sec


# ------------------------------------------------------------------ main
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
    ap.add_argument("--no-batching", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-structures", action="store_true")
    ap.add_argument("--adj-systems", type=int, default=8192)
    ap.add_argument("--adj-steps", type=int, default=20_000)
    ap.add_argument("--batch-systems", type=int, default=100_000)
    ap.add_argument("--serving-rps", type=float, default=1000.0)
    ap.add_argument("--serving-s", type=float, default=60.0)
    ap.add_argument("--serving-overload", type=float, nargs="*", default=[5000.0, 20000.0])
    ap.add_argument("--serving-overload-s", type=float, default=20.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: systems are independent and the path needs no "
              f"communication; --gpus {args.gpus} is not supported (each GPU would run its "
              f"own share of systems).", file=sys.stderr)
        sys.exit(2)

    from oracle import parity as PAR
    from pasif_amd import workloads as W

    w = W.two_dof(n_systems=args.systems, steps=args.rk_steps)
    sizes = w.sizes()
    cpu = duff = None
    if not args.no_cpu:
        _progress("cpu oracle (all cores, then one core; Duffing configuration in full)")
        duff = duffing_cpu()
        cpu = cpu_oracle(w, min(args.cpu_sample, w.n_systems))
        _progress(f"cpu oracle: {cpu['all_core']['seconds']:.1f} s on "
                  f"{cpu['all_core']['cores']} cores, {cpu['one_core']['seconds']:.1f} s one core")

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
    _progress(f"headline: {kms:.2f} ms per solve")

    flops = flops_per_step(p) * w.numsteps
    achieved_tf = flops / (kms * 1e-3) / 1e12
    waves = timing["tiles"] * max(1, timing["lanes"] // 64)
    used = min(1024, waves)
    roofline = {
        "bound": "valu",
        "pipe": "FP64 VALU (no matrix op on this path; MI355X FP64 vector peak)",
        "kernel": kname,
        "achieved": round(achieved_tf, 4),
        "peak": PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved_tf / PEAK_FP64_TFLOPS, 5),
        "traffic": read_traffic(f"{kname}_two_dof_4096x50000_traffic.txt"),
        "algorithmic_bytes": bytes_forward(p, w.n_systems * (w.numsteps + 1)),
        "flops_per_launch": flops,
        "kernel_ms": round(kms, 4),
        "tiles": timing["tiles"],
        "lanes_per_tile": timing["lanes"],
        # the workload's own parallelism: SIMDs its waves occupy (1024 on the chip)
        "simds_used": used,
        "frac_of_used_simds": round(achieved_tf / (PEAK_FP64_TFLOPS * used / 1024), 4),
    }
    out = {
        "metric": METRIC, "value": value, "unit": "system-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"two_dof {w.n_systems} systems x {w.numsteps} RK4 steps, "
                               f"dt 5e-4, sinusoidal forcing, forward, FP64",
                   "systems": w.n_systems, "rk_steps": w.numsteps, "solver": "persistent",
                   "kernel": kinfo["kernel"], "structure_classes": kinfo["classes"],
                   "lanes_per_system": kinfo.get("lanes_per_instance")},
        "roofline": roofline,
    }

    if not args.no_plain:
        p.setSolver("plain")
        qp, pw, pt = _timed(p, p.getAmplitudes)
        p.setSolver("persistent")
        out["plain"] = {"kernel": kinfo["kernel"], "value": sys_steps / pw,
                        "kernel_ms": pt["kernel_ms"], "launches": pt["launches"],
                        "bitwise_equal_to_persistent": bool(np.array_equal(qp, q))}
        out["persistent_vs_plain"] = pw / wall

    qt = None
    if kinfo["kernel"] != "tile":
        p.setKernel("tile")
        p.getAmplitudes()
        qt, tw, tt = _timed(p, p.getAmplitudes)
        out["tile_kernel"] = {"value": sys_steps / tw, "kernel_ms": tt["kernel_ms"]}
        if not args.no_plain:
            p.setSolver("plain")
            _, tpw, tpt = _timed(p, p.getAmplitudes)
            out["plain_tile"] = {"value": sys_steps / tpw, "kernel_ms": tpt["kernel_ms"],
                                 "launches": tpt["launches"]}
            out["persistent_vs_plain_tile"] = tpw / wall
            p.setSolver("persistent")
        p.setKernel("auto")

    p.setPrecision(32)
    p.getAmplitudes()
    q32, w32, t32 = _timed(p, p.getAmplitudes)
    out["fp32"] = {"value": sys_steps / w32, "kernel_ms": t32["kernel_ms"]}
    p.setPrecision(64)

    if cpu is not None:
        o = cpu["o"]
        par = {"sample": f"all {w.n_systems} systems x {w.numsteps} steps, every final state",
               "lane" if kinfo["kernel"] != "tile" else "tile":
                   PAR.report(q, o, sizes).as_dict(),
               "tolerance_fp64": PAR.TOL64}
        if qt is not None:
            par["tile"] = PAR.report(qt, o, sizes).as_dict()
        gold = os.path.join(GOLD, "fullsize_two_dof.npz")
        if os.path.exists(gold) and w.n_systems == 4096 and w.numsteps == 50_000:
            G = np.load(gold)
            par["oracle_matches_fixture"] = bool(np.array_equal(G["o"], o))
            par["oracle_one_ulp_spread"] = PAR.sensitivity(G["o"], G["op"], sizes)
            par["fp32_vs_fp32_oracle"] = PAR.report(q32, G["o32"], sizes).max_rel
            par["fp32_vs_fp64_oracle"] = PAR.report(q32, o, sizes).max_rel
            par["fp32_oracle_one_ulp_spread"] = PAR.sensitivity(G["o32"], G["p32"], sizes)
        out["parity"] = par
        out["cpu_baseline"] = {
            "value": cpu["all_core"]["value"], "unit": "system-steps/s",
            "cores": cpu["all_core"]["cores"], "kind": "port",
            "sample": f"FP64 NumPy oracle, all {w.n_systems} systems x {w.numsteps} steps "
                      f"sharded over {cpu['all_core']['cores']} processes "
                      f"({cpu['all_core']['seconds']:.1f} s)",
            "one_core": {"value": cpu["one_core"]["value"], "cores": 1,
                         "sample": f"first {cpu['one_core']['systems']} systems x all "
                                   f"{w.numsteps} steps ({cpu['one_core']['seconds']:.1f} s)"}}
    del p
    if duff is not None:
        out["duffing"] = duffing_section(duff)

    if not args.no_adjoint:
        _progress("fwd_adj")
        out["fwd_adj"] = fwd_adj_section(args)
        _progress("differentiable form")
        out["fwd_adj"]["differentiable"] = differentiable_section(args)
    if not args.no_batching:
        _progress("batching")
        out["batching"] = batching_section(args)
    if not args.no_serving:
        _progress("serving")
        out["serving"] = serving_section(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
