"""Autotuner: tile lanes (systems per workgroup) and slot configuration
(segment width per term kind) of the tile kernel, for the adjoint the
checkpoint interval (also the cap on the lane kernel's interval; a problem
entirely on the lane kernel probes only that), and the lane kernel's group
sizes (the cost model's choice, or every class one size narrower or wider).

Each candidate runs the loaded workload with every system's step count
scaled by one factor so the longest has ``probe_steps`` (the ratios between
systems stay, and with them the lane planner's group choice, which weighs
the longest chain of steps against the total work) and is timed with the
library's HIP events; the fastest wins.  Results depend only on the kernel choice (numerics differ at
round-off level between slot configurations: a row's terms are summed in
different segment groupings), so the tuner records the configuration, not
results.  Choices are cached per workload signature (with the precision) in
a JSON file.
"""
from __future__ import annotations

import hashlib
import json
import os
from typing import Iterable, Optional

import numpy as np

CACHE = os.environ.get("PASIF_AUTOTUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "pasif_amd",
                                    "autotune.json"))

LANES = (64, 128, 256)
SLOTS = (3, 1, 2)          # small, narrow, wide
CHECKPOINTS = (64, 128, 256, 512)


def signature(p, adjoint: bool) -> str:
    """Structure of the loaded problem: sizes, nnz per kind, steps, mode."""
    A, G, L, f, q0 = p.getMatrices()
    parts = [str(len(q0)), str(A.tocsr().nnz), str(len(np.asarray(G.val).reshape(-1))),
             str(len(np.asarray(L.val).reshape(-1))), str(p.numberOfSystems),
             str(int(p._max_steps())), str(adjoint), str(p._precision()),
             ",".join(str(int(x)) for x in np.bincount(np.asarray(p.system_sizes)))]
    return hashlib.sha1("|".join(parts).encode()).hexdigest()[:16]


def probe_caps(steps, probe_steps: int) -> np.ndarray:
    """Step counts of a probe: every system's steps times one factor that
    brings the longest to ``probe_steps`` (never more than its own; systems
    with steps keep at least one, systems without stay at none)."""
    full = np.asarray(steps, dtype=np.int64)
    if full.size == 0:
        return full
    scale = min(1.0, probe_steps / max(1, int(full.max())))
    return np.where(full > 0, np.maximum(1, np.ceil(full * scale)), 0).astype(np.int64)


def full_size_bytes(probe: dict, probe_tmax: int, full_tmax: int) -> float:
    """Run-buffer bytes of the full-size solve estimated from a probe's
    footprint: checkpoint storage grows with the step counts (ceil(T/k)
    slots per tile or wave, the probe's steps being the full ones times one
    factor), the rest does not.  Rounds up: one slot more per checkpoint
    block than the scaled count."""
    scale = full_tmax / max(1, probe_tmax)
    return float(probe["run"] - probe["ckpt"]) + probe["ckpt"] * scale * 1.02


def fits(probe: dict, probe_tmax: int, full_tmax: int, frac: float = 0.8) -> bool:
    """The full-size solve of a candidate stays within ``frac`` of the
    device memory a solve may use (the reference's 0.8 x free planner,
    src/__GpuDriver.cu:896-938).  Host-only handles (free == 0) always fit."""
    if probe["free"] == 0:
        return True
    return full_size_bytes(probe, probe_tmax, full_tmax) <= frac * probe["free"]


def _load_cache() -> dict:
    try:
        with open(CACHE) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _save_cache(c: dict):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE + ".tmp", "w") as fh:
            json.dump(c, fh, indent=1)
        os.replace(CACHE + ".tmp", CACHE)
    except OSError:
        pass


def autotune(p, adjoint: bool = False, probe_steps: int = 2000,
             lanes: Iterable[int] = LANES, slots: Iterable[int] = SLOTS,
             checkpoints: Optional[Iterable[int]] = None, use_cache: bool = True,
             apply: bool = True, verbose: bool = False) -> dict:
    """Pick the fastest (tile lanes, slot configuration[, checkpoint], lane
    group bias) for the problem loaded in PASIf instance ``p``; apply it
    unless ``apply=False``."""
    key = signature(p, adjoint)
    cache = _load_cache() if use_cache else {}
    if key in cache:
        best = cache[key]
    else:
        steps_full = p._sys_steps
        exc = p._sys_exc
        n = p.numberOfSystems
        full = steps_full if steps_full is not None else np.full(n, p.numsteps)
        cap = probe_caps(full, probe_steps)
        full_tmax, probe_tmax = int(np.max(full)) if n else 0, int(np.max(cap)) if n else 0
        p.setSystemExcitations(exc, cap)
        cks = list(checkpoints) if checkpoints is not None else (list(CHECKPOINTS) if adjoint
                                                                 else [0])
        # every system on the lane kernel: tile shapes change nothing, only the
        # checkpoint interval (a cap on the lane kernel's own) is probed
        kernel = p.kernelInfo(adjoint)["kernel"]
        lane_only = kernel == "lane"
        if lane_only:
            lanes, slots = list(lanes)[:1], list(slots)[:1]
        trials = []
        try:
            for ln in lanes:
                for sl in slots:
                    for ck in cks:
                        try:
                            p.setTuning(ln, sl, ck)
                            (p.getGradient if adjoint else p.getAmplitudes)()
                            ms = p._lastTiming()["kernel_ms"]
                            # the full-size solve of this candidate must fit
                            if not fits(p._runFootprint(adjoint), probe_tmax, full_tmax):
                                if verbose:
                                    print(f"[autotune] lanes {ln} slots {sl} ck {ck}: the "
                                          f"full-size run buffers do not fit")
                                continue
                        except Exception as e:  # configuration cannot hold the systems
                            if verbose:
                                print(f"[autotune] lanes {ln} slots {sl} ck {ck}: {e}")
                            continue
                        trials.append((ms, ln, sl, ck))
                        if verbose:
                            print(f"[autotune] lanes {ln} slots {sl} ck {ck}: {ms:.3f} ms")
            if not trials:
                raise RuntimeError("no tile configuration fits this problem (on the device "
                                   "and, at full size, in 80 % of its free memory)")
            ms, ln, sl, ck = min(trials)
            # the lane kernel's group sizes: the cost model's choice, and every
            # class one step narrower / wider where that fits
            bias = 0
            if kernel in ("lane", "mixed"):
                p.setTuning(ln, sl, ck)
                for b in (1, -1):
                    try:
                        p._setLaneBias(b)
                        (p.getGradient if adjoint else p.getAmplitudes)()
                        mb = p._lastTiming()["kernel_ms"]
                        if not fits(p._runFootprint(adjoint), probe_tmax, full_tmax):
                            continue
                    except Exception as e:
                        if verbose:
                            print(f"[autotune] lane bias {b}: {e}")
                        continue
                    finally:
                        p._setLaneBias(0)
                    if verbose:
                        print(f"[autotune] lane bias {b}: {mb:.3f} ms")
                    trials.append((mb, ln, sl, ck))
                    if mb < 0.98 * ms:       # a clear gain over the cost model's choice
                        ms, bias = mb, b
        finally:
            p.setSystemExcitations(exc, steps_full)
        best = {"tile_lanes": ln, "slot_config": sl, "checkpoint_every": ck, "probe_ms": ms,
                "probe_steps": probe_steps, "trials": len(trials), "lane_only": lane_only,
                "lane_bias": bias}
        if use_cache:
            cache[key] = best
            _save_cache(cache)
    if apply:
        p.setTuning(best["tile_lanes"], best["slot_config"], best["checkpoint_every"])
        p._setLaneBias(best.get("lane_bias", 0))
    return best
