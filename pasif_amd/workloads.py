"""Synthetic workloads of the benchmark configurations, built as PASIf user
inputs (per-system ``dia_matrix`` M, ``coo_matrix`` B and K, ``coo_tensor``
Gamma / Lambda, force pattern, initial conditions) plus excitations.

State layouts (first-order form, ``M = I``; the driver scales by ``-M^{-1}``
so the user matrices carry the opposite sign of the RHS, as in the
reference's examples):

* Duffing, ``[x, v, E]``: ``x' = v``, ``v' = -b v - k x - lam x^3 + A u``,
  ``E' = b v^2`` (the dissipated-energy probe);
* two-DOF system, ``[x0, x1, v0, v1]``;
* chains of d DOF, ``[x(d), v(d), E]`` with ``E' = v^T B v``.

Everything is seeded (``numpy.random.default_rng``); sizes follow the
benchmark configurations and can be reduced for tests.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
from scipy.sparse import coo_matrix, dia_matrix

from .PASIf import coo_tensor


@dataclass
class Workload:
    name: str
    vecM: list
    vecB: list
    vecK: list
    vecGamma: list
    vecLambda: list
    vecForce: list
    vecQ0: list
    excitations: np.ndarray          # [n_exc][L]
    sample_rate: int
    numsteps: int
    exc_of_system: Optional[np.ndarray] = None
    steps_of_system: Optional[np.ndarray] = None
    interp: Optional[np.ndarray] = None
    meta: dict = field(default_factory=dict)

    @property
    def n_systems(self) -> int:
        return len(self.vecM)

    @property
    def system_steps(self) -> int:
        if self.steps_of_system is not None:
            return int(np.sum(self.steps_of_system))
        n_inst = self.n_systems if self.exc_of_system is not None else \
            self.n_systems * len(self.excitations)
        return int(n_inst) * int(self.numsteps)

    def sizes(self) -> np.ndarray:
        return np.array([m.shape[0] for m in self.vecM], dtype=np.int64)


def _tensor(dims, val, idx):
    t = coo_tensor(list(dims))
    t.val = np.asarray(val, dtype=np.float64)
    t.indices = np.asarray(idx, dtype=np.int64).reshape(-1)
    return t


def duffing(n_systems: int = 64, steps: int = 10_000, dt: float = 1e-3, seed: int = 0,
            b: float = 0.1, lam: float = 0.5, amp: float = 1.0) -> Workload:
    """64 Duffing oscillators ``x'' + b x' + k x + lam x^3 = A sin(w t)``,
    ``k ~ U[0.5, 2]``, ``w = linspace(0.5, 2, 64)``, x0 = v0 = 0, zero-order
    hold forcing, state ``[x, v, E]`` with ``E' = b v^2``."""
    rng = np.random.default_rng(seed)
    k = rng.uniform(0.5, 2.0, n_systems)
    w = np.linspace(0.5, 2.0, n_systems)
    sr = int(round(1.0 / dt))
    t = np.arange(steps + 1) / sr
    exc = np.sin(w[:, None] * t[None, :])
    vM, vB, vK, vG, vL, vF, vQ = [], [], [], [], [], [], []
    for s in range(n_systems):
        vM.append(dia_matrix((np.ones(3), [0]), shape=(3, 3)))
        vB.append(coo_matrix(([-1.0, b], ([0, 1], [1, 1])), shape=(3, 3)))
        vK.append(coo_matrix(([k[s]], ([1], [0])), shape=(3, 3)))
        vG.append(_tensor([3] * 3, [-b], [2, 1, 1]))
        vL.append(_tensor([3] * 4, [lam], [1, 0, 0, 0]))
        vF.append(np.array([0.0, amp, 0.0]))
        vQ.append(np.zeros(3))
    return Workload("duffing", vM, vB, vK, vG, vL, vF, vQ, exc, sr, steps,
                    exc_of_system=np.arange(n_systems), meta={"k": k, "w": w, "b": b})


def two_dof(n_systems: int = 4096, steps: int = 50_000, dt: float = 5e-4,
            seed: int = 1) -> Workload:
    """2-DOF systems ``[x0, x1, v0, v1]``: ``B = 0.05 I``,
    ``K = [[2,-1],[-1,2]] s`` with ``s ~ U[0.8, 1.2]``; Gamma on 3 fixed
    entries ~ N(0, 0.1); Lambda on 4 fixed entries ~ U[0, 0.5]; forcing
    ``sin(w t)`` on DOF 0 with ``w ~ U[0.5, 3]``, one excitation per system."""
    rng = np.random.default_rng(seed)
    s_ = rng.uniform(0.8, 1.2, n_systems)
    w = rng.uniform(0.5, 3.0, n_systems)
    g = rng.normal(0.0, 0.1, (n_systems, 3))
    l_ = rng.uniform(0.0, 0.5, (n_systems, 4))
    sr = int(round(1.0 / dt))
    t = np.arange(steps + 1) / sr
    exc = np.sin(w[:, None] * t[None, :])
    g_idx = [2, 0, 1, 3, 0, 0, 3, 1, 1]
    l_idx = [2, 0, 0, 0, 3, 1, 1, 1, 2, 0, 1, 1, 3, 0, 0, 1]
    vM, vB, vK, vG, vL, vF, vQ = [], [], [], [], [], [], []
    M = dia_matrix((np.ones(4), [0]), shape=(4, 4))
    B = coo_matrix(([-1.0, -1.0, 0.05, 0.05], ([0, 1, 2, 3], [2, 3, 2, 3])), shape=(4, 4))
    for s in range(n_systems):
        vM.append(M)
        vB.append(B)
        vK.append(coo_matrix(([2 * s_[s], -s_[s], -s_[s], 2 * s_[s]], ([2, 2, 3, 3], [0, 1, 0, 1])),
                             shape=(4, 4)))
        vG.append(_tensor([4] * 3, g[s], g_idx))
        vL.append(_tensor([4] * 4, l_[s], l_idx))
        vF.append(np.array([0.0, 0.0, 1.0, 0.0]))
        vQ.append(np.zeros(4))
    return Workload("two_dof", vM, vB, vK, vG, vL, vF, vQ, exc, sr, steps,
                    exc_of_system=np.arange(n_systems), meta={"s": s_, "w": w})


def linear_interp_matrix(steps_per_sample: int) -> np.ndarray:
    """Rows 0 and 1 of a linear interpolation on the half-step grid:
    ``W = 2*steps_per_sample``, ``u = (1 - j/W) e[i] + (j/W) e[i+1]``."""
    W = 2 * steps_per_sample
    j = np.arange(W) / W
    return np.stack([j, 1.0 - j])


def band_limited(rng, n_samples: int, cutoff: float = 0.1) -> np.ndarray:
    """White Gaussian noise low-passed by an FFT brick-wall at ``cutoff``
    (fraction of the sample rate), unit RMS."""
    x = rng.normal(0.0, 1.0, n_samples)
    X = np.fft.rfft(x)
    f = np.fft.rfftfreq(n_samples)
    X[f > cutoff] = 0.0
    y = np.fft.irfft(X, n_samples)
    return y / (np.sqrt(np.mean(y ** 2)) + 1e-300)


_CHAIN = {}


def _chain_template(d: int, with_energy: bool) -> dict:
    """Index patterns of a chain of ``d`` DOF (the same for every scale):
    built once per size."""
    key = (d, with_energy)
    t = _CHAIN.get(key)
    if t is not None:
        return t
    n = 2 * d + (1 if with_energy else 0)
    T = (np.diag(np.full(d, 2.0)) - np.diag(np.ones(d - 1), 1) - np.diag(np.ones(d - 1), -1)) \
        if d > 1 else np.array([[2.0]])
    br, bc = np.nonzero(0.02 * T + 0.01 * np.eye(d))
    kr, kc = np.nonzero(T)
    gi = chain_gamma_index(d)
    G_idx = list(gi.reshape(-1))
    if with_energy:
        for j, k in zip(br, bc):
            G_idx += [2 * d, d + j, d + k]
    L_idx = []
    for i in range(d):
        L_idx += [d + i, i, i, i]
    f = np.zeros(n)
    f[d] = 1.0
    t = {"n": n, "T": T, "br": br, "bc": bc, "kr": kr, "kc": kc, "ngi": len(gi),
         "B_rows": np.concatenate([np.arange(d), d + br]),
         "B_cols": np.concatenate([np.arange(d, 2 * d), d + bc]),
         "K_rows": d + kr, "K_cols": kc,
         "G_idx": np.asarray(G_idx, dtype=np.int64), "L_idx": np.asarray(L_idx, dtype=np.int64),
         "M": dia_matrix((np.ones(n), [0]), shape=(n, n)), "f": f}
    _CHAIN[key] = t
    return t


def chain_system(d: int, s: float, gvals: np.ndarray, lvals: np.ndarray, with_energy=True):
    """User inputs of one chain of ``d`` DOF: ``K = tridiag(-1, 2, -1) s``,
    ``B = 0.02 K + 0.01 I``, nearest-neighbour Gamma (values ``gvals``),
    on-site Lambda (``lvals``), energy probe ``E' = v^T B v``.  The mass
    matrix and force pattern are shared by the chains of one size."""
    t = _chain_template(d, with_energy)
    n = t["n"]
    Kd = t["T"] * s
    Bd = 0.02 * Kd + 0.01 * np.eye(d)
    br, bc = t["br"], t["bc"]
    B = coo_matrix((np.concatenate([np.full(d, -1.0), Bd[br, bc]]), (t["B_rows"], t["B_cols"])),
                   shape=(n, n))
    K = coo_matrix((Kd[t["kr"], t["kc"]], (t["K_rows"], t["K_cols"])), shape=(n, n))
    G_val = np.asarray(gvals[: t["ngi"]], dtype=np.float64)
    if with_energy:
        G_val = np.concatenate([G_val, -Bd[br, bc]])
    G = _tensor([n] * 3, G_val, t["G_idx"])
    L = _tensor([n] * 4, lvals[:d], t["L_idx"])
    return t["M"], B, K, G, L, t["f"], np.zeros(n), Kd, Bd


def chain_gamma_index(d: int) -> np.ndarray:
    """Nearest-neighbour quadratic couplings of a chain: (v_i; x_i, x_{i+1})
    and (v_{i+1}; x_i, x_{i+1}) per bond, plus on-site (v_0; x_0, x_0) and
    (v_{d-1}; x_{d-1}, x_{d-1}) -- 2d entries."""
    idx = []
    for i in range(d - 1):
        idx.append((d + i, i, i + 1))
        idx.append((d + i + 1, i, i + 1))
    idx.append((d, 0, 0))
    idx.append((2 * d - 1, d - 1, d - 1))
    return np.array(idx[: 2 * d], dtype=np.int64)


def chains(n_systems: int = 8192, d: int = 16, steps: int = 20_000, dt: float = 1e-3,
           steps_per_sample: int = 10, seed: int = 2) -> Workload:
    """Chains of ``d`` DOF, K scaled by ``U[0.9, 1.1]``, Gamma ~ N(0, 0.05),
    Lambda ~ U[0, 1], band-limited forcing sampled every ``steps_per_sample``
    steps and linearly interpolated, applied to DOF 0."""
    rng = np.random.default_rng(seed)
    sr = int(round(1.0 / dt))
    n_samples = steps // steps_per_sample + 1
    vM, vB, vK, vG, vL, vF, vQ = [], [], [], [], [], [], []
    exc = np.empty((n_systems, n_samples))
    scales = rng.uniform(0.9, 1.1, n_systems)
    for s in range(n_systems):
        gv = rng.normal(0.0, 0.05, 2 * d)
        lv = rng.uniform(0.0, 1.0, d)
        M, B, K, G, L, f, q0, _, _ = chain_system(d, scales[s], gv, lv)
        vM.append(M); vB.append(B); vK.append(K); vG.append(G); vL.append(L)
        vF.append(f); vQ.append(q0)
        exc[s] = band_limited(rng, n_samples)
    return Workload("chains", vM, vB, vK, vG, vL, vF, vQ, exc, sr, steps,
                    exc_of_system=np.arange(n_systems),
                    interp=linear_interp_matrix(steps_per_sample),
                    meta={"d": d, "scales": scales})


DOF_CHOICES = np.array([1, 2, 4, 8, 16, 32, 64])
DOF_PROBS = np.array([0.3, 0.25, 0.15, 0.1, 0.1, 0.06, 0.04])


def structure_variant(d: int, v: int):
    """Sparsity variant ``v`` of the ``d``-DOF chain: which of its 2d
    nearest-neighbour quadratic couplings (Gamma) it keeps and up to three
    extra springs between DOF two or three apart (next-nearest couplings),
    drawn from a generator seeded by ``(d, v)``.  The parameters of the
    energy gradient (K diagonal, Lambda) are those of the plain chain in
    every variant.  Variant 0 is the plain chain; a chain of 1 DOF has 2
    variants, of 2 DOF 4."""
    if v == 0:
        return np.ones(2 * d, dtype=bool), []
    if d == 1:          # two copies of one monomial: all or none
        return np.zeros(2, dtype=bool), []
    if d == 2:
        keep = np.ones(4, dtype=bool)
        keep[v - 1] = False
        return keep, []
    r = np.random.default_rng([d, v])
    keep = r.uniform(size=2 * d) < 0.7
    starts = r.integers(0, d, size=6)
    gaps = r.integers(2, 4, size=6)
    pairs = {(int(i), int(i + g)) for i, g in zip(starts, gaps) if i + g < d}
    return keep, sorted(pairs)[:3]


def variant_count(d: int, per_size: int) -> int:
    """Distinct variants ``structure_variant`` gives a ``d``-DOF chain."""
    return min(per_size, 2 if d == 1 else (4 if d == 2 else per_size))


def apply_variant(d, K, G, keep, springs, s):
    """K with the extra springs (stiffness 0.5 s each, symmetric) and the
    Gamma tensor without the nearest-neighbour couplings ``keep`` drops (the
    energy-probe entries after them stay)."""
    if springs:
        K = K.tocoo()
        rows, cols, vals = [K.row], [K.col], [K.data]
        for i, j in springs:
            k = 0.5 * s
            rows.append(np.array([d + i, d + j, d + i, d + j]))
            cols.append(np.array([i, j, j, i]))
            vals.append(np.array([k, k, -k, -k]))
        K = coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                       shape=K.shape)
        K.sum_duplicates()
    if not keep.all():
        idx = np.asarray(G.indices).reshape(-1, 3)
        val = np.asarray(G.val)
        k = np.ones(len(val), dtype=bool)
        k[: len(keep)] = keep
        G = _tensor(list(G.dimensions), val[k], idx[k].reshape(-1))
    return K, G


def heterogeneous(n_systems: int = 100_000, tmin: float = 1e3, tmax: float = 1e5,
                  dt: float = 1e-3, steps_per_sample: int = 10, seed: int = 3,
                  n_signals: int = 64, dof_choices=None, dof_probs=None,
                  variants_per_size: int = 1) -> Workload:
    """Systems of uneven size and step count: DOF from {1,..,64} with the
    configured probabilities (or ``dof_choices`` / ``dof_probs``), steps
    log-uniform on [tmin, tmax], chain structure at each size, one of
    ``n_signals`` band-limited forcings each.  ``variants_per_size`` > 1
    draws each system's sparsity uniformly from that many variants of its
    size (``structure_variant``): 12 gives 66 distinct structures over the
    seven sizes."""
    rng = np.random.default_rng(seed)
    choices = DOF_CHOICES if dof_choices is None else np.asarray(dof_choices)
    probs = DOF_PROBS if dof_probs is None else np.asarray(dof_probs, dtype=np.float64)
    dofs = rng.choice(choices, size=n_systems, p=probs / np.sum(probs))
    steps = np.exp(rng.uniform(np.log(tmin), np.log(tmax), n_systems)).astype(np.int64)
    sr = int(round(1.0 / dt))
    n_samples = int(steps.max()) // steps_per_sample + 1
    exc = np.stack([band_limited(rng, n_samples) for _ in range(n_signals)])
    vM, vB, vK, vG, vL, vF, vQ = [], [], [], [], [], [], []
    scales = rng.uniform(0.9, 1.1, n_systems)
    variants = np.zeros(n_systems, dtype=np.int64)
    if variants_per_size > 1:
        nv = np.array([variant_count(int(d), variants_per_size) for d in dofs])
        variants = (rng.uniform(size=n_systems) * nv).astype(np.int64)
    shapes = {}
    for s in range(n_systems):
        d = int(dofs[s])
        gv = rng.normal(0.0, 0.05, 2 * d)
        lv = rng.uniform(0.0, 1.0, d)
        M, B, K, G, L, f, q0, _, _ = chain_system(d, scales[s], gv, lv)
        v = int(variants[s])
        if v:
            key = (d, v)
            if key not in shapes:
                shapes[key] = structure_variant(d, v)
            K, G = apply_variant(d, K, G, *shapes[key], scales[s])
        vM.append(M); vB.append(B); vK.append(K); vG.append(G); vL.append(L)
        vF.append(f); vQ.append(q0)
    return Workload("heterogeneous", vM, vB, vK, vG, vL, vF, vQ, exc, sr, int(steps.max()),
                    exc_of_system=rng.integers(0, n_signals, n_systems),
                    steps_of_system=steps, interp=linear_interp_matrix(steps_per_sample),
                    meta={"dofs": dofs, "variants": variants})


def make_solver(w: Workload, **kw):
    """A :class:`pasif_amd.PASIf` loaded with the workload."""
    from .PASIf import PASIf
    p = PASIf(list(w.excitations), w.sample_rate, w.numsteps, **kw)
    p.setSystems(w.vecM, w.vecB, w.vecK, w.vecGamma, w.vecLambda, w.vecForce, w.vecQ0)
    if w.exc_of_system is not None or w.steps_of_system is not None:
        p.setSystemExcitations(w.exc_of_system, w.steps_of_system)
    if w.interp is not None:
        p.setInterpolationMatrix(list(w.interp))
    return p
