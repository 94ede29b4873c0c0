"""Continuous-batching solve server on one GPU.

Requests (batches of 1-256 systems, forward-only or forward+adjoint) are
admitted into a device-resident pool (``pasif_pool_*`` in the C ABI) as
soon as the pool's HBM budget allows.  Every tick advances all admitted,
unfinished requests by a slice of steps in ONE kernel launch; the host admits
newly arrived requests while that launch runs and collects finished ones
after it, so requests that arrive while others are on the GPU join them at
the next tick.  A request's result is bitwise independent of its neighbours:
each request owns its tiles and no arithmetic crosses systems.

Requests are drawn from the heterogeneous-batch distributions (DOF sizes
{1,...,64}, log-uniform step counts, chain structure).  To keep admission
cheap at high request rates, systems are taken from a library of
pre-assembled chain systems per size (``SystemLibrary``) and forcing signals
from a bank uploaded once (``pasif_pool_set_bank``).
"""
from __future__ import annotations

import ctypes as C
import time
from collections import Counter, deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib
from . import adjoint as ADJ
from . import workloads as W
from .PASIf import assemble_arrays
from .PASIfgpu import __GpuDriver as _Driver

_u32p = C.POINTER(C.c_uint32)
_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)

_POOL_SIGS = {
    "pasif_pool_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_uint32,
                                    C.c_uint32, C.c_uint32, C.c_uint32, C.c_double]),
    "pasif_pool_destroy": (C.c_int, [C.c_void_p]),
    "pasif_pool_last_error": (C.c_char_p, [C.c_void_p]),
    "pasif_pool_set_bank": (C.c_int, [C.c_void_p, _f64p, C.c_uint32, C.c_uint32, C.c_uint32,
                                      _f64p, C.c_uint32]),
    "pasif_pool_job_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), _f64p, C.c_uint32,
                                        C.c_uint32, C.c_uint32, C.c_uint32]),
    "pasif_pool_submit": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "pasif_pool_tick": (C.c_int, [C.c_void_p, C.c_int, _u32p]),
    "pasif_pool_wait": (C.c_int, [C.c_void_p, _f64p]),
    "pasif_pool_job_finished": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]),
    "pasif_pool_collect": (C.c_int, [C.c_void_p, C.c_void_p, _f64p, C.c_size_t,
                                     C.POINTER(C.c_size_t), _f64p, C.c_size_t,
                                     C.POINTER(C.c_size_t)]),
    "pasif_pool_stats": (C.c_int, [C.c_void_p, _u32p, _f64p, _f64p, _u32p]),
    "pasif_pool_slots": (C.c_int, [C.c_void_p, _u32p, _u32p]),
    "pasif_pool_set_compile": (C.c_int, [C.c_void_p, C.c_int]),
    "pasif_pool_compile_wait": (C.c_int, [C.c_void_p, C.c_uint32, _u32p, _u32p, _u32p]),
    "pasif_pool_set_slots": (C.c_int, [C.c_void_p, C.c_uint32]),
    "pasif_pool_register": (C.c_int, [C.c_void_p, C.c_void_p, _u32p, _u32p]),
    "pasif_pool_submit_systems": (C.c_int, [C.c_void_p, C.c_uint32, _u32p, _i32p, _u32p,
                                            C.c_int, C.POINTER(C.c_void_p)]),
    "pasif_pool_poll": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "pasif_pool_submit_batch": (C.c_int, [C.c_void_p, C.c_uint32, _u32p, _u32p, _i32p, _u32p,
                                          _i32p, C.POINTER(C.c_void_p), _u32p]),
    "pasif_pool_collect_batch": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p),
                                           _f64p, C.POINTER(C.c_size_t), _f64p,
                                           C.POINTER(C.c_size_t)]),
}


def _load():
    lib = _lib.load()
    for name, (res, args) in _POOL_SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def pool_symbols():
    return list(_POOL_SIGS)


# --------------------------------------------------------------- payloads
@dataclass
class SystemArrays:
    """Driver arrays of one system (local indices), forward and adjoint."""
    n: int
    csr: tuple            # data, indices, indptr
    gamma: tuple          # (idx [nnz,3], val)
    lam: tuple
    force: np.ndarray
    q0: np.ndarray
    na: int = 0
    acsr: Optional[tuple] = None
    agamma: Optional[tuple] = None
    alam: Optional[tuple] = None
    apsi: Optional[tuple] = None
    lamT: Optional[np.ndarray] = None
    key: int = -1         # position in SystemLibrary.flat


def _tens(t, order):
    v = np.asarray(t.val, dtype=np.float64).reshape(-1)
    ix = (np.asarray(t.indices, dtype=np.int64).reshape(-1, order) if v.size
          else np.zeros((0, order), dtype=np.int64))
    return ix, v


def chain_arrays(d: int, s: float, gv: np.ndarray, lv: np.ndarray, adjoint: bool = True,
                 variant: int = 0) -> SystemArrays:
    """Driver arrays of one chain of ``d`` DOF (``workloads.chain_system``,
    sparsity variant ``variant`` of ``workloads.structure_variant``) with the
    energy-gradient adjoint (dE(T)/dK_ii, dE(T)/dLambda_i)."""
    M, B, K, G, L, f, q0, _, _ = W.chain_system(int(d), s, gv, lv)
    if variant:
        K, G = W.apply_variant(int(d), K, G, *W.structure_variant(int(d), int(variant)), s)
    r = assemble_arrays([M], [B], [K], [G], [L], [f], [q0])
    csr = r["csr"]
    e = SystemArrays(n=M.shape[0], csr=(csr.data.copy(), csr.indices.astype(np.int64),
                                        csr.indptr.astype(np.int64)),
                     gamma=_tens(r["gamma"], 3), lam=_tens(r["lam"], 4),
                     force=r["force"], q0=r["q0"])
    if adjoint:
        n = M.shape[0]
        params = [("K", d + i, i) for i in range(d)] + \
                 [("Lambda", t) for t in range(len(np.asarray(L.val).reshape(-1)))]
        term = np.zeros(n)
        term[-1] = 1.0
        Mp, Bp, Kp, Gp, Lp, fp, lT, Pp = ADJ.system_jacobian(M, B, K, G, L, params, term)
        ra = assemble_arrays([Mp], [Bp], [Kp], [Gp], [Lp], [fp], [lT], [Pp], adjoint=True)
        ac = ra["csr"]
        e.na = Mp.shape[0]
        e.acsr = (ac.data.copy(), ac.indices.astype(np.int64), ac.indptr.astype(np.int64))
        e.agamma = _tens(ra["gamma"], 3)
        e.alam = _tens(ra["lam"], 4)
        e.apsi = _tens(ra["psi"], 5)
        e.lamT = ra["q0"]
    return e


class SystemLibrary:
    """Pre-assembled chain systems of each DOF size with the energy-gradient
    adjoint (dE(T)/dK_ii, dE(T)/dLambda_i), ``variants`` per size."""

    def __init__(self, dofs: Sequence[int] = tuple(W.DOF_CHOICES), variants: int = 8,
                 seed: int = 4, adjoint: bool = True):
        rng = np.random.default_rng(seed)
        self.entries: Dict[int, List[SystemArrays]] = {}
        for d in dofs:
            lst = []
            for _ in range(variants):
                s = rng.uniform(0.9, 1.1)
                gv = rng.normal(0.0, 0.05, 2 * d)
                lv = rng.uniform(0.0, 1.0, d)
                lst.append(chain_arrays(int(d), s, gv, lv, adjoint))
            self.entries[int(d)] = lst
        self.flat: List[SystemArrays] = [e for d in sorted(self.entries)
                                         for e in self.entries[d]]
        for k, e in enumerate(self.flat):
            e.key = k


def _cat_csr(parts, offs):
    """Block-diagonal concatenation of CSR parts (data, indices, indptr), the
    column indices of part k shifted by offs[k]."""
    datas, inds, ptrs = zip(*parts)
    data = np.concatenate(datas)
    nnz = np.array([len(x) for x in datas], dtype=np.int64)
    ind = np.concatenate(inds)
    ind = (ind if ind.dtype == np.int64 else ind.astype(np.int64))
    ind += np.repeat(np.asarray(offs, dtype=np.int64), nnz)
    rows = np.array([len(x) - 1 for x in ptrs], dtype=np.int64)
    base = np.cumsum(nnz) - nnz
    ptr = np.empty(int(rows.sum()) + 1, dtype=np.int64)
    ptr[0] = 0
    ptr[1:] = np.concatenate([x[1:] for x in ptrs])
    ptr[1:] += np.repeat(base, rows)
    return data, ind, ptr


def _cat_tens(parts, offs_a, offs_b, n_a_axes):
    """indices: the first n_a_axes axes offset by offs_a, the rest by offs_b"""
    idxs, vals = zip(*parts)
    cnt = np.array([len(v) for v in vals], dtype=np.int64)
    idx = np.concatenate(idxs, axis=0)
    idx = idx if idx.dtype == np.int64 else idx.astype(np.int64)
    val = np.concatenate(vals)
    if idx.shape[0]:
        idx[:, :n_a_axes] += np.repeat(np.asarray(offs_a, dtype=np.int64), cnt)[:, None]
        idx[:, n_a_axes:] += np.repeat(np.asarray(offs_b, dtype=np.int64), cnt)[:, None]
    return idx, val


@dataclass
class Request:
    rid: int
    systems: List[SystemArrays]
    exc_ids: np.ndarray
    steps: np.ndarray
    adjoint: bool
    arrival: float = 0.0
    admitted: float = 0.0
    done: float = 0.0
    result: Optional[np.ndarray] = None
    grad: Optional[np.ndarray] = None
    keys: Optional[np.ndarray] = None     # library positions of the systems
    nq: int = 0                           # output lengths (final states, gradient)
    ng: int = 0

    def __post_init__(self):
        if self.systems is not None and not self.nq and \
                all(hasattr(x, "n") for x in self.systems):
            self.nq = sum(int(x.n) for x in self.systems)
            self.ng = sum(int(x.na) for x in self.systems) if self.adjoint else 0

    @property
    def system_steps(self) -> int:
        return int(np.sum(self.steps))


class RequestStream:
    """``n`` random requests of library systems, the distributions of
    ``draw_request``, drawn vectorised in blocks of ``block`` requests (each
    block from its own seed (seed, block) on first access, so an overloaded
    stream never builds the blocks the pool does not reach); the k-th is
    served as a :class:`Request` of array slices."""

    def __init__(self, lib: SystemLibrary, n: int, seed: int, n_signals: int, tmin=1e3,
                 tmax=1e5, max_systems: int = 256, adjoint_frac: float = 0.7,
                 block: int = 4096):
        self.lib, self.n, self.seed, self.block = lib, n, seed, block
        self.kw = dict(n_signals=n_signals, tmin=tmin, tmax=tmax, max_systems=max_systems,
                       adjoint_frac=adjoint_frac)
        self.lib_n = np.array([e.n for e in lib.flat], dtype=np.int64)
        self.lib_na = np.array([getattr(e, "na", 0) for e in lib.flat], dtype=np.int64)
        self._blocks = {}

    def __len__(self):
        return self.n

    def _draw(self, b: int) -> dict:
        idx, probs, nvar, ents, keys = _draw_tables(self.lib)
        kw = self.kw
        rng = np.random.default_rng((self.seed, b))
        m = min(self.block, self.n - b * self.block)
        cnt = rng.integers(1, kw["max_systems"] + 1, m)
        off = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
        tot = int(off[-1])
        di = rng.choice(idx, size=tot, p=probs)
        var = (rng.uniform(size=tot) * nvar[di]).astype(np.int64)
        first = np.concatenate([[0], np.cumsum([len(k) for k in keys])])[:-1]
        k = np.concatenate(keys)[first[di] + var]
        steps = np.exp(rng.uniform(np.log(kw["tmin"]), np.log(kw["tmax"]), tot)).astype(np.int64)
        exc = rng.integers(0, kw["n_signals"], tot)
        adj = rng.uniform(size=m) < kw["adjoint_frac"]
        cq = np.concatenate([[0], np.cumsum(self.lib_n[k])])
        cg = np.concatenate([[0], np.cumsum(self.lib_na[k])])
        nq = cq[off[1:]] - cq[off[:-1]]
        ng = np.where(adj, cg[off[1:]] - cg[off[:-1]], 0)
        return {"off": off, "keys": k, "steps": steps, "exc": exc, "adj": adj, "nq": nq,
                "ng": ng}

    def __getitem__(self, k: int) -> Request:
        b, i = divmod(int(k), self.block)
        blk = self._blocks.get(b)
        if blk is None:
            blk = self._blocks[b] = self._draw(b)
        a, e = int(blk["off"][i]), int(blk["off"][i + 1])
        return Request(int(k), None, blk["exc"][a:e], blk["steps"][a:e], bool(blk["adj"][i]),
                       keys=blk["keys"][a:e], nq=int(blk["nq"][i]), ng=int(blk["ng"][i]))


def _draw_tables(lib: SystemLibrary):
    """The library's sizes with their probabilities, variant counts and
    entries, computed once per library."""
    t = getattr(lib, "_draw", None)
    if t is None:
        sizes = [int(d) for d in W.DOF_CHOICES if int(d) in lib.entries]
        probs = np.array([p for d, p in zip(W.DOF_CHOICES, W.DOF_PROBS) if int(d) in lib.entries])
        ents = [lib.entries[d] for d in sizes]
        keys = [np.array([e.key for e in es], dtype=np.int64) for es in ents]
        t = (np.arange(len(sizes)), probs / probs.sum(),
             np.array([len(es) for es in ents], dtype=np.int64), ents, keys)
        lib._draw = t
    return t


def draw_request(lib: SystemLibrary, rng, rid: int, n_signals: int, tmin=1e3, tmax=1e5,
                 max_systems: int = 256, adjoint_frac: float = 0.7) -> Request:
    """A request of 1..max_systems library systems (sizes by the DOF
    probabilities, a uniform variant of each), log-uniform step counts on
    [tmin, tmax] and uniform signal ids; forward+adjoint with probability
    ``adjoint_frac``."""
    idx, probs, nvar, ents, keys = _draw_tables(lib)
    n = int(rng.integers(1, max_systems + 1))
    di = rng.choice(idx, size=n, p=probs)
    var = rng.integers(0, nvar[di])
    systems = [ents[a][b] for a, b in zip(di.tolist(), var.tolist())]
    steps = np.exp(rng.uniform(np.log(tmin), np.log(tmax), n)).astype(np.int64)
    k = np.array([keys[a][b] for a, b in zip(di.tolist(), var.tolist())], dtype=np.int64)
    return Request(rid, systems, rng.integers(0, n_signals, n), steps,
                   bool(rng.uniform() < adjoint_frac), keys=k)


class FreshSystems:
    """``n`` systems none of which a pool has registered: each a chain of a
    DOF size drawn with the heterogeneous probabilities, coefficients drawn
    afresh (K scale, Gamma, Lambda) and its sparsity one of ``variants``
    structure variants of its size (``workloads.structure_variant``).
    Requests of them carry the full arrays; the pool packs them into tiles.

    ``like``: a :class:`SystemLibrary` whose sparsity structures the systems
    keep (its DOF sizes, the base structure of each), with coefficients
    drawn afresh: the pool recognises the structures and runs such systems on
    its lane modules."""

    def __init__(self, n: int, seed: int = 9, variants: int = 12, adjoint: bool = True,
                 like: Optional["SystemLibrary"] = None):
        rng = np.random.default_rng(seed)
        choices = np.asarray(W.DOF_CHOICES)
        probs = np.asarray(W.DOF_PROBS, dtype=np.float64)
        if like is not None:
            keep = np.isin(choices, np.array(sorted(like.entries)))
            choices, probs = choices[keep], probs[keep]
        self.dofs = rng.choice(choices, size=n, p=probs / probs.sum())
        self.systems: List[SystemArrays] = []
        for d in self.dofs.tolist():
            v = 0 if like is not None else int(rng.integers(0, W.variant_count(int(d), variants)))
            s = rng.uniform(0.9, 1.1)
            gv = rng.normal(0.0, 0.05, 2 * d)
            lv = rng.uniform(0.0, 1.0, d)
            self.systems.append(chain_arrays(int(d), s, gv, lv, adjoint, variant=v))


class FreshRequestStream:
    """``n`` random requests of :class:`FreshSystems` (1..max_systems systems
    drawn uniformly from the bank, log-uniform steps, uniform signals,
    forward+adjoint with probability ``adjoint_frac``), each drawn from its
    own seed (seed, k) on first access."""

    def __init__(self, bank: FreshSystems, n: int, seed: int, n_signals: int, tmin=1e3,
                 tmax=1e5, max_systems: int = 256, adjoint_frac: float = 0.7):
        self.bank, self.n, self.seed = bank, n, seed
        self.kw = dict(n_signals=n_signals, tmin=tmin, tmax=tmax, max_systems=max_systems,
                       adjoint_frac=adjoint_frac)

    def __len__(self):
        return self.n

    def __getitem__(self, k: int) -> Request:
        kw = self.kw
        rng = np.random.default_rng((self.seed, int(k)))
        m = int(rng.integers(1, kw["max_systems"] + 1))
        pick = rng.integers(0, len(self.bank.systems), m)
        steps = np.exp(rng.uniform(np.log(kw["tmin"]), np.log(kw["tmax"]), m)).astype(np.int64)
        exc = rng.integers(0, kw["n_signals"], m)
        adj = bool(rng.uniform() < kw["adjoint_frac"])
        return Request(int(k), [self.bank.systems[i] for i in pick.tolist()], exc, steps, adj)


class MixedRequests:
    """Request k from ``fresh`` with probability ``fresh_frac`` (decided by
    its own seed), from ``registered`` otherwise."""

    def __init__(self, registered, fresh, fresh_frac: float, seed: int):
        self.a, self.b = registered, fresh
        n = min(len(registered), len(fresh))
        self.is_fresh = np.random.default_rng(seed).uniform(size=n) < fresh_frac

    def __len__(self):
        return len(self.is_fresh)

    def __getitem__(self, k: int) -> Request:
        return self.b[k] if self.is_fresh[k] else self.a[k]


# --------------------------------------------------------------- pool
class Pool:
    """Device-resident pool of jobs advanced together, one launch per tick."""

    def __init__(self, lanes: int = 512, slot_config: int = 1, checkpoint_every: int = 256,
                 steps_per_tick: int = 2000, hbm_budget: float = 200e9, precision: int = 64,
                 gpu: int = 0):
        self.lib = _load()
        self.ids = None
        self.refusals = Counter()       # refused group submissions by reason
        # slots and HBM bytes a registered system took in admitted groups, by
        # adjoint flag (running estimates: a group predicted not to fit is
        # not packed)
        # keyed by (registered ids, adjoint flag)
        self._slots_per_system = {}
        self._bytes_per_system = {}
        self.hbm_budget = float(hbm_budget)
        h = C.c_void_p()
        rc = self.lib.pasif_pool_create(C.byref(h), gpu, precision, lanes, slot_config,
                                        checkpoint_every, steps_per_tick, float(hbm_budget))
        self.h = h
        if rc:
            msg = self.lib.pasif_pool_last_error(h)
            self.lib.pasif_pool_destroy(h)
            self.h = None
            raise _lib.PasifError(rc, msg.decode() if msg else "")
        self.steps_per_tick = steps_per_tick

    def _chk(self, rc):
        if rc:
            raise _lib.PasifError(rc, self.lib.pasif_pool_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.lib.pasif_pool_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_bank(self, exc: np.ndarray, sample_rate: int, interp: Optional[np.ndarray]):
        e = np.ascontiguousarray(exc, dtype=np.float64)
        m = None if interp is None else np.ascontiguousarray(interp[:2], dtype=np.float64)
        self._chk(self.lib.pasif_pool_set_bank(
            self.h, e.ctypes.data_as(_f64p), e.shape[0], e.shape[1], int(sample_rate),
            None if m is None else m.ctypes.data_as(_f64p), 0 if m is None else m.shape[1]))

    def _job(self):
        j = C.c_void_p()
        self._chk(self.lib.pasif_pool_job_create(self.h, C.byref(j), None, 0, 0, 0, 0))
        d = object.__new__(_Driver)
        d._pl = self.lib
        d._h = j
        d._layout = {"n": 0, "sizes_f": None, "sizes_a": None, "exc": None, "steps": None}
        return d

    @staticmethod
    def _load_systems(d, sy: Sequence[SystemArrays], adjoint: bool):
        """Concatenate systems' arrays into driver handle ``d`` (block-diagonal)."""
        sizes = np.array([s.n for s in sy], dtype=np.int64)
        offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
        n = int(sizes.sum())
        data, ind, ptr = _cat_csr([s.csr for s in sy], offs)
        d._setFwdK((n, n), data, ind, ptr)
        gi, gv = _cat_tens([s.gamma for s in sy], offs, offs, 3)
        d._setFwdGamma((n, n, n), gv, gi.reshape(-1))
        li, lv = _cat_tens([s.lam for s in sy], offs, offs, 4)
        d._setFwdLambda((n, n, n, n), lv, li.reshape(-1))
        d._setFwdPsi((1, 1, 1, 1, 1), [], [])
        d._setFwdForcePattern(np.concatenate([s.force for s in sy]))
        d._setFwdInitialConditions(np.concatenate([s.q0 for s in sy]))
        sizes_a = None
        if adjoint:
            sizes_a = np.array([s.na for s in sy], dtype=np.int64)
            aoffs = np.concatenate([[0], np.cumsum(sizes_a)[:-1]])
            na = int(sizes_a.sum())
            data, ind, ptr = _cat_csr([s.acsr for s in sy], aoffs)
            d._setBwdK((na, na), data, ind, ptr)
            gi, gv = _cat_tens([s.agamma for s in sy], aoffs, offs, 2)
            d._setBwdGamma((na, na, n), gv, gi.reshape(-1))
            li, lv = _cat_tens([s.alam for s in sy], aoffs, offs, 2)
            d._setBwdLambda((na, na, n, n), lv, li.reshape(-1))
            pi, pv = _cat_tens([s.apsi for s in sy], aoffs, offs, 2)
            d._setBwdPsi((na, na, n, n, n), pv, pi.reshape(-1))
            d._setBwdForcePattern(np.zeros(na))
            d._setBwdInitialConditions(np.concatenate([s.lamT for s in sy]))
        return sizes, sizes_a

    def submit(self, req: Request):
        """Load and admit a request with its full system arrays; returns the
        job or None when the slots or the HBM budget are exhausted (retry
        after collecting)."""
        d = self._job()
        sizes, sizes_a = self._load_systems(d, req.systems, req.adjoint)
        d._setSystemsLayout(sizes_fwd=sizes, sizes_adj=sizes_a, exc_of_system=req.exc_ids,
                            steps_of_system=req.steps)
        d._allocateSystemOnDevice()
        if req.adjoint:
            d._allocateAdjointOnDevice()
        rc = self.lib.pasif_pool_submit(self.h, d._h, int(req.adjoint))
        if rc == 4:          # slots / HBM budget: keep queued
            self.lib.pasif_destroy(d._h)
            d._h = None
            return None
        self._chk(rc)
        return d

    def register_library(self, lib: SystemLibrary):
        """Register every library system (with its adjoint) with the pool, so
        requests of library systems travel as ids (``submit_ids``)."""
        host = _Driver([np.zeros(2)], 1, 1, GPUId_=-1)      # host-only handle
        sizes, sizes_a = self._load_systems(host, lib.flat, True)
        host._setSystemsLayout(sizes_fwd=sizes, sizes_adj=sizes_a)
        host._allocateSystemOnDevice()
        host._allocateAdjointOnDevice()
        first, count = C.c_uint32(0), C.c_uint32(0)
        try:
            self._chk(self.lib.pasif_pool_register(self.h, host._h, C.byref(first),
                                                   C.byref(count)))
        finally:
            self.lib.pasif_destroy(host._h)
            host._h = None
        if count.value != len(lib.flat):
            raise RuntimeError(f"registered {count.value} systems, the library has "
                               f"{len(lib.flat)}")
        self.ids = first.value + np.arange(count.value, dtype=np.uint32)

    def submit_ids(self, req: Request):
        """Admit a request of registered library systems: one C call carries
        the system ids, excitation ids and step counts.  None when the slots
        or the HBM budget are exhausted."""
        ids = np.ascontiguousarray(self.ids[req.keys], dtype=np.uint32)
        exc = np.ascontiguousarray(req.exc_ids, dtype=np.int32)
        steps = np.ascontiguousarray(req.steps, dtype=np.uint32)
        j = C.c_void_p()
        rc = self.lib.pasif_pool_submit_systems(
            self.h, len(ids), ids.ctypes.data_as(_u32p), exc.ctypes.data_as(_i32p),
            steps.ctypes.data_as(_u32p), int(req.adjoint), C.byref(j))
        if rc == 4:
            return None
        self._chk(rc)
        d = object.__new__(_Driver)
        d._pl = self.lib
        d._h = j
        return d

    def submit_batch(self, reqs: Sequence[Request]) -> list:
        """Admit requests of registered systems in order with one C call;
        returns the jobs of those admitted (a prefix of ``reqs``: the first
        that does not fit the slots or the HBM budget stops the batch)."""
        n = len(reqs)
        if n == 0:
            return []
        cnt = np.array([len(r.keys) for r in reqs], dtype=np.int64)
        off = np.ascontiguousarray(np.concatenate([[0], np.cumsum(cnt)]), dtype=np.uint32)
        ids = np.ascontiguousarray(self.ids[np.concatenate([r.keys for r in reqs])],
                                   dtype=np.uint32)
        exc = np.ascontiguousarray(np.concatenate([r.exc_ids for r in reqs]), dtype=np.int32)
        steps = np.ascontiguousarray(np.concatenate([r.steps for r in reqs]), dtype=np.uint32)
        adj = np.ascontiguousarray([int(r.adjoint) for r in reqs], dtype=np.int32)
        jobs = (C.c_void_p * n)()
        got = C.c_uint32(0)
        rc = self.lib.pasif_pool_submit_batch(
            self.h, n, off.ctypes.data_as(_u32p), ids.ctypes.data_as(_u32p),
            exc.ctypes.data_as(_i32p), steps.ctypes.data_as(_u32p), adj.ctypes.data_as(_i32p),
            jobs, C.byref(got))
        out = []
        for i in range(got.value):
            d = object.__new__(_Driver)
            d._pl = self.lib
            d._h = C.c_void_p(jobs[i])
            out.append(d)
        self._chk(rc)
        return out

    def set_compile(self, on: bool = True):
        """Compile the structures of array-carrying requests that no lane
        module runs into further modules on a background thread (later
        requests of those structures run on them)."""
        self._chk(self.lib.pasif_pool_set_compile(self.h, int(bool(on))))

    def compile_wait(self, max_builds: int = 0):
        """Block until every queued structure is compiled (``max_builds`` > 0:
        at most that many module compilations); returns (modules built,
        builds failed, structures still queued)."""
        m, f, q = C.c_uint32(0), C.c_uint32(0), C.c_uint32(0)
        self._chk(self.lib.pasif_pool_compile_wait(self.h, int(max_builds), C.byref(m),
                                                   C.byref(f), C.byref(q)))
        return m.value, f.value, q.value

    def slots(self):
        """(slots held now, slot cap): tiles and lane workgroups ticks still
        launch; cap 0 = none."""
        u, m = C.c_uint32(0), C.c_uint32(0)
        self._chk(self.lib.pasif_pool_slots(self.h, C.byref(u), C.byref(m)))
        return u.value, m.value

    def group_fit(self, reqs: Sequence[Request]) -> int:
        """How many leading requests of ``reqs`` (one adjoint flag) the free
        slots and HBM budget are expected to take as one group, from the
        slots and bytes per system earlier groups took (no limit before any
        group was admitted; at least one into an empty pool)."""
        if not reqs:
            return 0
        flag = (reqs[0].keys is not None, bool(reqs[0].adjoint))
        used, cap = self.slots()
        free = float("inf")
        per = self._slots_per_system.get(flag)
        if cap and per:
            free = (cap - used) / (1.05 * per)
        per_b = self._bytes_per_system.get(flag)
        if per_b:
            free = min(free, (self.hbm_budget - self.stats()["used_bytes"]) / (1.05 * per_b))
        n, tot = 0, 0
        for r in reqs:
            tot += len(r.steps)
            if tot > free:
                break
            n += 1
        return max(n, 1) if used == 0 else n

    def submit_group(self, reqs: Sequence[Request]):
        """Admit requests of one kind (registered ids, or carrying their own
        arrays) and one adjoint flag as one job: the packer sorts all their
        systems by step count together and fills each structure class's
        waves across requests, instead of one request's log-uniform spread
        (and, for array requests, a few systems per class) per wave.  None
        when the slots or the HBM budget do not take them all."""
        used0 = self.slots()[0]
        bytes0 = self.stats()["used_bytes"]
        exc = np.ascontiguousarray(np.concatenate([r.exc_ids for r in reqs]), dtype=np.int32)
        steps = np.ascontiguousarray(np.concatenate([r.steps for r in reqs]), dtype=np.uint32)
        adjoint = bool(reqs[0].adjoint)
        if reqs[0].keys is None:
            d = self._job()
            sizes, sizes_a = self._load_systems(d, [x for r in reqs for x in r.systems], adjoint)
            d._setSystemsLayout(sizes_fwd=sizes, sizes_adj=sizes_a, exc_of_system=exc,
                                steps_of_system=steps)
            d._allocateSystemOnDevice()
            if adjoint:
                d._allocateAdjointOnDevice()
            rc = self.lib.pasif_pool_submit(self.h, d._h, int(adjoint))
            if rc == 4:
                self.refusals[self.lib.pasif_pool_last_error(self.h).decode().split(":")[0]] += 1
                self.lib.pasif_destroy(d._h)
                d._h = None
                return None
            self._chk(rc)
        else:
            ids = np.ascontiguousarray(self.ids[np.concatenate([r.keys for r in reqs])],
                                       dtype=np.uint32)
            j = C.c_void_p()
            rc = self.lib.pasif_pool_submit_systems(
                self.h, len(ids), ids.ctypes.data_as(_u32p), exc.ctypes.data_as(_i32p),
                steps.ctypes.data_as(_u32p), int(adjoint), C.byref(j))
            if rc == 4:
                self.refusals[self.lib.pasif_pool_last_error(self.h).decode().split(":")[0]] += 1
                return None
            self._chk(rc)
            d = object.__new__(_Driver)
            d._pl = self.lib
            d._h = j
        flag = (reqs[0].keys is not None, adjoint)
        for est, now, before in ((self._slots_per_system, self.slots()[0], used0),
                                 (self._bytes_per_system, self.stats()["used_bytes"], bytes0)):
            per = (now - before) / len(steps)
            est[flag] = per if est.get(flag) is None else 0.7 * est[flag] + 0.3 * per
        return d

    def collect_groups(self, jobs: Sequence, groups: Sequence[Sequence[Request]]):
        """Collect finished jobs (each a group of requests, in submission
        order) with one C call; sets ``result`` / ``grad`` of every request."""
        n = len(jobs)
        if n == 0:
            return
        nq = np.array([sum(r.nq for r in g) for g in groups], dtype=np.int64)
        ng = np.array([sum(r.ng for r in g) if g[0].adjoint else 0 for g in groups],
                      dtype=np.int64)
        q_off = np.ascontiguousarray(np.concatenate([[0], np.cumsum(nq)]), dtype=np.uint64)
        g_off = np.ascontiguousarray(np.concatenate([[0], np.cumsum(ng)]), dtype=np.uint64)
        q = np.empty(int(q_off[-1]))
        g = np.empty(max(int(g_off[-1]), 1))
        hs = (C.c_void_p * n)(*[j._h for j in jobs])
        szp = C.POINTER(C.c_size_t)
        rc = self.lib.pasif_pool_collect_batch(self.h, n, hs, q.ctypes.data_as(_f64p),
                                               q_off.ctypes.data_as(szp), g.ctypes.data_as(_f64p),
                                               g_off.ctypes.data_as(szp))
        for j in jobs:
            j._h = None
        self._chk(rc)
        for i, grp in enumerate(groups):
            a, b = int(q_off[i]), int(g_off[i])
            for r in grp:
                r.result = q[a:a + r.nq]
                a += r.nq
                if r.adjoint:
                    r.grad = g[b:b + r.ng]
                    b += r.ng
                else:
                    r.grad = None

    def collect_batch(self, jobs: Sequence, reqs: Sequence[Request]):
        """Collect finished jobs with one C call (their handles are destroyed):
        sets ``result`` / ``grad`` of each request."""
        n = len(jobs)
        if n == 0:
            return
        nq = np.array([r.nq for r in reqs], dtype=np.int64)
        ng = np.array([r.ng if r.adjoint else 0 for r in reqs], dtype=np.int64)
        q_off = np.ascontiguousarray(np.concatenate([[0], np.cumsum(nq)]), dtype=np.uint64)
        g_off = np.ascontiguousarray(np.concatenate([[0], np.cumsum(ng)]), dtype=np.uint64)
        q = np.empty(int(q_off[-1]))
        g = np.empty(max(int(g_off[-1]), 1))
        hs = (C.c_void_p * n)(*[j._h for j in jobs])
        szp = C.POINTER(C.c_size_t)
        rc = self.lib.pasif_pool_collect_batch(self.h, n, hs, q.ctypes.data_as(_f64p),
                                               q_off.ctypes.data_as(szp), g.ctypes.data_as(_f64p),
                                               g_off.ctypes.data_as(szp))
        for j in jobs:
            j._h = None
        self._chk(rc)
        for i, r in enumerate(reqs):
            r.result = q[q_off[i]:q_off[i + 1]]
            r.grad = g[g_off[i]:g_off[i + 1]] if r.adjoint else None

    def set_slots(self, max_tiles: int):
        self._chk(self.lib.pasif_pool_set_slots(self.h, int(max_tiles)))

    def running(self) -> bool:
        r = C.c_int(0)
        self._chk(self.lib.pasif_pool_poll(self.h, C.byref(r)))
        return bool(r.value)

    def tick(self, wait: bool) -> int:
        n = C.c_uint32(0)
        self._chk(self.lib.pasif_pool_tick(self.h, int(wait), C.byref(n)))
        return n.value

    def wait(self) -> float:
        ms = C.c_double(0)
        self._chk(self.lib.pasif_pool_wait(self.h, C.byref(ms)))
        return ms.value

    def finished(self, d) -> bool:
        f = C.c_int(0)
        self._chk(self.lib.pasif_pool_job_finished(self.h, d._h, C.byref(f)))
        return bool(f.value)

    def collect(self, d, req: "Request"):
        adjoint = req.adjoint
        lq, lg = C.c_size_t(0), C.c_size_t(0)
        q = np.empty(req.nq)
        g = np.empty(req.ng if adjoint else 0)
        self._chk(self.lib.pasif_pool_collect(self.h, d._h, q.ctypes.data_as(_f64p), q.size,
                                              C.byref(lq), g.ctypes.data_as(_f64p), g.size,
                                              C.byref(lg)))
        self.lib.pasif_destroy(d._h)
        d._h = None
        return q[:lq.value], (g[:lg.value] if adjoint else None)

    def drop(self, d):
        """Remove an unfinished job from the pool (its results are discarded)."""
        if d._h:
            self.lib.pasif_destroy(d._h)
            d._h = None

    def stats(self) -> dict:
        j, b, ms, blk = C.c_uint32(0), C.c_double(0), C.c_double(0), C.c_uint32(0)
        self._chk(self.lib.pasif_pool_stats(self.h, C.byref(j), C.byref(b), C.byref(ms),
                                            C.byref(blk)))
        return {"jobs": j.value, "used_bytes": b.value, "kernel_ms": ms.value,
                "blocks": blk.value}


# --------------------------------------------------------------- server
@dataclass
class ServeReport:
    offered_rps: float
    duration_s: float
    completed: int
    admitted: int
    arrived: int
    p50_ms: float
    p99_ms: float
    mean_ms: float
    system_steps_per_s: float
    ticks: int
    gpu_busy_frac: float
    latencies_ms: List[float] = field(default_factory=list)
    backlog: int = 0
    p99_all_ms: float = 0.0
    admit_ms_mean: float = 0.0    # host time per admitted request (packing, upload)


class SolveServer:
    """``group=True`` (default): queued requests of registered systems are
    admitted together, as pool jobs of one adjoint flag and at most
    ``group_systems`` systems each (``Pool.submit_group``), so their systems
    are sorted by step count across requests and a wave's systems run
    similar lengths; every request of a group completes when the group does.
    ``group=False``: one job per request (``submit_batch``)."""

    def __init__(self, pool: Pool, group: bool = True, group_systems: int = 4096):
        self.pool = pool
        self.group = group
        self.group_systems = group_systems

    def run(self, requests: Sequence[Request], horizon_s: Optional[float] = None,
            drain: bool = True, max_wall_s: float = 600.0,
            arrivals: Optional[np.ndarray] = None) -> ServeReport:
        """Serve ``requests`` (arrival times relative to start).  Requests that
        arrive while a tick runs are admitted during it (host work overlaps the
        launch) and join at the next tick.

        With ``arrivals`` (sorted arrival times) ``requests`` may be a lazy
        sequence: ``requests[k]`` is only touched when request ``k`` is
        submitted, so an overloaded stream does not build requests the pool
        never reaches.  Latency percentiles cover completed requests; requests
        still queued or in flight at the end are reported as ``backlog`` with
        their waiting time so far in ``p99_all_ms`` (a lower bound).  Jobs
        still resident when the run stops (horizon or ``max_wall_s``) are
        removed from the pool, so a later run starts from an empty pool."""
        if arrivals is None:
            reqs = sorted(requests, key=lambda r: r.arrival)
            arr = np.array([r.arrival for r in reqs], dtype=np.float64)
        else:
            reqs = requests
            arr = np.asarray(arrivals, dtype=np.float64)
        n_req = len(arr)
        queue = deque()             # indices of arrived, not yet admitted requests
        done: List[Request] = []
        i = 0
        t0 = time.perf_counter()
        ticks = 0
        busy_ms = 0.0
        admitted = 0
        blocked = False             # a submit was refused; retry after a tick or a collect
        admit_s = 0.0
        by_id = self.pool.ids is not None

        pending = []                # finished after the last tick, not collected yet
        nres = [0]                  # resident (admitted, not finished) jobs
        # admitted jobs by the tick after which their work is covered: a job
        # of at most T steps per system needs ceil(T (3 with the adjoint) /
        # steps_per_tick) ticks, so only the jobs due are asked whether they
        # finished
        due = {}                    # tick -> [(job, request)]

        def arrivals_until(now):
            nonlocal i
            while i < n_req and arr[i] <= now:
                queue.append(i)
                i += 1

        batched = by_id and hasattr(self.pool, "submit_batch")
        grouped = batched and self.group and hasattr(self.pool, "submit_group")
        # (array-carrying requests are grouped too, among themselves)
        # due / pending entries are (job, [requests]): one request per job, or
        # a group of them admitted as one job

        def collect_pending():
            nonlocal blocked
            if not pending:
                return
            if batched:
                if hasattr(self.pool, "collect_groups"):
                    self.pool.collect_groups([j for j, _ in pending], [m for _, m in pending])
                else:
                    self.pool.collect_batch([j for j, _ in pending], [m[0] for _, m in pending])
                now = time.perf_counter() - t0
                for _, m in pending:
                    for r in m:
                        r.done = now
                        done.append(r)
            else:
                for job, m in pending:
                    r = m[0]
                    r.result, r.grad = self.pool.collect(job, r)
                    r.done = time.perf_counter() - t0
                    done.append(r)
            blocked = False
            pending.clear()

        def admit_batch(limit: int = 256) -> bool:
            """Admit queued requests in one call; False when nothing was
            admitted (none queued, or the pool is full)."""
            nonlocal admitted, blocked, admit_s
            if blocked or not queue:
                return False
            first = reqs[queue[0]]
            if first.keys is None and not grouped:    # carries its arrays: one submit
                return admit_one()
            ta = time.perf_counter()
            # the queued requests, each fetched once (a stream may build a new
            # object per access)
            take, obj = [], {}
            for k in range(min(limit, len(queue))):
                r = first if k == 0 else reqs[queue[k]]
                if r.keys is None and not grouped:
                    break
                r.arrival = float(arr[queue[k]])
                take.append(queue[k])
                obj[queue[k]] = r
            entries = []
            if grouped:
                # per kind (registered ids / own arrays) and adjoint flag,
                # groups of consecutive queued requests of at most
                # group_systems systems (a single larger request is a group of
                # its own), cut to what the pool is expected to take; a group
                # the pool refuses is halved until it fits, and the admission
                # of that kind and flag stops at the first cut.  Array requests
                # cost host time to assemble: one group of them per call, so
                # the loop gets back to launching ticks.
                got = set()
                fit = getattr(self.pool, "group_fit", None)
                for by_id, flag in ((True, False), (True, True), (False, False), (False, True)):
                    members = [k for k in take if bool(obj[k].adjoint) == flag and
                               (obj[k].keys is not None) == by_id]
                    stop = False
                    while members and not stop:
                        n, tot = 0, 0
                        for k in members:
                            tot += len(obj[k].steps)
                            if n and tot > self.group_systems:
                                break
                            n += 1
                        chunk = members[:n]
                        if fit is not None:
                            m = fit([obj[k] for k in chunk])
                            stop = m < len(chunk)
                            chunk = chunk[:m]
                        job = None
                        while chunk:
                            job = self.pool.submit_group([obj[k] for k in chunk])
                            if job is not None:
                                break
                            stop = True
                            chunk = chunk[:len(chunk) // 2]
                        if not chunk:
                            break
                        entries.append((job, [obj[k] for k in chunk]))
                        got.update(chunk)
                        members = members[len(chunk):]
                        stop = stop or not by_id
                for _ in take:
                    queue.popleft()
                for k in reversed(take):
                    if k not in got:
                        queue.appendleft(k)
                if len(got) < len(take):
                    blocked = True
            else:
                batch = [obj[k] for k in take]
                jobs = self.pool.submit_batch(batch)
                entries = [(j, [r]) for j, r in zip(jobs, batch)]
                for _ in jobs:
                    queue.popleft()
                if len(jobs) < len(batch):
                    blocked = True
            admit_s += time.perf_counter() - ta
            now = time.perf_counter() - t0
            for job, members in entries:
                work = max(int(np.max(r.steps)) for r in members) * \
                    (3 if members[0].adjoint else 1)
                ready = ticks + max(1, -(-work // self.pool.steps_per_tick))
                due.setdefault(ready, []).append((job, members))
                nres[0] += 1
                for r in members:
                    r.admitted = now
                admitted += len(members)
            return len(entries) > 0

        def admit_one() -> bool:
            """Admit the oldest queued request; False if none or refused."""
            nonlocal admitted, blocked, admit_s
            if blocked or not queue:
                return False
            ta = time.perf_counter()
            r = reqs[queue[0]]
            r.arrival = float(arr[queue[0]])
            job = (self.pool.submit_ids(r) if by_id and r.keys is not None
                   else self.pool.submit(r))
            admit_s += time.perf_counter() - ta
            if job is None:
                blocked = True
                return False
            r.admitted = time.perf_counter() - t0
            work = int(np.max(r.steps)) * (3 if r.adjoint else 1)
            ready = ticks + max(1, -(-work // self.pool.steps_per_tick))
            due.setdefault(ready, []).append((job, [r]))
            nres[0] += 1
            admitted += 1
            queue.popleft()
            return True

        while True:
            now = time.perf_counter() - t0
            if now > max_wall_s:
                break
            if horizon_s is not None and now > horizon_s and not drain:
                break
            arrivals_until(now)
            admit = admit_batch if batched else admit_one
            if not nres[0]:
                collect_pending()
                while admit():
                    pass
            if nres[0]:
                self.pool.tick(wait=False)
                ticks += 1
                # while the launch runs: collect what the previous tick
                # finished (ticks skip finished jobs), then admit; new jobs
                # join the next tick
                collect_pending()
                while self.pool.running():
                    arrivals_until(time.perf_counter() - t0)
                    if not admit():
                        break
                busy_ms += self.pool.wait()
                blocked = False             # the tick may have freed slots
                for t in [t for t in due if t <= ticks]:
                    for job, r in due.pop(t):
                        if self.pool.finished(job):
                            pending.append((job, r))
                            nres[0] -= 1
                        else:               # not expected: ask again after the next tick
                            due.setdefault(ticks + 1, []).append((job, r))
            elif pending:
                collect_pending()
            elif i >= n_req and not queue:
                break
            else:
                nxt = arr[i] if i < n_req else now
                time.sleep(max(0.0, min(0.01, nxt - now)))
        collect_pending()
        wall = time.perf_counter() - t0
        for lst in due.values():            # unfinished at the stop: leave the pool
            for job, _ in lst:
                self.pool.drop(job)
        lat = np.array([(r.done - r.arrival) * 1e3 for r in done]) if done else np.zeros(1)
        waiting = [(wall - r.arrival) * 1e3 for lst in due.values() for _, m in lst
                   for r in m] + [(wall - arr[k]) * 1e3 for k in queue]
        lat_all = np.concatenate([lat, np.array(waiting)]) if waiting else lat
        steps = sum(r.system_steps for r in done)
        return ServeReport(
            offered_rps=n_req / max(float(arr[-1]), 1e-9) if n_req else 0.0,
            duration_s=wall, completed=len(done), admitted=admitted, arrived=i,
            p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
            mean_ms=float(lat.mean()), system_steps_per_s=steps / wall, ticks=ticks,
            gpu_busy_frac=busy_ms / 1e3 / wall, latencies_ms=list(lat),
            backlog=len(waiting), p99_all_ms=float(np.percentile(lat_all, 99)),
            admit_ms_mean=admit_s * 1e3 / max(admitted, 1))


class LazyRequests:
    """Indexable stream of random requests, each drawn on first access from
    its own seed (``seed``, index), so the k-th request does not depend on
    which others were drawn."""

    def __init__(self, lib: SystemLibrary, n: int, seed: int, n_signals: int, **kw):
        self.lib, self.n, self.seed, self.n_signals, self.kw = lib, n, seed, n_signals, kw

    def __len__(self):
        return self.n

    def __getitem__(self, k: int) -> Request:
        rng = np.random.default_rng((self.seed, int(k)))
        return draw_request(self.lib, rng, int(k), self.n_signals, **self.kw)


def poisson_arrivals(rng, rate: float, duration: float) -> np.ndarray:
    n = rng.poisson(rate * duration)
    return np.sort(rng.uniform(0.0, duration, n))


def signal_bank(rng, n_signals: int, tmax: float, steps_per_sample: int = 10):
    n_samples = int(tmax) // steps_per_sample + 2
    return np.stack([W.band_limited(rng, n_samples) for _ in range(n_signals)])
