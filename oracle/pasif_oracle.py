"""FP64 NumPy restatement of PASIf's batched RK4 forward integrator and its
adjoint recursion.

TEST INFRASTRUCTURE ONLY -- this module is the checker.  It is imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py``; the product (``pasif_amd``) never imports it and never falls
back to it.

Every function cites the reference file:line it restates (paths relative to
the reference repository root).  The reference itself cannot be imported or
compiled here (CUDA sources, cuSPARSE/cuBLAS, an empty pybind11 submodule,
mpi4py), so each part is pinned by closed forms, brute force and invariants in
``tests/test_oracle_*.py``:

=========================  ===================================================
part                       pinned by
=========================  ===================================================
assembly / scaling         scipy closed form ``-(diag(1/m))(B+K)``, dense
                           round trip, reference example systems
RHS                        dense ``einsum`` brute force + pure-Python loops
RK4 step                   matrix polynomial of a linear system, utest KATs
                           (``src/utest.cu:220-296``), forced oscillator
                           convergence against the analytic solution
forcing index rules        hand-written sample/index tables incl. guards
adjoint recursion          linear closed form ``lam_{n-1} = T(hA') lam_n``,
                           pure-Python loop transcription, checkpoint
                           interval invariance, central finite differences
                           (to the scheme's discretisation order)
exact discrete adjoint     the same closed form, central finite differences
                           of the forward oracle to FD precision (1e-9),
                           checkpoint interval invariance
trajectory / windows       slot semantics tables on a linear system
=========================  ===================================================

Deliberate, documented departures from the reference (its defects are not
reproduced; see the tests for the exact cases):

* ``SpTdV``'s Lambda indexing (``src/kernels.cu:55``) drops / repeats cubic
  terms once nnz(Gamma)+nnz(Lambda)+nnz(Psi) > 128.  The oracle applies every
  Lambda entry exactly once.
* The adjoint uses the clean schedule over all steps T..1 with the true
  states x_n, x_{n-1}; the reference's checkpoint off-by-one, skipped first
  chunk and baked scratch slot (``src/__GpuDriver.cu:523-573, 727-779``) are
  not reproduced.
* Interpolated forcing reads ``e[i+1]`` even when ``i == L-1``; the reference
  then reads past the end of that excitation (``src/kernels.cu:124``).  The
  oracle uses 0 for that sample.  Windows W that are not a power of two use
  integer division (the reference's ``log2f`` shift is only defined for
  powers of two, ``src/kernels.cu:115-117``).

Summation order inside one RHS row (fixed, documented): CSR row order of A,
then the Gamma, Lambda, Psi entries in COO order (each kind summed from 0 in
its own accumulator), then ``((((A + G) + L) + Psi) + force)``.  Products are
evaluated left to right as in ``src/kernels.cu:52-61``.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import scipy.sparse as sp

__all__ = [
    "Assembled", "Forcing", "step_constants", "drop_small_forces",
    "assemble_forward", "assemble_adjoint", "assemble_from_arrays",
    "stage_sample_index_zoh", "stage_sample_index_interp",
    "excitation_stage_values", "rhs", "update_slope", "integrate", "rk4_forward", "adjoint_gradient",
    "discrete_adjoint_gradient",
    "InstanceLayout", "cartesian_layout", "mapped_layout",
]


# ---------------------------------------------------------------------------
# constants
# ---------------------------------------------------------------------------
def step_constants(sample_rate: int, dtype=np.float64):
    """h = 1/sampleRate, h2 = h/2, h6 = h/6 exactly as the reference forms
    them (``src/__GpuDriver.cu:632-636``): ``h`` is ``1.0/sampleRate`` rounded
    to the working type, ``h2``/``h6`` are computed from that rounded ``h`` in
    double and rounded again."""
    dt = np.dtype(dtype)
    h = dt.type(1.0 / float(sample_rate))
    h2 = dt.type(float(h) / 2.0)
    h6 = dt.type(float(h) / 6.0)
    return h, h2, h6


def drop_small_forces(force: np.ndarray, dtype=np.float64) -> np.ndarray:
    """COOVector keeps only entries with ``|v| > eps`` of the working type
    (``src/helpers.cu:674-680``, ``reel_eps`` in ``src/helpers.cuh:30``)."""
    f = np.asarray(force, dtype=np.float64).copy()
    eps = float(np.finfo(np.dtype(dtype)).eps)
    f[np.abs(f) <= eps] = 0.0
    return f


# ---------------------------------------------------------------------------
# assembled (global, block-diagonal) system
# ---------------------------------------------------------------------------
@dataclass
class Assembled:
    """One assembled system as the driver receives it.

    ``A`` is the CSR of ``-M^{-1}(B+K)`` with sorted indices
    (``PASIf.py:250-255``); tensors hold ``-M^{-1}``-scaled values and their
    indices as an ``(nnz, order)`` array (the interleaved reference layout
    reshaped, ``PASIf.py:82-83``).  For the forward system ``n_src == n_set ==
    n``.  For an adjoint system ``n_src`` is the adjoint size and ``n_set`` the
    forward size (tensor dims ``[n_adj, n_adj, n_fwd, ...]``,
    ``PASIf.py:469-472``).
    """
    n: int
    A: sp.csr_matrix
    gamma_idx: np.ndarray
    gamma_val: np.ndarray
    lam_idx: np.ndarray
    lam_val: np.ndarray
    psi_idx: np.ndarray
    psi_val: np.ndarray
    force: np.ndarray
    q0: np.ndarray
    n_set: int = 0
    sizes: Optional[np.ndarray] = None      # per-system block sizes

    def __post_init__(self):
        if self.n_set == 0:
            self.n_set = self.n
        if self.sizes is None:
            self.sizes = np.array([self.n], dtype=np.int64)

    def astype(self, dtype) -> "Assembled":
        dt = np.dtype(dtype)
        return Assembled(
            n=self.n, A=self.A.astype(dt),
            gamma_idx=self.gamma_idx, gamma_val=self.gamma_val.astype(dt),
            lam_idx=self.lam_idx, lam_val=self.lam_val.astype(dt),
            psi_idx=self.psi_idx, psi_val=self.psi_val.astype(dt),
            force=drop_small_forces(self.force, dt).astype(dt),
            q0=self.q0.astype(dt), n_set=self.n_set, sizes=self.sizes)


def _coo_tensor_arrays(t, order):
    val = np.asarray(t.val, dtype=np.float64).reshape(-1)
    idx = np.asarray(t.indices, dtype=np.int64).reshape(-1, order) if val.size else \
        np.zeros((0, order), dtype=np.int64)
    return idx, val


def _unfold(vecM, vecB, vecK, tensors, orders, vecF, vecQ0):
    """Block-diagonal concatenation of per-system inputs
    (``PASIf.py:475-543``, ``coo_tensor.concatenateTensor`` 86-100): every
    index axis of system ``s`` is offset by that axis' running size."""
    sizes = np.array([m.shape[0] for m in vecM], dtype=np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    n = int(sizes.sum())
    diag = np.concatenate([np.asarray(m.diagonal(), dtype=np.float64) for m in vecM])

    def cat_mat(vec):
        rows, cols, vals = [], [], []
        for o, m in zip(offs, vec):
            c = sp.coo_matrix(m)
            rows.append(c.row.astype(np.int64) + o)
            cols.append(c.col.astype(np.int64) + o)
            vals.append(np.asarray(c.data, dtype=np.float64))
        return sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                             shape=(n, n))

    B = cat_mat(vecB)
    K = cat_mat(vecK)
    out_t = []
    for tl, order in zip(tensors, orders):
        idxs, vals = [], []
        dim_off = np.zeros(order, dtype=np.int64)
        for t in tl:
            i, v = _coo_tensor_arrays(t, order)
            idxs.append(i + dim_off[None, :])
            vals.append(v)
            dim_off += np.asarray(t.dimensions, dtype=np.int64)
        out_t.append((np.concatenate(idxs, axis=0), np.concatenate(vals), dim_off.copy()))
    f = np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x in vecF])
    q0 = np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x in vecQ0])
    return sizes, diag, B, K, out_t, f, q0


def assemble_from_arrays(diag, B, K, tensors, f, q0, scale_force=True, n_set=None,
                         sizes=None) -> Assembled:
    """Preprocessing ``PASIf.py:607-618`` (forward) / ``620-633`` (adjoint):
    ``D = -M^{-1}``; ``A = (D.B + D.K)`` as sorted CSR (``250-255``);
    every tensor value is multiplied by ``D[first index]``
    (``multiplyByDiagMatrix``, 102-115); the forward force pattern becomes
    ``f * (-D)`` (618), the adjoint one is left unscaled (633)."""
    n = len(diag)
    D = -(1.0 / np.asarray(diag, dtype=np.float64))
    Dm = sp.dia_matrix((D, [0]), shape=(n, n)).tocsc()
    Bs = sp.coo_matrix(Dm.dot(sp.coo_matrix(B)))
    Ks = sp.coo_matrix(Dm.dot(sp.coo_matrix(K)))
    A = (Bs + Ks).tocsr()
    A.sort_indices()
    scaled = []
    for idx, val in tensors:
        v = np.asarray(val, dtype=np.float64) * D[idx[:, 0]] if len(val) else np.zeros(0)
        scaled.append((np.asarray(idx, dtype=np.int64), v))
    fp = np.asarray(f, dtype=np.float64) * (-1.0 * D) if scale_force else np.asarray(f, dtype=np.float64).copy()
    (gi, gv), (li, lv), (pi, pv) = scaled
    return Assembled(n=n, A=A, gamma_idx=gi, gamma_val=gv, lam_idx=li, lam_val=lv,
                     psi_idx=pi, psi_val=pv, force=drop_small_forces(fp),
                     q0=np.asarray(q0, dtype=np.float64).copy(),
                     n_set=int(n_set or n), sizes=sizes)


def assemble_forward(vecM, vecB, vecK, vecGamma, vecLambda, vecF, vecQ0) -> Assembled:
    """``PASIf.setSystems`` (``PASIf.py:202-277``) restated: unfold, scale,
    CSR.  Psi of the forward system is empty (``PASIf.py:271``)."""
    sizes, diag, B, K, tens, f, q0 = _unfold(vecM, vecB, vecK, [vecGamma, vecLambda],
                                             [3, 4], vecF, vecQ0)
    (gi, gv, _), (li, lv, _) = tens
    empty5 = (np.zeros((0, 5), dtype=np.int64), np.zeros(0))
    return assemble_from_arrays(diag, B, K, [(gi, gv), (li, lv), empty5], f, q0,
                                scale_force=True, sizes=sizes)


def assemble_adjoint(vecM, vecB, vecK, vecGamma, vecLambda, vecF, vecQ0, vecPsi,
                     n_fwd: int) -> Assembled:
    """``PASIf.setJacobian`` (``PASIf.py:279-369, 545-605, 620-633``): the
    adjoint system; its tensors mix adjoint indices (first two axes) and
    forward indices (remaining axes).  The terminal condition lambda_T is the
    "initial conditions" vector (``src/__GpuDriver.cu:320-326``)."""
    sizes, diag, B, K, tens, f, q0 = _unfold(vecM, vecB, vecK, [vecGamma, vecLambda, vecPsi],
                                             [3, 4, 5], vecF, vecQ0)
    (gi, gv, _), (li, lv, _), (pi, pv, _) = tens
    return assemble_from_arrays(diag, B, K, [(gi, gv), (li, lv), (pi, pv)], f, q0,
                                scale_force=False, n_set=n_fwd, sizes=sizes)


# ---------------------------------------------------------------------------
# forcing
# ---------------------------------------------------------------------------
# (offset, half) per RK4 stage: stage 1 (0, F), 2 (0, T), 3 (0, T), 4 (1, F)
# -- src/__GpuDriver.cu:697,701,705,709
STAGE_OFFSET = (0, 0, 0, 1)
STAGE_HALF = (0, 1, 1, 0)


def stage_sample_index_zoh(step: int, stage: int, length: int) -> Optional[int]:
    """Zero-order hold: sample ``s = step + offset`` used only when
    ``0 < s < L`` (``src/kernels.cu:83-87``); the half-step flag is ignored
    so stages 2 and 3 reuse ``e[step]``."""
    s = step + STAGE_OFFSET[stage]
    return s if 0 < s < length else None


def stage_sample_index_interp(step: int, stage: int, length: int, window: int):
    """Interpolation on the half-step grid (``src/kernels.cu:114-124``):
    ``hs = 2*(step+offset) + half``, ``i = hs // W``, ``j = hs % W``; used only
    when ``0 < i < L``.  Returns ``(i, j)`` or ``None``."""
    hs = 2 * (step + STAGE_OFFSET[stage]) + STAGE_HALF[stage]
    i, j = divmod(hs, window)
    return (i, j) if 0 < i < length else None


@dataclass
class Forcing:
    """Excitation set (``_loadExcitationsSet``, ``src/__GpuDriver.cu:343-381``)
    and optional interpolation matrix (``_setInterpolationMatrix`` 383-398).
    ``exc`` is ``[n_exc][L]``; ``interp`` is ``[rows][W]`` (rows >= 2)."""
    exc: np.ndarray
    interp: Optional[np.ndarray] = None
    _cast: dict = field(default_factory=dict, repr=False, compare=False)

    @property
    def length(self) -> int:
        return self.exc.shape[1]

    def as_type(self, dt):
        """Samples and interpolation rows rounded to ``dt`` (once per type)."""
        dt = np.dtype(dt)
        if dt not in self._cast:
            self._cast[dt] = (self.exc.astype(dt, copy=False),
                              None if self.interp is None else self.interp.astype(dt, copy=False))
        return self._cast[dt]


def excitation_stage_values(forcing: Forcing, step: int, dtype=np.float64) -> np.ndarray:
    """u for every excitation at the four stages of ``step``: ``[4][n_exc]``.

    ZOH: ``e[s]`` (``src/kernels.cu:86``).  Interpolation:
    ``Mi[1][j]*e[i] + Mi[0][j]*e[i+1]`` (``src/kernels.cu:124``) with
    ``e[L] := 0``."""
    dt = np.dtype(dtype)
    exc, M = forcing.as_type(dt)
    n_exc, L = exc.shape
    u = np.zeros((4, n_exc), dtype=dt)
    if forcing.interp is None:
        for s in range(4):
            k = stage_sample_index_zoh(step, s, L)
            if k is not None:
                u[s] = exc[:, k]
    else:
        W = M.shape[1]
        for s in range(4):
            ij = stage_sample_index_interp(step, s, L, W)
            if ij is None:
                continue
            i, j = ij
            e_next = exc[:, i + 1] if i + 1 < L else np.zeros(n_exc, dtype=dt)
            u[s] = M[1, j] * exc[:, i] + M[0, j] * e_next
    return u


# ---------------------------------------------------------------------------
# instances (replicas / per-system excitation / per-system step count)
# ---------------------------------------------------------------------------
@dataclass
class InstanceLayout:
    """Which excitation drives which DOF of which replica, and how many steps
    each DOF integrates.

    The reference replicates the whole global system once per excitation
    (``parallelizeThroughExcitations``, ``src/__GpuDriver.cu:896-938``;
    replica ``p`` reads excitation ``d_indice/systemStride == p``,
    ``src/kernels.cu:85``): that is ``cartesian_layout``.  ``mapped_layout``
    is an extension beyond the reference: one replica, each system its own
    excitation id and its own step count.
    """
    exc_id: np.ndarray          # [P][N] excitation index per replica/dof
    steps: np.ndarray           # [P][N] step count per replica/dof

    @property
    def P(self) -> int:
        return self.exc_id.shape[0]


def cartesian_layout(n: int, n_exc: int, numsteps: int) -> InstanceLayout:
    exc = np.repeat(np.arange(n_exc, dtype=np.int64)[:, None], n, axis=1)
    return InstanceLayout(exc, np.full((n_exc, n), numsteps, dtype=np.int64))


def mapped_layout(sizes: Sequence[int], exc_of_system, steps_of_system) -> InstanceLayout:
    sizes = np.asarray(sizes, dtype=np.int64)
    exc = np.repeat(np.asarray(exc_of_system, dtype=np.int64), sizes)[None, :]
    st = np.repeat(np.asarray(steps_of_system, dtype=np.int64), sizes)[None, :]
    return InstanceLayout(exc, st)


# ---------------------------------------------------------------------------
# right-hand side
# ---------------------------------------------------------------------------
class _Scatter:
    """Sum of per-entry contributions into output rows in COO order: a CSR
    ``S[o, t] = 1`` applied to the contribution vector sums each row's entries
    in ascending entry order, i.e. sequentially in COO order."""

    def __init__(self, out_idx: np.ndarray, n_out: int, dtype):
        nnz = len(out_idx)
        self.empty = nnz == 0
        self.S = sp.csr_matrix((np.ones(nnz, dtype=dtype), (out_idx, np.arange(nnz))),
                               shape=(n_out, max(nnz, 1)))
        self.S.sort_indices()

    def __call__(self, contrib: np.ndarray) -> np.ndarray:
        # contrib: [nnz][P] -> [n_out][P]
        return self.S @ contrib


class RHS:
    """``derivatives`` (``src/__GpuDriver.cu:785-838``):

    ``Y = A.S + sum G[o,j,k] S_j X_k + sum L[o,j,k,l] S_j X_k X_l
    + sum Psi[o,j,k,l,m] S_j X_k X_l X_m`` (SpMM then ``SpTdV``,
    ``src/kernels.cu:21-63``) and the forcing term ``P_o u`` added by
    ``applyForces`` / ``interpolateForces_fwd`` (``src/kernels.cu:71-126``).
    ``S`` is the integrated state (``X_source``), ``X`` the set-point state
    (``X_setpoint``; equal to ``S`` in the forward pass,
    ``src/__GpuDriver.cu:789-791``).  Arrays are ``[N][P]`` (dof-major) here.
    """

    def __init__(self, sys: Assembled, dtype=np.float64):
        self.dt = np.dtype(dtype)
        s = sys.astype(self.dt)
        self.sys = s
        self.A = s.A
        self.g = _Scatter(s.gamma_idx[:, 0], s.n, self.dt)
        self.l = _Scatter(s.lam_idx[:, 0], s.n, self.dt)
        self.p = _Scatter(s.psi_idx[:, 0], s.n, self.dt)
        self.forced = np.nonzero(s.force)[0]
        self.fval = s.force[self.forced]

    def __call__(self, S: np.ndarray, X: np.ndarray, u_dof: Optional[np.ndarray]) -> np.ndarray:
        s = self.sys
        y = self.A @ S
        if not self.g.empty:
            gi = s.gamma_idx
            c = (s.gamma_val[:, None] * S[gi[:, 1]]) * X[gi[:, 2]]
            y = y + self.g(c)
        if not self.l.empty:
            li = s.lam_idx
            c = ((s.lam_val[:, None] * S[li[:, 1]]) * X[li[:, 2]]) * X[li[:, 3]]
            y = y + self.l(c)
        if not self.p.empty:
            pi = s.psi_idx
            c = (((s.psi_val[:, None] * S[pi[:, 1]]) * X[pi[:, 2]]) * X[pi[:, 3]]) * X[pi[:, 4]]
            y = y + self.p(c)
        if u_dof is not None and len(self.forced):
            y[self.forced] = y[self.forced] + self.fval[:, None] * u_dof[self.forced]
        return y


def rhs(sys: Assembled, S, X=None, u_dof=None, dtype=np.float64):
    """Convenience single evaluation; ``S``/``X``/``u_dof`` are ``[N]`` or
    ``[N][P]``."""
    r = RHS(sys, dtype)
    S = np.asarray(S, dtype=dtype)
    squeeze = S.ndim == 1
    S2 = S[:, None] if squeeze else S
    X2 = S2 if X is None else (np.asarray(X, dtype=dtype)[:, None] if squeeze else np.asarray(X, dtype=dtype))
    U2 = None
    if u_dof is not None:
        U2 = np.asarray(u_dof, dtype=dtype)
        U2 = U2[:, None] if squeeze else U2
    out = r(S2, X2, U2)
    return out[:, 0] if squeeze else out


# ---------------------------------------------------------------------------
# RK4 forward
# ---------------------------------------------------------------------------
def update_slope(q, rk, dt):
    """``rki = q + dt*rk`` (``updateSlope``, ``src/kernels.cu:201-210``)."""
    return q + dt * rk


def integrate(q, rk1, rk2, rk3, rk4, h6):
    """``q + h6*(rk1 + 2*rk2 + 2*rk3 + rk4)`` evaluated left to right
    (``integrate``, ``src/kernels.cu:218-229``)."""
    two = np.asarray(q).dtype.type(2)
    return q + h6 * (((rk1 + two * rk2) + two * rk3) + rk4)


@dataclass
class ForwardResult:
    q: np.ndarray                      # final state [P][N]
    traj: Optional[np.ndarray] = None  # [slots][P][N]
    maxabs: Optional[np.ndarray] = None
    checkpoints: dict = field(default_factory=dict)


def _u_dof(forcing: Optional[Forcing], layout: InstanceLayout, step: int, dt):
    if forcing is None:
        return [None] * 4
    u = excitation_stage_values(forcing, step, dt)    # [4][E]
    return [u[s][layout.exc_id].T for s in range(4)]  # each [N][P]


def rk4_forward(sys: Assembled, forcing: Optional[Forcing], sample_rate: int,
                layout: InstanceLayout, *, dtype=np.float64, save_steps: int = 0,
                window: int = 0, q_init: Optional[np.ndarray] = None,
                step_begin: int = 0, step_end: Optional[int] = None,
                keep_states: Optional[Sequence[int]] = None) -> ForwardResult:
    """``forwardRungeKutta`` + ``fwdStep`` (``src/__GpuDriver.cu:642-713``),
    ``updateSlope`` / ``integrate`` (``src/kernels.cu:201-229``)::

        k1 = f(q, n, st1); y = q + h2 k1
        k2 = f(y, n, st2); y = q + h2 k2
        k3 = f(y, n, st3); y = q + h  k3
        k4 = f(y, n, st4)
        q  = q + h6 (((k1 + 2 k2) + 2 k3) + k4)

    A DOF whose step count is reached keeps its state.  ``save_steps > 0``
    records the post-step state after every step ``t`` with
    ``t % save_steps == 0`` (``src/__GpuDriver.cu:669-687``): slot ``j`` holds
    ``q_{j*s+1}``; ``floor(T/s)`` slots (``PASIf.py:412``).  ``window > 0``
    returns ``max |q_m|`` over the post-step states ``m`` in
    ``(T-window, T]`` of each DOF (a fused reduction the reference does not have).
    ``keep_states`` returns the states ``q_n`` for the listed step numbers
    (0 = initial state)."""
    dt = np.dtype(dtype)
    h, h2, h6 = step_constants(sample_rate, dt)
    f = RHS(sys, dt)
    P = layout.P
    steps = layout.steps.T                       # [N][P]
    Tmax = int(steps.max()) if step_end is None else step_end
    q = (np.repeat(sys.q0.astype(dt)[:, None], P, axis=1) if q_init is None
         else np.array(q_init, dtype=dt).reshape(P, -1).T.copy())
    res = ForwardResult(q=None)
    keep = set(keep_states or [])
    if 0 in keep and step_begin == 0:
        res.checkpoints[0] = q.T.copy()
    nslots = (Tmax // save_steps) if save_steps else 0
    traj = np.zeros((nslots, sys.n, P), dtype=dt) if nslots else None
    mx = np.zeros_like(q) if window else None
    for n in range(step_begin, Tmax):
        u1, u2, u3, u4 = _u_dof(forcing, layout, n, dt)
        k1 = f(q, q, u1)
        y = update_slope(q, k1, h2)
        k2 = f(y, y, u2)
        y = update_slope(q, k2, h2)
        k3 = f(y, y, u3)
        y = update_slope(q, k3, h)
        k4 = f(y, y, u4)
        qn = integrate(q, k1, k2, k3, k4, h6)
        act = n < steps
        q = np.where(act, qn, q)
        if traj is not None and n % save_steps == 0 and n // save_steps < nslots:
            traj[n // save_steps] = q
        if mx is not None:
            inwin = act & (n + 1 > steps - window)
            mx = np.where(inwin, np.maximum(mx, np.abs(q)), mx)
        if (n + 1) in keep:
            res.checkpoints[n + 1] = q.T.copy()
    res.q = q.T.copy()
    if traj is not None:
        res.traj = np.transpose(traj, (0, 2, 1)).copy()
    if mx is not None:
        res.maxabs = mx.T.copy()
    return res


# ---------------------------------------------------------------------------
# adjoint
# ---------------------------------------------------------------------------
def adjoint_gradient(fwd: Assembled, adj: Assembled, forcing: Optional[Forcing],
                     sample_rate: int, layout: InstanceLayout, adj_steps: np.ndarray,
                     *, dtype=np.float64, checkpoint_every: int = 256) -> np.ndarray:
    """Adjoint recursion of ``backwardRungeKutta`` / ``bwdStep``
    (``src/__GpuDriver.cu:715-783``) with the set-point kernels
    ``request_setpoint`` / ``half_setpoint`` / ``previous_setpoint``
    (``src/kernels.cu:162-195``), run over the clean schedule n = T .. 1::

        m1 = g(lam, x_n);             y = lam + h2 m1
        m2 = g(y, xbar);              y = lam + h2 m2
        m3 = g(y, xbar);              y = lam + h  m3
        m4 = g(y, x_{n-1})
        lam = lam + h6 (((m1 + 2 m2) + 2 m3) + m4),   xbar = .5 (x_n + x_{n-1})

    starting from lam_T = the adjoint "initial conditions"; no forcing term
    (all backward ``modterpolator`` calls are commented out,
    ``src/__GpuDriver.cu:756-780``).  Forward states come from a segment-wise
    recompute between checkpoints every ``checkpoint_every`` steps; the
    result does not depend on that interval.  An adjoint DOF of a system with
    step count T_s only updates for n <= T_s.  Returns ``lam_0`` as ``[P][Na]``.

    ``adj_steps`` is ``[P][Na]``: the step count of the system owning each
    adjoint DOF.
    """
    dt = np.dtype(dtype)
    h, h2, h6 = step_constants(sample_rate, dt)
    half = dt.type(0.5)
    g = RHS(adj, dt)
    P = layout.P
    Tmax = int(layout.steps.max())
    k = max(1, int(checkpoint_every))
    seg_starts = list(range(0, Tmax, k))
    ck = rk4_forward(fwd, forcing, sample_rate, layout, dtype=dt,
                     keep_states=seg_starts).checkpoints
    lam = np.repeat(adj.q0.astype(dt)[:, None], P, axis=1)
    asteps = np.asarray(adj_steps).T            # [Na][P]
    for s0 in reversed(seg_starts):
        s1 = min(s0 + k, Tmax)
        rec = rk4_forward(fwd, forcing, sample_rate, layout, dtype=dt, q_init=ck[s0],
                          step_begin=s0, step_end=s1, keep_states=range(s0 + 1, s1 + 1))
        xs = {s0: ck[s0].T}
        xs.update({m: v.T for m, v in rec.checkpoints.items()})
        for n in range(s1, s0, -1):
            xn, xp = xs[n], xs[n - 1]
            xm = half * (xn + xp)
            m1 = g(lam, xn, None)
            y = update_slope(lam, m1, h2)
            m2 = g(y, xm, None)
            y = update_slope(lam, m2, h2)
            m3 = g(y, xm, None)
            y = update_slope(lam, m3, h)
            m4 = g(y, xp, None)
            ln = integrate(lam, m1, m2, m3, m4, h6)
            lam = np.where(n <= asteps, ln, lam)
    return lam.T.copy()


def discrete_adjoint_gradient(fwd: Assembled, adj: Assembled, forcing: Optional[Forcing],
                              sample_rate: int, layout: InstanceLayout, adj_steps: np.ndarray,
                              *, dtype=np.float64, checkpoint_every: int = 256) -> np.ndarray:
    """Exact derivative of the computed forward map (an extension: the
    reference's recursion, ``adjoint_gradient``, integrates the continuous
    adjoint with linearly interpolated set points and so differs from the
    derivative of the discrete forward map by O(h^2)).

    Reverse-mode differentiation of ``rk4_forward``'s step from q = x_{n-1}
    (stage inputs y1 = q, y2 = q + h2 k1, y3 = q + h2 k2, y4 = q + h k3 with
    the step's forcing), for n = T .. 1::

        a4 = h6 lam;              mu4 = g(a4, y4)
        a3 = 2 h6 lam + h mu4;    mu3 = g(a3, y3)
        a2 = 2 h6 lam + h2 mu3;   mu2 = g(a2, y2)
        a1 = h6 lam + h2 mu2;     mu1 = g(a1, y1)
        lam = lam + (((mu4 + mu3) + mu2) + mu1)

    where ``g(v, x)`` is the assembled adjoint right-hand side with set point
    x -- ``J(x)^T v`` when the adjoint system is the transposed Jacobian of
    the forward one (parameter rows included: their rows collect
    ``(df/dtheta)^T v``).  Terminal condition, step counts and checkpoints as
    in ``adjoint_gradient``.  Returns ``lam_0`` as ``[P][Na]``."""
    dt = np.dtype(dtype)
    h, h2, h6 = step_constants(sample_rate, dt)
    two = dt.type(2)
    f = RHS(fwd, dt)
    g = RHS(adj, dt)
    P = layout.P
    Tmax = int(layout.steps.max())
    k = max(1, int(checkpoint_every))
    seg_starts = list(range(0, Tmax, k))
    ck = rk4_forward(fwd, forcing, sample_rate, layout, dtype=dt,
                     keep_states=seg_starts).checkpoints
    lam = np.repeat(adj.q0.astype(dt)[:, None], P, axis=1)
    asteps = np.asarray(adj_steps).T            # [Na][P]
    for s0 in reversed(seg_starts):
        s1 = min(s0 + k, Tmax)
        rec = rk4_forward(fwd, forcing, sample_rate, layout, dtype=dt, q_init=ck[s0],
                          step_begin=s0, step_end=s1, keep_states=range(s0 + 1, s1 + 1))
        xs = {s0: ck[s0].T}
        xs.update({m: v.T for m, v in rec.checkpoints.items()})
        for n in range(s1, s0, -1):
            q = xs[n - 1]
            u1, u2, u3, _ = _u_dof(forcing, layout, n - 1, dt)
            k1 = f(q, q, u1)
            y2 = update_slope(q, k1, h2)
            k2 = f(y2, y2, u2)
            y3 = update_slope(q, k2, h2)
            k3 = f(y3, y3, u3)
            y4 = update_slope(q, k3, h)
            mu4 = g(h6 * lam, y4, None)
            mu3 = g((two * h6) * lam + h * mu4, y3, None)
            mu2 = g((two * h6) * lam + h2 * mu3, y2, None)
            mu1 = g(h6 * lam + h2 * mu2, q, None)
            ln = lam + (((mu4 + mu3) + mu2) + mu1)
            lam = np.where(n <= asteps, ln, lam)
    return lam.T.copy()
