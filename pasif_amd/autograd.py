"""Differentiable batched RK4: a ``torch.autograd.Function`` whose forward is
the GPU forward solve and whose backward is the GPU checkpointed adjoint:
by default the exact discrete adjoint (the derivative of the computed RK4
map, ``setAdjointMode("exact")``), or with ``mode="reference"`` the
reference recursion of ``src/__GpuDriver.cu:502-587, 715-783``.

Inputs are the initial states ``q0`` (all systems, flattened ``[N]``) and a
parameter vector ``theta`` -- selected user coefficients (entries of K or B,
Gamma or Lambda values; see :mod:`pasif_amd.adjoint`).  The output is the
final state ``q(T)`` ``[N]``.  ``backward`` receives ``dL/dq(T)``, runs the
adjoint with that terminal condition and returns ``dL/dq0`` and
``dL/dtheta``.  In the exact mode the gradient is that of the computed
forward map (it agrees with finite differences to their own precision); the
reference recursion agrees with them to the scheme's order (O(dt^2)).  The
adjoint systems are assembled vectorised (``adjoint.jacobian_blocks``, the
arrays ``setJacobian`` would unfold) for the current coefficients.

PyTorch is only the tensor / autograd glue here; both solves run in
``libpasif_hip.so``.
"""
from __future__ import annotations

from typing import List, Sequence

import numpy as np
from scipy.sparse import coo_matrix

from . import adjoint as ADJ
from .PASIf import _coo_parts, _diagonal, _neg_inverse_diag, assemble_arrays, coo_tensor
from .workloads import Workload, make_solver


class _ValuePlan:
    """The driver's input arrays of a problem for new coefficient values,
    initial states and terminal conditions, by gathers and scatters into
    positions computed once (no per-system objects):

    * the user values concatenated in system order -- B entries, K entries,
      Gamma values, Lambda values -- with every parameter's entries (a matrix
      position's duplicates zeroed, the first set) scattered in;
    * the forward arrays as the assembly forms them: CSR of
      ``D_r B + D_r K`` (D = -1/diag(M); B entries first, then K added at
      their CSR positions), tensor values times D of their output row;
    * the adjoint arrays (the energy-gradient adjoint systems of
      ``adjoint.jacobian_blocks``): every entry is ``D_r`` times one user
      value or ``D_r`` alone; which value and which row is read once from an
      evaluation of ``jacobian_blocks`` on tagged inputs (M = -I so D = 1,
      user values replaced by distinct integers).

    Both are checked bit for bit against the full assembly at construction;
    ``ok`` is False (and the problem keeps the full path) if any differs."""

    def __init__(self, prob: "DifferentiableProblem", terms0):
        w = prob.w
        S = len(w.vecM)
        sizes = prob.sizes
        off = prob.offsets[:-1]
        n = int(prob.offsets[-1])
        self.n = n
        parts = [[_coo_parts(w.vecB[s]), _coo_parts(w.vecK[s])] for s in range(S)]
        cnt_b = np.array([len(p[0][2]) for p in parts], dtype=np.int64)
        cnt_k = np.array([len(p[1][2]) for p in parts], dtype=np.int64)
        gvals = [np.asarray(w.vecGamma[s].val, dtype=np.float64).reshape(-1) for s in range(S)]
        lvals = [np.asarray(w.vecLambda[s].val, dtype=np.float64).reshape(-1) for s in range(S)]
        cnt_g = np.array([v.size for v in gvals], dtype=np.int64)
        cnt_l = np.array([v.size for v in lvals], dtype=np.int64)
        rB = np.concatenate([np.asarray(p[0][0], dtype=np.int64) + o for p, o in zip(parts, off)])
        cB = np.concatenate([np.asarray(p[0][1], dtype=np.int64) + o for p, o in zip(parts, off)])
        rK = np.concatenate([np.asarray(p[1][0], dtype=np.int64) + o for p, o in zip(parts, off)])
        cK = np.concatenate([np.asarray(p[1][1], dtype=np.int64) + o for p, o in zip(parts, off)])
        self.u0 = np.concatenate(
            [np.concatenate([np.asarray(p[0][2], dtype=np.float64) for p in parts]),
             np.concatenate([np.asarray(p[1][2], dtype=np.float64) for p in parts]),
             np.concatenate(gvals) if gvals else np.zeros(0),
             np.concatenate(lvals) if lvals else np.zeros(0)])
        nb, nk, ng = int(cnt_b.sum()), int(cnt_k.sum()), int(cnt_g.sum())
        ob = np.concatenate([[0], np.cumsum(cnt_b)])
        ok_ = np.concatenate([[0], np.cumsum(cnt_k)]) + nb
        og = np.concatenate([[0], np.cumsum(cnt_g)]) + nb + nk
        ol = np.concatenate([[0], np.cumsum(cnt_l)]) + nb + nk + ng
        self.nb, self.nk, self.ng = nb, nk, ng
        # parameter scatter: theta_j -> user position, and positions zeroed
        tpos, zpos = np.zeros(int(prob.poffsets[-1]), dtype=np.int64), []
        for s in range(S):
            base = int(prob.poffsets[s])
            for j, which, sel in prob._mat_par[s]:
                start = ob[s] if which == "B" else ok_[s]
                zpos.append(start + np.asarray(sel, dtype=np.int64))
                tpos[base + j] = start + int(sel[0])
            for kind, start in (("Gamma", og[s]), ("Lambda", ol[s])):
                js, ts = prob._ten_par[s][kind]
                if js.size:
                    tpos[base + js] = start + ts
        self.tpos = tpos
        self.zpos = np.concatenate(zpos) if zpos else np.zeros(0, dtype=np.int64)
        diag = np.concatenate([_diagonal(m) for m in w.vecM])
        self.D = -(1.0 / diag)
        # forward: CSR structure and the place of every B / K entry in it,
        # from the full assembly of the inputs at the initial coefficients
        theta0 = prob.theta0()
        self.ref_inputs = prob._inputs(prob.q00(), theta0)
        vB0, vK0, vG0, vL0, vQ0 = self.ref_inputs
        A = assemble_arrays(w.vecM, vB0, vK0, vG0, vL0, w.vecForce, vQ0)
        csr = A["csr"]
        self.indices = np.asarray(csr.indices)
        self.indptr = np.asarray(csr.indptr)
        keys = np.repeat(np.arange(n, dtype=np.int64), np.diff(self.indptr)) * n + self.indices
        self.posB = np.searchsorted(keys, rB * n + cB)
        self.posK = np.searchsorted(keys, rK * n + cK)
        self.rB, self.rK = rB, rK
        gi = np.asarray(A["gamma"].indices, dtype=np.int64)
        li = np.asarray(A["lam"].indices, dtype=np.int64)
        self.g_idx, self.l_idx = gi, li
        self.g_dims, self.l_dims = list(A["gamma"].dimensions), list(A["lam"].dimensions)
        self.g_row = gi.reshape(-1, 3)[:, 0] if gi.size else np.zeros(0, dtype=np.int64)
        self.l_row = li.reshape(-1, 4)[:, 0] if li.size else np.zeros(0, dtype=np.int64)
        self.force = A["force"]
        f = self.forward(prob.q00(), theta0)
        self.ok = (np.array_equal(f["data"], csr.data)
                   and np.array_equal(f["gval"], np.asarray(A["gamma"].val))
                   and np.array_equal(f["lval"], np.asarray(A["lam"].val)))
        if self.ok:
            self.ok = self._adjoint_plan(prob, parts, gvals, lvals, theta0, terms0)

    def values(self, theta):
        u = self.u0.copy()
        u[self.zpos] = 0.0
        u[self.tpos] = theta
        return u

    def forward(self, q0, theta) -> dict:
        u = self.values(theta)
        nb, nk, ng = self.nb, self.nk, self.ng
        data = np.zeros(self.indices.size)
        data[self.posB] = self.D[self.rB] * u[:nb]
        data[self.posK] = data[self.posK] + self.D[self.rK] * u[nb:nb + nk]
        return {"data": data, "gval": u[nb + nk:nb + nk + ng] * self.D[self.g_row],
                "lval": u[nb + nk + ng:] * self.D[self.l_row], "q0": q0, "u": u}

    def _adjoint_plan(self, prob, parts, gvals, lvals, theta0, terms0) -> bool:
        from scipy.sparse import coo_matrix as coo, dia_matrix as dia
        w = prob.w
        S = len(w.vecM)
        tag = 2.0 + np.arange(self.u0.size, dtype=np.float64)
        nb, nk, ng = self.nb, self.nk, self.ng
        cut = np.cumsum([[len(p[0][2]) for p in parts], [len(p[1][2]) for p in parts]], axis=1)
        vM, vB, vK, vG, vL = [], [], [], [], []
        a_b, a_k = 0, nb
        a_g, a_l = nb + nk, nb + nk + ng
        for s in range(S):
            ns = int(self.n_of(prob, s))
            vM.append(dia((-np.ones(ns), [0]), shape=(ns, ns)))
            (r, c, v), (r2, c2, v2) = parts[s]
            vB.append(coo((tag[a_b:a_b + len(v)], (r, c)), shape=(ns, ns)))
            a_b += len(v)
            vK.append(coo((tag[a_k:a_k + len(v2)], (r2, c2)), shape=(ns, ns)))
            a_k += len(v2)
            g = _copy_tensor(w.vecGamma[s])
            g.val = tag[a_g:a_g + gvals[s].size].copy()
            a_g += gvals[s].size
            vG.append(g)
            lt = _copy_tensor(w.vecLambda[s])
            lt.val = tag[a_l:a_l + lvals[s].size].copy()
            a_l += lvals[s].size
            vL.append(lt)
        del cut
        T = ADJ.jacobian_blocks(vM, vB, vK, vG, vL, prob.params, [None] * S)
        self.a_sizes = np.asarray(T.sizes)
        off_a = np.concatenate([[0], np.cumsum(self.a_sizes)])
        off_f = prob.offsets

        def plan(vals, idx, order):
            vals = np.asarray(vals, dtype=np.float64)
            src = np.where(vals == 1.0, -1, (vals - 2.0).astype(np.int64))
            ix = np.asarray(idx, dtype=np.int64).reshape(-1, order)
            # the row whose D scales the entry: the source's own row (B / K
            # row, tensor output row); for a constant, the forward row on
            # the second axis
            u_row = np.concatenate([self.rB, self.rK, self.g_row, self.l_row])
            sys_a = np.searchsorted(off_a, ix[:, 1], side="right") - 1
            const_row = off_f[sys_a] + (ix[:, 1] - off_a[sys_a])
            drow = np.where(src >= 0, u_row[np.maximum(src, 0)], const_row)
            return src, drow

        Kc = T.K.tocoo() if hasattr(T.K, "tocoo") else T.K
        self.ak_src, self.ak_drow = plan(Kc.data, np.stack([Kc.row, Kc.col], axis=1), 2)
        self.ak_rows, self.ak_cols = np.asarray(Kc.row), np.asarray(Kc.col)
        self.ag_src, self.ag_drow = plan(T.gamma.val, T.gamma.indices, 3)
        self.al_src, self.al_drow = plan(T.lam.val, T.lam.indices, 4)
        self.ap_src, self.ap_drow = plan(T.psi.val, T.psi.indices, 5)
        self.ag_idx, self.al_idx, self.ap_idx = (np.asarray(t.indices) for t in
                                                 (T.gamma, T.lam, T.psi))
        self.ag_dims, self.al_dims, self.ap_dims = (list(t.dimensions) for t in
                                                    (T.gamma, T.lam, T.psi))
        Na = int(off_a[-1])
        self.Na = Na
        # the adjoint K as the driver receives it: CSR, duplicates summed
        st = coo((np.arange(1, Kc.data.size + 1, dtype=np.float64), (Kc.row, Kc.col)),
                 shape=(Na, Na)).tocsr()
        st.sort_indices()
        self.ak_indices, self.ak_indptr = np.asarray(st.indices), np.asarray(st.indptr)
        akeys = np.repeat(np.arange(Na, dtype=np.int64), np.diff(st.indptr)) * Na + st.indices
        self.ak_pos = np.searchsorted(akeys, self.ak_rows.astype(np.int64) * Na + self.ak_cols)
        # positions of the terminal condition in lambda_T
        self.init_pos = np.concatenate([off_a[s] + np.arange(prob.sizes[s]) for s in range(S)])
        # check against the full assembly at theta0
        a = self.adjoint(self.values(theta0), np.concatenate(
            [np.asarray(t, dtype=np.float64) for t in terms0]))
        vB0, vK0, vG0, vL0, _ = self.ref_inputs
        self.ref_inputs = None
        J = ADJ.jacobian_blocks(w.vecM, vB0, vK0, vG0, vL0, prob.params, terms0)
        # as the solver's setJacobianBlocks hands them on: scaled by
        # -(M')^-1 = 1 (M' = -I), B' + K' in CSR with sorted indices
        Dm = _neg_inverse_diag(dia((-np.ones(Na), [0]), shape=(Na, Na))).tocsc()
        ref = (coo(Dm.dot(coo((Na, Na)))) + coo(Dm.dot(J.K))).tocsr()
        ref.sort_indices()
        return (np.array_equal(a["k_data"], ref.data)
                and np.array_equal(self.ak_indices, ref.indices)
                and np.array_equal(a["g_val"], J.gamma.val)
                and np.array_equal(a["l_val"], J.lam.val)
                and np.array_equal(a["p_val"], J.psi.val)
                and np.array_equal(a["init"], J.init))

    @staticmethod
    def n_of(prob, s):
        return prob.sizes[s]

    def adjoint(self, u, grad_qT) -> dict:
        def vals(src, drow):
            x = np.where(src >= 0, u[np.maximum(src, 0)], 1.0)
            return self.D[drow] * x
        kv = vals(self.ak_src, self.ak_drow)
        k_data = np.zeros(self.ak_indices.size)
        np.add.at(k_data, self.ak_pos, kv)
        init = np.zeros(self.Na)
        init[self.init_pos] = grad_qT
        return {"k_data": k_data, "g_val": vals(self.ag_src, self.ag_drow),
                "l_val": vals(self.al_src, self.al_drow), "p_val": vals(self.ap_src, self.ap_drow),
                "init": init}


def _copy_tensor(t):
    c = coo_tensor(list(t.dimensions))
    c.val = np.asarray(t.val, dtype=np.float64).copy()
    c.indices = np.asarray(t.indices).copy()
    return c


class DifferentiableProblem:
    """A batch of systems (one excitation each) with differentiable
    coefficients ``params_of_system[s]``: ``("K", r, c)`` / ``("B", r, c)``
    (the user matrix entry; duplicates of that position are summed into the
    first), ``("Gamma", t)`` / ``("Lambda", t)`` (the t-th tensor value)."""

    def __init__(self, w: Workload, params_of_system: Sequence[Sequence[tuple]],
                 precision: int = 64, checkpoint_every: int = 256, mode: str = "exact"):
        if w.exc_of_system is None:
            raise ValueError("differentiation needs one excitation per system "
                             "(Workload.exc_of_system)")
        self.w = w
        self.params = [list(p) for p in params_of_system]
        self.sizes = w.sizes()
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])
        self.npar = np.array([len(p) for p in self.params])
        self.poffsets = np.concatenate([[0], np.cumsum(self.npar)])
        # per system: the COO forms of K and B and, for every parameter, the
        # positions it reads / writes (computed once)
        self._coo = []
        self._mat_par = []      # [(param j, "K"/"B", data positions)]
        self._ten_par = []      # ("Gamma"/"Lambda") -> (param js, value indices)
        for s, ps in enumerate(self.params):
            Kc, Bc = coo_matrix(w.vecK[s]), coo_matrix(w.vecB[s])
            self._coo.append((Kc, Bc))
            mp, tg, tl = [], ([], []), ([], [])
            for j, par in enumerate(ps):
                if par[0] in ("K", "B"):
                    m = Kc if par[0] == "K" else Bc
                    sel = np.flatnonzero((m.row == par[1]) & (m.col == par[2]))
                    if not sel.size:
                        raise ValueError(f"system {s}: no {par[0]} entry at ({par[1]}, "
                                         f"{par[2]}) to differentiate")
                    mp.append((j, par[0], sel))
                elif par[0] == "Gamma":
                    tg[0].append(j)
                    tg[1].append(int(par[1]))
                elif par[0] == "Lambda":
                    tl[0].append(j)
                    tl[1].append(int(par[1]))
                else:
                    raise ValueError(f"unknown parameter kind {par[0]!r}")
            self._mat_par.append(mp)
            self._ten_par.append({"Gamma": tuple(np.asarray(x, dtype=np.int64) for x in tg),
                                  "Lambda": tuple(np.asarray(x, dtype=np.int64) for x in tl)})
        self.solver = make_solver(w)
        self.solver.setPrecision(precision)
        self.solver.setTuning(checkpoint_every=checkpoint_every)
        self.mode = mode
        self.solver.setAdjointMode(mode)
        self.timing = {"host_s": 0.0, "solves": 0}
        # value plans (checked against the full assembly); the first solves
        # load through the full path, which sets up the solver's layout
        self.plan = _ValuePlan(self, [np.ones(int(n)) for n in self.sizes])
        self._fwd_key = None        # (q0, theta) of the forward arrays loaded
        self._adj_loaded = False

    def theta0(self) -> np.ndarray:
        out = np.zeros(int(self.poffsets[-1]))
        for s in range(len(self.params)):
            th = out[self.poffsets[s]:self.poffsets[s + 1]]
            Kc, Bc = self._coo[s]
            for j, which, sel in self._mat_par[s]:
                th[j] = float((Kc if which == "K" else Bc).data[sel].sum())
            for kind, src in (("Gamma", self.w.vecGamma[s]), ("Lambda", self.w.vecLambda[s])):
                js, ts = self._ten_par[s][kind]
                if js.size:
                    th[js] = np.asarray(src.val, dtype=np.float64).reshape(-1)[ts]
        return out

    def q00(self) -> np.ndarray:
        return np.concatenate([np.asarray(q, dtype=np.float64) for q in self.w.vecQ0])

    def _inputs(self, q0: np.ndarray, theta: np.ndarray):
        """Per-system user inputs for coefficients ``theta`` and states ``q0``."""
        w = self.w
        vB, vK = list(w.vecB), list(w.vecK)
        vG, vL = list(w.vecGamma), list(w.vecLambda)
        for s in range(len(self.params)):
            th = theta[self.poffsets[s]:self.poffsets[s + 1]]
            mp = self._mat_par[s]
            if mp:
                Kc, Bc = self._coo[s]
                data = {"K": None, "B": None}
                for j, which, sel in mp:
                    m = Kc if which == "K" else Bc
                    if data[which] is None:
                        data[which] = m.data.astype(np.float64).copy()
                    d = data[which]
                    d[sel] = 0.0
                    d[sel[0]] = th[j]
                if data["K"] is not None:
                    vK[s] = coo_matrix((data["K"], (Kc.row, Kc.col)), shape=Kc.shape)
                if data["B"] is not None:
                    vB[s] = coo_matrix((data["B"], (Bc.row, Bc.col)), shape=Bc.shape)
            for kind, lst in (("Gamma", vG), ("Lambda", vL)):
                js, ts = self._ten_par[s][kind]
                if js.size:
                    t = _copy_tensor(lst[s])
                    t.val[ts] = th[js]
                    lst[s] = t
        vQ = [q0[self.offsets[s]:self.offsets[s + 1]] for s in range(len(self.sizes))]
        return vB, vK, vG, vL, vQ

    def _load(self, q0: np.ndarray, theta: np.ndarray):
        """Full path: the solver's systems rebuilt from per-system inputs."""
        w = self.w
        vB, vK, vG, vL, vQ = self._inputs(q0, theta)
        self.solver.setSystems(w.vecM, vB, vK, vG, vL, w.vecForce, vQ)
        self.solver.setSystemExcitations(w.exc_of_system, w.steps_of_system)
        if w.interp is not None:
            self.solver.setInterpolationMatrix(list(w.interp))
        self._fwd_key = None
        self._adj_loaded = False
        return vB, vK, vG, vL

    def _load_forward(self, q0: np.ndarray, theta: np.ndarray):
        """The forward arrays for (q0, theta): through the value plan once the
        solver is set up (same structure, new values), else the full path."""
        key = (q0.tobytes(), theta.tobytes())
        if self._fwd_key == key:
            return
        pl = self.plan
        if pl.ok and self._fwd_key is not None:
            f = pl.forward(q0, theta)
            s = self.solver
            s._setFwdK((pl.n, pl.n), f["data"], pl.indices, pl.indptr)
            s._setFwdGamma(pl.g_dims, f["gval"], pl.g_idx)
            s._setFwdLambda(pl.l_dims, f["lval"], pl.l_idx)
            s._setFwdPsi([1] * 5, [], [])
            s._setFwdForcePattern(pl.force)
            s._setFwdInitialConditions(np.ascontiguousarray(q0, dtype=np.float64))
            s._push_layout()
            s._allocateSystemOnDevice()
        else:
            self._load(q0, theta)
        self._fwd_key = key
        self._adj_loaded = False

    def forward_solve(self, q0: np.ndarray, theta: np.ndarray) -> np.ndarray:
        import time
        t0 = time.perf_counter()
        self._load_forward(q0, theta)
        self.timing["host_s"] += time.perf_counter() - t0
        return self.solver.getAmplitudes()

    def adjoint_solve(self, q0: np.ndarray, theta: np.ndarray, grad_qT: np.ndarray):
        import time
        t0 = time.perf_counter()
        self._load_forward(q0, theta)
        pl = self.plan
        if pl.ok and self.solver.jacobianSet:
            a = pl.adjoint(pl.values(theta), np.ascontiguousarray(grad_qT, dtype=np.float64))
            s = self.solver
            s._setBwdK((pl.Na, pl.Na), a["k_data"], pl.ak_indices, pl.ak_indptr)
            s._setBwdGamma(pl.ag_dims, a["g_val"], pl.ag_idx)
            s._setBwdLambda(pl.al_dims, a["l_val"], pl.al_idx)
            s._setBwdForcePattern(np.zeros(pl.Na))
            s._setBwdInitialConditions(a["init"])
            s._setBwdPsi(pl.ap_dims, a["p_val"], pl.ap_idx)
            s._push_layout()
            s._allocateAdjointOnDevice()
            sizes_a = pl.a_sizes
        else:
            vB, vK, vG, vL, _ = self._inputs(q0, theta)
            terms = [grad_qT[self.offsets[s]:self.offsets[s + 1]] for s in range(len(self.sizes))]
            J = ADJ.jacobian_blocks(self.w.vecM, vB, vK, vG, vL, self.params, terms)
            self.solver.setJacobianBlocks(J)
            sizes_a = J.sizes
        self._adj_loaded = True
        self.timing["host_s"] += time.perf_counter() - t0
        lam0 = self.solver.getGradient()
        self.timing["solves"] += 1
        aoff = np.concatenate([[0], np.cumsum(sizes_a)])
        gq, gt = [], []
        for s in range(len(self.sizes)):
            blk = lam0[aoff[s]:aoff[s + 1]]
            gq.append(blk[:self.sizes[s]])
            gt.append(blk[self.sizes[s]:])
        return np.concatenate(gq), (np.concatenate(gt) if gt else np.zeros(0))


def _torch():
    import torch
    return torch


def make_function():
    torch = _torch()

    class RK4Solve(torch.autograd.Function):
        """q(T) = RK4(q0, theta); backward by the GPU adjoint."""

        @staticmethod
        def forward(ctx, q0, theta, problem: DifferentiableProblem):
            qT = problem.forward_solve(q0.detach().double().cpu().numpy(),
                                       theta.detach().double().cpu().numpy())
            ctx.problem = problem
            ctx.save_for_backward(q0, theta)
            return torch.from_numpy(qT).to(device=q0.device, dtype=q0.dtype)

        @staticmethod
        def backward(ctx, grad_qT):
            q0, theta = ctx.saved_tensors
            gq, gt = ctx.problem.adjoint_solve(q0.detach().double().cpu().numpy(),
                                               theta.detach().double().cpu().numpy(),
                                               grad_qT.detach().double().cpu().numpy())
            return (torch.from_numpy(gq).to(device=q0.device, dtype=q0.dtype),
                    torch.from_numpy(gt).to(device=theta.device, dtype=theta.dtype), None)

    return RK4Solve


_FN = None


def rk4_solve(problem: DifferentiableProblem, q0, theta):
    """Differentiable final state ``q(T)`` of every system."""
    global _FN
    if _FN is None:
        _FN = make_function()
    return _FN.apply(q0, theta, problem)
