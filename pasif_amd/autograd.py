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
from .PASIf import coo_tensor
from .workloads import Workload, make_solver


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

    def _load(self, q0: np.ndarray, theta: np.ndarray):
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
        self.solver.setSystems(w.vecM, vB, vK, vG, vL, w.vecForce, vQ)
        self.solver.setSystemExcitations(w.exc_of_system, w.steps_of_system)
        if w.interp is not None:
            self.solver.setInterpolationMatrix(list(w.interp))
        return vB, vK, vG, vL

    def forward_solve(self, q0: np.ndarray, theta: np.ndarray) -> np.ndarray:
        import time
        t0 = time.perf_counter()
        self._load(q0, theta)
        self.timing["host_s"] += time.perf_counter() - t0
        return self.solver.getAmplitudes()

    def adjoint_solve(self, q0: np.ndarray, theta: np.ndarray, grad_qT: np.ndarray):
        import time
        t0 = time.perf_counter()
        vB, vK, vG, vL = self._load(q0, theta)
        terms = [grad_qT[self.offsets[s]:self.offsets[s + 1]] for s in range(len(self.sizes))]
        J = ADJ.jacobian_blocks(self.w.vecM, vB, vK, vG, vL, self.params, terms)
        self.solver.setJacobianBlocks(J)
        self.timing["host_s"] += time.perf_counter() - t0
        lam0 = self.solver.getGradient()
        self.timing["solves"] += 1
        aoff = np.concatenate([[0], np.cumsum(J.sizes)])
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
