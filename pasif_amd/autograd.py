"""Differentiable batched RK4: a ``torch.autograd.Function`` whose forward is
the GPU forward solve and whose backward is the GPU checkpointed adjoint
(the reference recursion of ``src/__GpuDriver.cu:502-587, 715-783``, run by
the tile kernel's recompute / reverse phases).

Inputs are the initial states ``q0`` (all systems, flattened ``[N]``) and a
parameter vector ``theta`` -- selected user coefficients (entries of K or B,
Gamma or Lambda values; see :mod:`pasif_amd.adjoint`).  The output is the
final state ``q(T)`` ``[N]``.  ``backward`` receives ``dL/dq(T)``, runs the
adjoint with that terminal condition and returns ``dL/dq0`` and
``dL/dtheta``.  The gradient is that of the reference's adjoint scheme: it
agrees with finite differences of the forward to the scheme's order
(O(dt^2)), not to round-off.

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


def _set_entry(mat, r: int, c: int, value: float):
    m = coo_matrix(mat)
    sel = (m.row == r) & (m.col == c)
    if not np.any(sel):
        raise ValueError(f"no entry at ({r}, {c}) to differentiate")
    data = m.data.astype(np.float64).copy()
    idx = np.flatnonzero(sel)
    data[idx] = 0.0
    data[idx[0]] = value
    return coo_matrix((data, (m.row, m.col)), shape=m.shape)


def _get_entry(mat, r: int, c: int) -> float:
    m = coo_matrix(mat)
    sel = (m.row == r) & (m.col == c)
    return float(m.data[sel].sum())


def _copy_tensor(t):
    c = coo_tensor(list(t.dimensions))
    c.val = np.asarray(t.val, dtype=np.float64).copy()
    c.indices = np.asarray(t.indices).copy()
    return c


class DifferentiableProblem:
    """A batch of systems (one excitation each) with differentiable
    coefficients ``params_of_system[s]``."""

    def __init__(self, w: Workload, params_of_system: Sequence[Sequence[tuple]],
                 precision: int = 64, checkpoint_every: int = 256):
        if w.exc_of_system is None:
            raise ValueError("differentiation needs one excitation per system "
                             "(Workload.exc_of_system)")
        self.w = w
        self.params = [list(p) for p in params_of_system]
        self.sizes = w.sizes()
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])
        self.npar = np.array([len(p) for p in self.params])
        self.poffsets = np.concatenate([[0], np.cumsum(self.npar)])
        self.solver = make_solver(w)
        self.solver.setPrecision(precision)
        self.solver.setTuning(checkpoint_every=checkpoint_every)

    def theta0(self) -> np.ndarray:
        vals = []
        for s, ps in enumerate(self.params):
            for par in ps:
                if par[0] == "K":
                    vals.append(_get_entry(self.w.vecK[s], par[1], par[2]))
                elif par[0] == "B":
                    vals.append(_get_entry(self.w.vecB[s], par[1], par[2]))
                elif par[0] == "Gamma":
                    vals.append(float(np.asarray(self.w.vecGamma[s].val)[par[1]]))
                else:
                    vals.append(float(np.asarray(self.w.vecLambda[s].val)[par[1]]))
        return np.array(vals)

    def q00(self) -> np.ndarray:
        return np.concatenate([np.asarray(q, dtype=np.float64) for q in self.w.vecQ0])

    def _load(self, q0: np.ndarray, theta: np.ndarray):
        w = self.w
        vB, vK = list(w.vecB), list(w.vecK)
        vG, vL = list(w.vecGamma), list(w.vecLambda)
        for s, ps in enumerate(self.params):
            th = theta[self.poffsets[s]:self.poffsets[s + 1]]
            for par, v in zip(ps, th):
                if par[0] == "K":
                    vK[s] = _set_entry(vK[s], par[1], par[2], v)
                elif par[0] == "B":
                    vB[s] = _set_entry(vB[s], par[1], par[2], v)
                elif par[0] == "Gamma":
                    if vG[s] is w.vecGamma[s]:
                        vG[s] = _copy_tensor(vG[s])
                    vG[s].val[par[1]] = v
                else:
                    if vL[s] is w.vecLambda[s]:
                        vL[s] = _copy_tensor(vL[s])
                    vL[s].val[par[1]] = v
        vQ = [q0[self.offsets[s]:self.offsets[s + 1]] for s in range(len(self.sizes))]
        self.solver.setSystems(w.vecM, vB, vK, vG, vL, w.vecForce, vQ)
        self.solver.setSystemExcitations(w.exc_of_system, w.steps_of_system)
        if w.interp is not None:
            self.solver.setInterpolationMatrix(list(w.interp))
        return vB, vK, vG, vL

    def forward_solve(self, q0: np.ndarray, theta: np.ndarray) -> np.ndarray:
        self._load(q0, theta)
        return self.solver.getAmplitudes()

    def adjoint_solve(self, q0: np.ndarray, theta: np.ndarray, grad_qT: np.ndarray):
        vB, vK, vG, vL = self._load(q0, theta)
        terms = [grad_qT[self.offsets[s]:self.offsets[s + 1]] for s in range(len(self.sizes))]
        J = ADJ.jacobian_inputs(self.w.vecM, vB, vK, vG, vL, self.params, terms)
        self.solver.setJacobian(*J.as_args())
        lam0 = self.solver.getGradient()
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
