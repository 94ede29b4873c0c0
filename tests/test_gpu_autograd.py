"""GPU checks of the differentiable form (torch.autograd.Function over the
forward kernel and the hand-written adjoint kernel):

* gradients of ``L = sum(w * q(T))`` w.r.t. q0 and theta (K diagonal and
  Lambda coefficients of chain systems) equal the oracle adjoint of the same
  assembled adjoint system within 1e-10 relative, in both modes (the
  reference recursion: ``adjoint_gradient``; the exact discrete adjoint, the
  default: ``discrete_adjoint_gradient``);
* they agree with central finite differences of the forward (h = 1e-6):
  the exact mode to 1e-6 relative (finite-difference precision), the
  reference recursion within 2e-4 (its scheme is second order in dt).
"""
import numpy as np
import pytest

import oracle as O
from pasif_amd import adjoint as ADJ
from pasif_amd import workloads as W
from tests.test_gpu_forward import TOL64, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,fd_tol", [("reference", 2e-4), ("exact", 1e-6)])
def test_autograd_matches_oracle_and_finite_differences(mode, fd_tol):
    print(f"[gpu case] autograd chains 4 x d4 x 600, {mode} adjoint", flush=True)
    import torch
    from pasif_amd.autograd import DifferentiableProblem, rk4_solve
    w = W.chains(n_systems=4, d=4, steps=600, steps_per_sample=10)
    d = 4
    params = [[("K", d + i, i) for i in range(d)] + [("Lambda", t) for t in range(d)]
              for _ in range(w.n_systems)]
    prob = DifferentiableProblem(w, params, checkpoint_every=64, mode=mode)
    rng = np.random.default_rng(3)
    q0 = torch.tensor(prob.q00() + rng.normal(0, 0.05, prob.q00().size), dtype=torch.float64,
                      requires_grad=True)
    theta = torch.tensor(prob.theta0(), dtype=torch.float64, requires_grad=True)
    wts = torch.tensor(rng.normal(0, 1, q0.numel()), dtype=torch.float64)
    qT = rk4_solve(prob, q0, theta)
    loss = (wts * qT).sum()
    loss.backward()
    gq, gt = q0.grad.numpy(), theta.grad.numpy()

    # oracle adjoint of the same problem
    vB, vK, vG, vL = prob._load(q0.detach().numpy(), theta.detach().numpy())
    terms = [wts.numpy()[prob.offsets[s]:prob.offsets[s + 1]] for s in range(w.n_systems)]
    J = ADJ.jacobian_inputs(w.vecM, vB, vK, vG, vL, params, terms)
    vQ = [q0.detach().numpy()[prob.offsets[s]:prob.offsets[s + 1]] for s in range(w.n_systems)]
    fwd = O.assemble_forward(w.vecM, vB, vK, vG, vL, w.vecForce, vQ)
    adj = O.assemble_adjoint(*J.as_args(), n_fwd=fwd.n)
    lay = O.mapped_layout(fwd.sizes, w.exc_of_system, np.full(w.n_systems, w.numsteps))
    grad_fn = O.adjoint_gradient if mode == "reference" else O.discrete_adjoint_gradient
    lam0 = grad_fn(fwd, adj, O.Forcing(w.excitations, w.interp), w.sample_rate, lay,
                   np.full((1, adj.n), w.numsteps), checkpoint_every=64)[0]
    aoff = np.concatenate([[0], np.cumsum(J.sizes)])
    oq = np.concatenate([lam0[aoff[s]:aoff[s] + prob.sizes[s]] for s in range(w.n_systems)])
    ot = np.concatenate([lam0[aoff[s] + prob.sizes[s]:aoff[s + 1]] for s in range(w.n_systems)])
    assert rel_err(gq, oq) < TOL64
    assert rel_err(gt, ot) < TOL64

    # central finite differences of the GPU forward
    h = 1e-6

    def L(q, th):
        return float((wts.numpy() * prob.forward_solve(q, th)).sum())

    th0 = theta.detach().numpy()
    qq = q0.detach().numpy()
    for j in (0, 3, 5, 12, 30):
        e = np.zeros_like(th0)
        e[j] = h
        fd = (L(qq, th0 + e) - L(qq, th0 - e)) / (2 * h)
        assert abs(gt[j] - fd) <= fd_tol * abs(fd) + 1e-9, (j, gt[j], fd)
    for i in (0, 4, 9, 20):
        e = np.zeros_like(qq)
        e[i] = h
        fd = (L(qq + e, th0) - L(qq - e, th0)) / (2 * h)
        assert abs(gq[i] - fd) <= fd_tol * abs(fd) + 1e-9, (i, gq[i], fd)


def test_value_plan_reloads_equal_full_reloads():
    # after the first solves, new coefficients and terminal conditions go
    # through the value plan (no per-system objects; the driver keeps its
    # lane plan); three steps of a descent give the same results bit for bit
    # as the full path
    print("[gpu case] autograd value-plan reloads vs full reloads, chains 16 x d8 x 500",
          flush=True)
    import torch
    from pasif_amd.autograd import DifferentiableProblem, rk4_solve
    w = W.chains(n_systems=16, d=8, steps=500, steps_per_sample=10)
    d = 8
    params = [[("K", d + i, i) for i in range(d)] + [("Lambda", t) for t in range(d)]
              for _ in range(w.n_systems)]
    runs = []
    for fast in (True, False):
        prob = DifferentiableProblem(w, params, checkpoint_every=64)
        assert prob.plan.ok
        prob.plan.ok = fast
        th = torch.tensor(prob.theta0(), dtype=torch.float64, requires_grad=True)
        q0 = torch.tensor(prob.q00(), dtype=torch.float64, requires_grad=True)
        out = []
        for it in range(3):
            qT = rk4_solve(prob, q0, th)
            (qT * qT).sum().backward()
            out.append((qT.detach().numpy().copy(), q0.grad.numpy().copy(),
                        th.grad.numpy().copy()))
            with torch.no_grad():
                th -= 1e-3 * th.grad
                q0 -= 1e-3 * q0.grad
            th.grad = None
            q0.grad = None
        runs.append(out)
    for a, b in zip(*runs):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
