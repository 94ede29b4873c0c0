"""The differentiable form's value plan (no GPU): for new coefficients,
initial states and terminal conditions, the driver arrays it gathers equal
the full assembly's (``assemble_arrays`` of the per-system inputs, and
``jacobian_blocks`` as ``setJacobianBlocks`` hands them on) bit for bit."""
import numpy as np
from scipy.sparse import coo_matrix, dia_matrix

from pasif_amd import adjoint as ADJ
from pasif_amd import workloads as W
from pasif_amd.PASIf import _neg_inverse_diag, assemble_arrays


def _problem(w, params):
    import os
    from pasif_amd.autograd import DifferentiableProblem
    os.environ["LOCAL_RANK"] = "-1"
    try:
        return DifferentiableProblem(w, params, checkpoint_every=64)
    finally:
        del os.environ["LOCAL_RANK"]


def test_plan_equals_full_assembly_for_new_values():
    w = W.heterogeneous(40, 100, 400, steps_per_sample=10, n_signals=4,
                        dof_choices=(1, 2, 4, 8), dof_probs=(1, 1, 1, 1), variants_per_size=4)
    params = []
    rng = np.random.default_rng(1)
    for s in range(w.n_systems):
        n = w.vecM[s].shape[0]
        d = (n - 1) // 2
        ps = [("K", d + i, i) for i in range(d)]
        nl = len(np.asarray(w.vecLambda[s].val).reshape(-1))
        ps += [("Lambda", t) for t in range(nl)]
        ng = len(np.asarray(w.vecGamma[s].val).reshape(-1))
        if ng:
            ps.append(("Gamma", int(rng.integers(ng))))
        params.append(ps)
    prob = _problem(w, params)
    pl = prob.plan
    assert pl.ok
    th = prob.theta0() * (1.0 + 0.1 * rng.standard_normal(prob.theta0().size))
    q0 = prob.q00() + rng.standard_normal(prob.q00().size)
    g = rng.standard_normal(q0.size)
    f = pl.forward(q0, th)
    vB, vK, vG, vL, vQ = prob._inputs(q0, th)
    A = assemble_arrays(w.vecM, vB, vK, vG, vL, w.vecForce, vQ)
    np.testing.assert_array_equal(f["data"], A["csr"].data)
    np.testing.assert_array_equal(pl.indices, A["csr"].indices)
    np.testing.assert_array_equal(f["gval"], A["gamma"].val)
    np.testing.assert_array_equal(f["lval"], A["lam"].val)
    a = pl.adjoint(pl.values(th), g)
    terms = [g[prob.offsets[s]:prob.offsets[s + 1]] for s in range(w.n_systems)]
    J = ADJ.jacobian_blocks(w.vecM, vB, vK, vG, vL, params, terms)
    Na = int(np.sum(J.sizes))
    Dm = _neg_inverse_diag(dia_matrix((-np.ones(Na), [0]), shape=(Na, Na))).tocsc()
    ref = (coo_matrix(Dm.dot(coo_matrix((Na, Na)))) + coo_matrix(Dm.dot(J.K))).tocsr()
    ref.sort_indices()
    np.testing.assert_array_equal(a["k_data"], ref.data)
    np.testing.assert_array_equal(pl.ak_indices, ref.indices)
    np.testing.assert_array_equal(a["g_val"], J.gamma.val)
    np.testing.assert_array_equal(a["l_val"], J.lam.val)
    np.testing.assert_array_equal(a["p_val"], J.psi.val)
    np.testing.assert_array_equal(a["init"], J.init)
