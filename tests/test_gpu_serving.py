"""GPU checks of the continuous-batching pool:

* a request's results are bitwise identical alone and in a mix of other
  requests sharing every launch, and match the oracle (forward 1e-10,
  adjoint 1e-10);
* admission respects the HBM budget: requests wait until finished ones free
  memory;
* the server drains a small Poisson trace with latencies recorded.
"""
import numpy as np
import pytest

import oracle as O
from pasif_amd import serving as S
from tests.test_gpu_forward import TOL64, rel_err

pytestmark = pytest.mark.gpu


def _run_alone(pool, req):
    job = pool.submit(req)
    assert job is not None
    while not pool.finished(job):
        pool.tick(wait=True)
    return pool.collect(job, req)


def _oracle(req, bank, lib_rate, interp, dtype=np.float64):
    sy = req.systems
    sizes = np.array([s.n for s in sy])
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    n = int(sizes.sum())
    import scipy.sparse as sp
    data, ind, ptr = S._cat_csr([s.csr for s in sy], offs)
    A = sp.csr_matrix((data, ind, ptr), shape=(n, n))
    gi, gv = S._cat_tens([s.gamma for s in sy], offs, offs, 3)
    li, lv = S._cat_tens([s.lam for s in sy], offs, offs, 4)
    fwd = O.Assembled(n=n, A=A, gamma_idx=gi, gamma_val=gv, lam_idx=li, lam_val=lv,
                      psi_idx=np.zeros((0, 5), dtype=np.int64), psi_val=np.zeros(0),
                      force=np.concatenate([s.force for s in sy]),
                      q0=np.concatenate([s.q0 for s in sy]), sizes=sizes)
    lay = O.mapped_layout(sizes, req.exc_ids, req.steps)
    F = O.Forcing(bank, interp)
    q = O.rk4_forward(fwd, F, lib_rate, lay, dtype=dtype).q[0]
    if not req.adjoint:
        return q, None
    sa = np.array([s.na for s in sy])
    aoffs = np.concatenate([[0], np.cumsum(sa)[:-1]])
    na = int(sa.sum())
    data, ind, ptr = S._cat_csr([s.acsr for s in sy], aoffs)
    Aa = sp.csr_matrix((data, ind, ptr), shape=(na, na))
    gi, gv = S._cat_tens([s.agamma for s in sy], aoffs, offs, 2)
    li, lv = S._cat_tens([s.alam for s in sy], aoffs, offs, 2)
    pi, pv = S._cat_tens([s.apsi for s in sy], aoffs, offs, 2)
    adj = O.Assembled(n=na, A=Aa, gamma_idx=gi, gamma_val=gv, lam_idx=li, lam_val=lv,
                      psi_idx=pi, psi_val=pv, force=np.zeros(na),
                      q0=np.concatenate([s.lamT for s in sy]), n_set=n, sizes=sa)
    ast = np.repeat(req.steps, sa)[None, :]
    g = O.adjoint_gradient(fwd, adj, F, lib_rate, lay, ast, checkpoint_every=256,
                           dtype=dtype)[0]
    return q, g


def test_alone_equals_in_mix_and_matches_oracle():
    print("[gpu case] serving identity", flush=True)
    rng = np.random.default_rng(41)
    lib = S.SystemLibrary(dofs=(1, 2, 4, 8), variants=3)
    bank = S.signal_bank(rng, 6, 3000, steps_per_sample=10)
    interp = __import__("pasif_amd.workloads", fromlist=["x"]).linear_interp_matrix(10)

    def new_pool():
        pool = S.Pool(lanes=256, slot_config=1, steps_per_tick=150, hbm_budget=4e9)
        pool.set_bank(bank, 1000, interp)
        return pool

    mk = lambda rid, adj: S.draw_request(lib, rng, rid, 6, tmin=200, tmax=1200,
                                         max_systems=12, adjoint_frac=1.0 if adj else 0.0)
    target = mk(0, True)
    others = [mk(i, i % 2 == 0) for i in range(1, 7)]
    q1, g1 = _run_alone(new_pool(), target)
    pool = new_pool()
    jobs = [(pool.submit(r), r) for r in [others[0], others[1], target] + others[2:]]
    results = {}
    while jobs:
        pool.tick(wait=True)
        keep = []
        for j, r in jobs:
            if pool.finished(j):
                results[r.rid] = pool.collect(j, r)
            else:
                keep.append((j, r))
        jobs = keep
    q2, g2 = results[0]
    np.testing.assert_array_equal(q1, q2)
    np.testing.assert_array_equal(g1, g2)
    oq, og = _oracle(target, bank, 1000, interp)
    assert rel_err(q1, oq) < TOL64
    assert rel_err(g1, og) < TOL64


def test_admission_waits_for_budget():
    print("[gpu case] serving budget", flush=True)
    rng = np.random.default_rng(42)
    lib = S.SystemLibrary(dofs=(2, 4), variants=2, adjoint=False)
    bank = S.signal_bank(rng, 4, 2000)
    pool = S.Pool(lanes=256, slot_config=1, steps_per_tick=500, hbm_budget=1.0)
    pool.set_bank(bank, 1000, None)
    r = S.draw_request(lib, rng, 0, 4, tmin=100, tmax=300, max_systems=4, adjoint_frac=0.0)
    assert pool.submit(r) is None          # does not fit in 1 byte
    pool2 = S.Pool(lanes=256, slot_config=1, steps_per_tick=500, hbm_budget=2e9)
    pool2.set_bank(bank, 1000, None)
    reqs = [S.draw_request(lib, rng, i, 4, tmin=100, tmax=300, max_systems=4,
                           adjoint_frac=0.0) for i in range(5)]
    for i, r in enumerate(reqs):
        r.arrival = 0.001 * i
    rep = S.SolveServer(pool2).run(reqs, max_wall_s=120)
    assert rep.completed == 5


def test_server_drains_poisson_trace():
    print("[gpu case] serving poisson", flush=True)
    rng = np.random.default_rng(43)
    lib = S.SystemLibrary(dofs=(1, 2, 4, 8, 16), variants=2)
    bank = S.signal_bank(rng, 8, 2000, steps_per_sample=10)
    from pasif_amd.workloads import linear_interp_matrix
    pool = S.Pool(lanes=512, slot_config=1, steps_per_tick=400, hbm_budget=20e9)
    pool.set_bank(bank, 1000, linear_interp_matrix(10))
    arr = S.poisson_arrivals(rng, 40.0, 1.0)
    reqs = []
    for i, t in enumerate(arr):
        r = S.draw_request(lib, rng, i, 8, tmin=100, tmax=1500, max_systems=16)
        r.arrival = float(t)
        reqs.append(r)
    rep = S.SolveServer(pool).run(reqs, max_wall_s=120)
    assert rep.completed == len(reqs)
    assert rep.p99_ms >= rep.p50_ms > 0


def test_registered_ids_match_oracle_under_slot_cap():
    """Requests submitted as registered system ids (lane kernel for the sizes
    that fit it, tiles for the rest) match the same requests submitted with
    their full arrays (tile kernel) within 1e-10; a slot cap refuses admission
    until jobs are collected, and device blocks freed by collected jobs are
    reused by later ones."""
    print("[gpu case] serving registry + slots", flush=True)
    rng = np.random.default_rng(44)
    lib = S.SystemLibrary(dofs=(1, 2, 4, 8, 16), variants=2)
    bank = S.signal_bank(rng, 6, 3000, steps_per_sample=10)
    from pasif_amd.workloads import linear_interp_matrix
    interp = linear_interp_matrix(10)

    def new_pool():
        pool = S.Pool(lanes=256, slot_config=1, steps_per_tick=300, hbm_budget=4e9)
        pool.set_bank(bank, 1000, interp)
        return pool

    reqs = [S.draw_request(lib, rng, i, 6, tmin=200, tmax=1500, max_systems=24)
            for i in range(8)]
    ref = {}
    for r in reqs:
        ref[r.rid] = _run_alone(new_pool(), r)
    pool = new_pool()
    pool.register_library(lib)
    pool.set_slots(1)                      # one tile: every extra request waits
    first = pool.submit_ids(reqs[0])
    assert first is not None               # an empty pool always admits
    assert pool.submit_ids(reqs[1]) is None
    jobs = [(first, reqs[0])]
    pending = list(reqs[1:])
    got = {}
    while jobs or pending:
        while pending:
            j = pool.submit_ids(pending[0])
            if j is None:
                break
            jobs.append((j, pending.pop(0)))
        pool.tick(wait=True)
        keep = []
        for j, r in jobs:
            if pool.finished(j):
                got[r.rid] = pool.collect(j, r)
            else:
                keep.append((j, r))
        jobs = keep
    for r in reqs:
        assert rel_err(got[r.rid][0], ref[r.rid][0]) < TOL64
        if r.adjoint:
            assert rel_err(got[r.rid][1], ref[r.rid][1]) < TOL64
    st = pool.stats()
    assert st["jobs"] == 0 and st["used_bytes"] == 0


def test_lane_served_request_alone_equals_in_mix_and_oracle():
    """A request of registered systems runs on the pool's lane modules (one
    lane per system for small sizes, a workgroup per system for 16-DOF) and
    on tiles (64-DOF) together with other requests: its results are bitwise
    the same alone and in the mix, and match the oracle."""
    print("[gpu case] serving lane modules identity", flush=True)
    rng = np.random.default_rng(45)
    lib = S.SystemLibrary(dofs=(1, 2, 4, 16, 64), variants=2)
    bank = S.signal_bank(rng, 6, 3000, steps_per_sample=10)
    from pasif_amd.workloads import linear_interp_matrix
    interp = linear_interp_matrix(10)

    def new_pool():
        pool = S.Pool(lanes=512, slot_config=1, steps_per_tick=170, hbm_budget=4e9)
        pool.set_bank(bank, 1000, interp)
        pool.register_library(lib)
        return pool

    def run(pool, reqs):
        jobs = [(pool.submit_ids(r), r) for r in reqs]
        assert all(j is not None for j, _ in jobs)
        out = {}
        while jobs:
            pool.tick(wait=True)
            keep = []
            for j, r in jobs:
                if pool.finished(j):
                    out[r.rid] = pool.collect(j, r)
                else:
                    keep.append((j, r))
            jobs = keep
        return out

    mk = lambda rid, adj: S.draw_request(lib, rng, rid, 6, tmin=200, tmax=1200,
                                         max_systems=14, adjoint_frac=1.0 if adj else 0.0)
    target = mk(0, True)
    others = [mk(i, i % 2 == 0) for i in range(1, 6)]
    alone = run(new_pool(), [target])[0]
    mixed = run(new_pool(), others[:2] + [target] + others[2:])[0]
    np.testing.assert_array_equal(alone[0], mixed[0])
    np.testing.assert_array_equal(alone[1], mixed[1])
    oq, og = _oracle(target, bank, 1000, interp)
    assert rel_err(alone[0], oq) < TOL64
    assert rel_err(alone[1], og) < TOL64


def test_server_results_equal_alone_runs():
    """The server collects a finished job while the next tick runs (ticks
    leave finished jobs out): every served result is bitwise the result of
    its request run alone on a fresh pool."""
    print("[gpu case] serving collect during tick", flush=True)
    rng = np.random.default_rng(46)
    lib = S.SystemLibrary(dofs=(1, 2, 4, 16, 64), variants=2)
    bank = S.signal_bank(rng, 6, 3000, steps_per_sample=10)
    from pasif_amd.workloads import linear_interp_matrix
    interp = linear_interp_matrix(10)

    def new_pool():
        pool = S.Pool(lanes=512, slot_config=1, steps_per_tick=150, hbm_budget=4e9)
        pool.set_bank(bank, 1000, interp)
        pool.register_library(lib)
        return pool

    reqs = [S.draw_request(lib, rng, i, 6, tmin=100, tmax=1500, max_systems=10)
            for i in range(10)]
    for i, r in enumerate(reqs):
        r.arrival = 0.002 * i
    rep = S.SolveServer(new_pool()).run(reqs, max_wall_s=120)
    assert rep.completed == len(reqs)
    for r in reqs:
        pool = new_pool()
        j = pool.submit_ids(r)
        while not pool.finished(j):
            pool.tick(wait=True)
        q, g = pool.collect(j, r)
        np.testing.assert_array_equal(r.result, q)
        if r.adjoint:
            np.testing.assert_array_equal(r.grad, g)


def test_job_handles_leave_and_outlive_the_pool():
    """A job handle destroyed while resident leaves the pool (at once, or once
    the tick in flight that runs it is over) and frees its memory; the other
    jobs' results are unaffected (bitwise equal to their runs alone); a pool
    destroyed with jobs still resident leaves their handles to their owners,
    who destroy them afterwards."""
    print("[gpu case] serving job handle lifetime", flush=True)
    rng = np.random.default_rng(45)
    lib = S.SystemLibrary(dofs=(1, 2, 4), variants=2)
    bank = S.signal_bank(rng, 4, 3000, steps_per_sample=10)
    from pasif_amd.workloads import linear_interp_matrix
    interp = linear_interp_matrix(10)

    def new_pool():
        pool = S.Pool(lanes=256, slot_config=1, steps_per_tick=200, hbm_budget=4e9)
        pool.set_bank(bank, 1000, interp)
        pool.register_library(lib)
        return pool

    reqs = [S.draw_request(lib, rng, i, 4, tmin=300, tmax=1200, max_systems=12)
            for i in range(4)]
    alone = new_pool()
    j = alone.submit_ids(reqs[0])
    while not alone.finished(j):
        alone.tick(wait=True)
    ref = alone.collect(j, reqs[0])
    alone.close()

    pool = new_pool()
    jobs = [pool.submit_ids(r) for r in reqs]
    assert all(x is not None for x in jobs)
    used4 = pool.stats()["used_bytes"]
    # destroyed while resident, no tick in flight: leaves at once
    assert pool.lib.pasif_destroy(jobs[1]._h) == 0
    jobs[1]._h = None
    st = pool.stats()
    assert st["jobs"] == 3 and st["used_bytes"] < used4
    # destroyed while a tick runs it: the pool deletes it after the tick
    pool.tick(wait=False)
    assert pool.lib.pasif_destroy(jobs[2]._h) == 0
    jobs[2]._h = None
    pool.wait()
    assert pool.stats()["jobs"] == 2
    while not pool.finished(jobs[0]):
        pool.tick(wait=True)
    got = pool.collect(jobs[0], reqs[0])
    assert np.array_equal(got[0], ref[0])
    if reqs[0].adjoint:
        assert np.array_equal(got[1], ref[1])
    # the pool goes first with job 3 resident; its handle is destroyed after
    pool.close()
    assert pool.lib.pasif_destroy(jobs[3]._h) == 0
    jobs[3]._h = None


def test_batch_submit_prefix_and_batch_collect_equal_single_calls():
    """pasif_pool_submit_batch admits requests in order until the slot cap
    refuses one (the rest wait); pasif_pool_collect_batch returns every job's
    results bitwise equal to the same request submitted and collected alone."""
    print("[gpu case] serving batch submit / collect", flush=True)
    rng = np.random.default_rng(47)
    lib = S.SystemLibrary(dofs=(1, 2, 8, 16), variants=2)
    bank = S.signal_bank(rng, 5, 2000, steps_per_sample=10)
    from pasif_amd.workloads import linear_interp_matrix
    interp = linear_interp_matrix(10)

    def new_pool(slots=0):
        pool = S.Pool(lanes=512, slot_config=1, steps_per_tick=400, hbm_budget=4e9)
        pool.set_bank(bank, 1000, interp)
        pool.register_library(lib)
        if slots:
            pool.set_slots(slots)
        return pool

    stream = S.RequestStream(lib, 12, 5, 5, tmin=100, tmax=1200, max_systems=12)
    reqs = [stream[k] for k in range(12)]
    pool = new_pool(slots=6)
    jobs = pool.submit_batch(reqs)
    assert 0 < len(jobs) < len(reqs)          # the slot cap stops the batch
    rest = reqs[len(jobs):]
    first = reqs[:len(jobs)]
    while not all(pool.finished(j) for j in jobs):
        pool.tick(wait=True)
    pool.collect_batch(jobs, first)
    jobs2 = pool.submit_batch(rest)
    assert len(jobs2) > 0
    while not all(pool.finished(j) for j in jobs2):
        pool.tick(wait=True)
    pool.collect_batch(jobs2, rest[:len(jobs2)])
    for r in reqs[:len(jobs) + len(jobs2)]:
        solo = new_pool()
        j = solo.submit_ids(r)
        while not solo.finished(j):
            solo.tick(wait=True)
        q, g = solo.collect(j, r)
        np.testing.assert_array_equal(r.result, q)
        if r.adjoint:
            np.testing.assert_array_equal(r.grad, g)


def test_group_submit_equals_single_requests():
    """Requests admitted together as one job per adjoint flag
    (``Pool.submit_group``) and split back at ``collect_groups`` get results
    bitwise equal to each request submitted and collected alone."""
    print("[gpu case] serving group submit / collect", flush=True)
    rng = np.random.default_rng(48)
    lib = S.SystemLibrary(dofs=(1, 2, 8, 16), variants=2)
    bank = S.signal_bank(rng, 5, 2000, steps_per_sample=10)
    from pasif_amd.workloads import linear_interp_matrix
    interp = linear_interp_matrix(10)

    def new_pool():
        pool = S.Pool(lanes=512, slot_config=1, steps_per_tick=400, hbm_budget=4e9)
        pool.set_bank(bank, 1000, interp)
        pool.register_library(lib)
        return pool

    stream = S.RequestStream(lib, 14, 6, 5, tmin=100, tmax=1200, max_systems=12)
    reqs = [stream[k] for k in range(14)]
    groups = [[r for r in reqs if bool(r.adjoint) == f] for f in (False, True)]
    groups = [g for g in groups if g]
    assert len(groups) == 2
    pool = new_pool()
    jobs = [pool.submit_group(g) for g in groups]
    assert all(j is not None for j in jobs)
    # slots held fall as workgroups finish (step counts 100..1200 against
    # 400 steps per tick), to none once both jobs have finished
    held = [pool.slots()[0]]
    while not all(pool.finished(j) for j in jobs):
        pool.tick(wait=True)
        held.append(pool.slots()[0])
    pool.tick(wait=True)
    held.append(pool.slots()[0])
    assert all(b <= a for a, b in zip(held, held[1:])), held
    assert held[0] > 0 and held[-1] == 0 and len(set(held)) > 2, held
    pool.collect_groups(jobs, groups)
    for r in reqs:
        solo = new_pool()
        j = solo.submit_ids(r)
        while not solo.finished(j):
            solo.tick(wait=True)
        q, g = solo.collect(j, r)
        np.testing.assert_array_equal(r.result, q)
        if r.adjoint:
            np.testing.assert_array_equal(r.grad, g)
        else:
            assert r.grad is None


def test_fresh_request_alone_equals_mix_with_registered_and_matches_oracle():
    # requests that carry their own coefficients and unregistered sparsity
    # structures (tile pool kernel) share the launches of registered-id
    # requests (lane modules); the server admits both kinds from one stream
    print("[gpu case] fresh-coefficient requests beside registered ones", flush=True)
    from pasif_amd.workloads import linear_interp_matrix
    rng = np.random.default_rng(43)
    lib = S.SystemLibrary(dofs=(1, 2, 4, 16), variants=2)
    fresh = S.FreshSystems(24, seed=5)
    bank = S.signal_bank(rng, 6, 3000, steps_per_sample=10)
    interp = linear_interp_matrix(10)

    def new_pool():
        # 512-lane tiles: the adjoint of a 64-DOF variant spans 384 lanes
        pool = S.Pool(lanes=512, slot_config=1, steps_per_tick=150, hbm_budget=4e9)
        pool.set_bank(bank, 1000, interp)
        pool.register_library(lib)
        return pool

    fs = S.FreshRequestStream(fresh, 8, 77, 6, tmin=200, tmax=1200, max_systems=10,
                              adjoint_frac=1.0)
    target = fs[0]
    assert target.keys is None and target.adjoint
    reg = [S.draw_request(lib, rng, 100 + i, 6, tmin=200, tmax=1200, max_systems=12,
                          adjoint_frac=0.5) for i in range(4)]
    alone = new_pool()
    q1, g1 = _run_alone(alone, target)
    alone.close()
    pool = new_pool()
    order = [(reg[0], True), (fs[1], False), (target, False), (reg[1], True), (fs[2], False),
             (reg[2], True), (reg[3], True)]
    jobs = [((pool.submit_ids(r) if by_id else pool.submit(r)), r) for r, by_id in order]
    assert all(j is not None for j, _ in jobs)
    results = {}
    while jobs:
        pool.tick(wait=True)
        keep = []
        for j, r in jobs:
            if pool.finished(j):
                results[r.rid] = pool.collect(j, r)
            else:
                keep.append((j, r))
        jobs = keep
    q2, g2 = results[target.rid]
    np.testing.assert_array_equal(q1, q2)
    np.testing.assert_array_equal(g1, g2)
    oq, og = _oracle(target, bank, 1000, interp)
    assert rel_err(q1, oq) < TOL64
    assert rel_err(g1, og) < TOL64
    oq, og = _oracle(reg[0], bank, 1000, interp)
    rq, rg = results[reg[0].rid]
    assert rel_err(rq, oq) < TOL64
    if reg[0].adjoint:
        assert rel_err(rg, og) < TOL64
    pool.close()
    # the server admits a mixed stream (registered ids in batches, fresh
    # requests one submit each) and completes it
    arr = np.sort(rng.uniform(0, 0.5, 24))
    stream = S.MixedRequests(S.RequestStream(lib, 24, 5, 6, tmin=200, tmax=1200,
                                             max_systems=12),
                             S.FreshRequestStream(fresh, 24, 6, 6, tmin=200, tmax=1200,
                                                  max_systems=12), 0.4, seed=3)
    assert 0 < int(stream.is_fresh.sum()) < 24
    pool = new_pool()
    rep = S.SolveServer(pool).run(stream, arrivals=arr, max_wall_s=120)
    pool.close()
    assert rep.completed == rep.arrived == 24 and rep.backlog == 0


def test_fresh_coefficients_on_registered_structures_run_on_lane_modules():
    """A request carrying its own arrays whose sparsity structures the pool
    has registered (coefficients drawn afresh) runs on the pool's lane
    modules; its results match the oracle and are bitwise the same alone and
    beside requests of unregistered structures (tiles) and registered ids."""
    print("[gpu case] fresh coefficients on registered structures", flush=True)
    from pasif_amd.workloads import linear_interp_matrix
    rng = np.random.default_rng(49)
    lib = S.SystemLibrary(dofs=(1, 2, 4, 16), variants=2)
    same = S.FreshSystems(12, seed=8, like=lib)
    new = S.FreshSystems(24, seed=5)
    bank = S.signal_bank(rng, 6, 3000, steps_per_sample=10)
    interp = linear_interp_matrix(10)

    def new_pool():
        pool = S.Pool(lanes=512, slot_config=1, steps_per_tick=150, hbm_budget=4e9)
        pool.set_bank(bank, 1000, interp)
        pool.register_library(lib)
        return pool

    fs = S.FreshRequestStream(same, 6, 79, 6, tmin=200, tmax=1200, max_systems=10,
                              adjoint_frac=0.5)
    reqs = [fs[k] for k in range(6)]
    assert any(r.adjoint for r in reqs) and not all(r.adjoint for r in reqs)
    alone = {}
    for r in reqs:
        pool = new_pool()
        job = pool.submit(r)
        assert job._kernelInfo(r.adjoint)["kernel"] == "lane", r.rid
        while not pool.finished(job):
            pool.tick(wait=True)
        alone[r.rid] = pool.collect(job, r)
        pool.close()
        oq, og = _oracle(r, bank, 1000, interp)
        assert rel_err(alone[r.rid][0], oq) < TOL64, r.rid
        if r.adjoint:
            assert rel_err(alone[r.rid][1], og) < TOL64, r.rid
    other = S.FreshRequestStream(new, 2, 80, 6, tmin=200, tmax=1200, max_systems=10,
                                 adjoint_frac=1.0)
    ids = [S.draw_request(lib, rng, 200 + i, 6, tmin=200, tmax=1200, max_systems=12)
           for i in range(2)]
    pool = new_pool()
    jobs = [(pool.submit(other[0]), other[0]), (pool.submit_ids(ids[0]), ids[0])]
    jobs += [(pool.submit(r), r) for r in reqs]
    jobs += [(pool.submit(other[1]), other[1]), (pool.submit_ids(ids[1]), ids[1])]
    assert all(j is not None for j, _ in jobs)
    got = {}
    while jobs:
        pool.tick(wait=True)
        keep = []
        for j, r in jobs:
            if pool.finished(j):
                got[id(r)] = pool.collect(j, r)
            else:
                keep.append((j, r))
        jobs = keep
    pool.close()
    for r in reqs:
        q, g = got[id(r)]
        np.testing.assert_array_equal(q, alone[r.rid][0])
        if r.adjoint:
            np.testing.assert_array_equal(g, alone[r.rid][1])


def test_new_structures_compiled_in_the_background_then_run_on_lanes():
    """With set_compile, structures of array-carrying requests that no module
    runs are compiled into further lane modules in the background: requests
    admitted before run on tiles, the same requests admitted after the
    compile run on the new modules; both match the oracle."""
    print("[gpu case] background compilation of new structures", flush=True)
    from pasif_amd.workloads import linear_interp_matrix
    rng = np.random.default_rng(50)
    lib = S.SystemLibrary(dofs=(1, 2, 4, 16), variants=2)
    new = S.FreshSystems(40, seed=5)
    bank = S.signal_bank(rng, 6, 3000, steps_per_sample=10)
    interp = linear_interp_matrix(10)
    pool = S.Pool(lanes=512, slot_config=1, steps_per_tick=150, hbm_budget=4e9)
    pool.set_bank(bank, 1000, interp)
    pool.register_library(lib)
    pool.set_compile(True)
    # systems of at most 8 DOF: their classes fit one wave
    small = [k for k, d in enumerate(new.dofs.tolist()) if d <= 8]
    assert len(small) >= 6
    reqs = []
    for i, adj in enumerate((False, True, True)):
        pick = rng.choice(small, 4, replace=False)
        reqs.append(S.Request(900 + i, [new.systems[k] for k in pick], rng.integers(0, 6, 4),
                              rng.integers(200, 1200, 4), adj))

    def run(kind_ok):
        jobs = []
        for r in reqs:
            j = pool.submit(r)
            assert j is not None
            assert kind_ok(j._kernelInfo(r.adjoint)["kernel"]), r.rid
            jobs.append((j, r))
        out = {}
        while jobs:
            pool.tick(wait=True)
            keep = []
            for j, r in jobs:
                if pool.finished(j):
                    out[r.rid] = pool.collect(j, r)
                else:
                    keep.append((j, r))
            jobs = keep
        return out

    before = run(lambda k: k in ("tile", "mixed", "lane"))
    built, failed, queued = pool.compile_wait()
    assert built >= 1 and failed == 0 and queued == 0, (built, failed, queued)
    after = run(lambda k: k == "lane")
    pool.close()
    for r in reqs:
        oq, og = _oracle(r, bank, 1000, interp)
        for res in (before, after):
            q, g = res[r.rid]
            assert rel_err(q, oq) < TOL64, r.rid
            if r.adjoint:
                assert rel_err(g, og) < TOL64, r.rid


def test_group_of_array_requests_equals_each_alone():
    """Array-carrying requests admitted as one job (``submit_group``: fresh
    coefficients of registered structures on lane modules, new structures on
    tiles) get results bitwise equal to each request submitted alone, and
    the oracle's."""
    print("[gpu case] group of array-carrying requests", flush=True)
    from pasif_amd.workloads import linear_interp_matrix
    rng = np.random.default_rng(51)
    lib = S.SystemLibrary(dofs=(1, 2, 4, 16), variants=2)
    same = S.FreshSystems(12, seed=9, like=lib)
    new = S.FreshSystems(24, seed=6)
    bank = S.signal_bank(rng, 6, 3000, steps_per_sample=10)
    interp = linear_interp_matrix(10)

    def new_pool():
        pool = S.Pool(lanes=512, slot_config=1, steps_per_tick=150, hbm_budget=4e9)
        pool.set_bank(bank, 1000, interp)
        pool.register_library(lib)
        return pool

    a = S.FreshRequestStream(same, 3, 82, 6, tmin=200, tmax=1200, max_systems=8,
                             adjoint_frac=1.0)
    b = S.FreshRequestStream(new, 3, 83, 6, tmin=200, tmax=1200, max_systems=8,
                             adjoint_frac=1.0)
    reqs = [a[0], b[0], a[1], b[1], a[2], b[2]]
    for i, r in enumerate(reqs):
        r.rid = 700 + i
    pool = new_pool()
    job = pool.submit_group(reqs)
    assert job is not None
    while not pool.finished(job):
        pool.tick(wait=True)
    pool.collect_groups([job], [reqs])
    pool.close()
    for r in reqs:
        solo = new_pool()
        q, g = _run_alone(solo, r)
        solo.close()
        np.testing.assert_array_equal(r.result, q)
        np.testing.assert_array_equal(r.grad, g)
        oq, og = _oracle(r, bank, 1000, interp)
        assert rel_err(q, oq) < TOL64 and rel_err(g, og) < TOL64, r.rid


def test_background_compilation_of_large_new_structures():
    """Background compilation also takes structures of 32 and 64 DOF that no
    registered module runs (forward-only and forward + adjoint keys compiled
    separately); results before (tiles) and after (lane modules) match the
    oracle."""
    print("[gpu case] background compilation, 32 / 64 DOF new structures", flush=True)
    from pasif_amd.workloads import linear_interp_matrix
    rng = np.random.default_rng(52)
    lib = S.SystemLibrary(dofs=(1, 2, 4, 16), variants=2)
    bank = S.signal_bank(rng, 6, 3000, steps_per_sample=10)
    interp = linear_interp_matrix(10)
    pool = S.Pool(lanes=1024, slot_config=1, steps_per_tick=150, hbm_budget=8e9)
    pool.set_bank(bank, 1000, interp)
    pool.register_library(lib)
    pool.set_compile(True)
    sysm = [S.chain_arrays(64, 1.02, rng.normal(0, 0.05, 128), rng.uniform(0, 1, 64), True,
                           variant=3),
            S.chain_arrays(32, 0.97, rng.normal(0, 0.05, 64), rng.uniform(0, 1, 32), True,
                           variant=2)]
    reqs = [S.Request(1, sysm, np.array([1, 2]), np.array([600, 900]), True),
            S.Request(2, sysm, np.array([3, 4]), np.array([700, 500]), False)]

    def run():
        out = {}
        for r in reqs:
            j = pool.submit(r)
            kind = j._kernelInfo(r.adjoint)["kernel"]
            while not pool.finished(j):
                pool.tick(wait=True)
            out[r.rid] = (kind, pool.collect(j, r))
        return out

    before = run()
    assert all(v[0] == "tile" for v in before.values()), before
    while True:
        built, failed, queued = pool.compile_wait(1)
        print(f"  compiled {built}, failed {failed}, queued {queued}", flush=True)
        if not queued:
            break
    assert failed == 0 and built >= 2
    after = run()
    assert all(v[0] == "lane" for v in after.values()), after
    pool.close()
    for r in reqs:
        oq, og = _oracle(r, bank, 1000, interp)
        for res in (before, after):
            q, g = res[r.rid][1]
            assert rel_err(q, oq) < TOL64, r.rid
            if r.adjoint:
                assert rel_err(g, og) < TOL64, r.rid


def test_zero_order_hold_bank_on_lane_modules_matches_oracle():
    """Lane modules compiled for a zero-order-hold bank (no interpolation
    window), then rebuilt when the bank switches to interpolation: id requests
    match the oracle under each rule."""
    print("[gpu case] pool lane modules, zero-order hold then interpolation", flush=True)
    from pasif_amd.workloads import linear_interp_matrix
    rng = np.random.default_rng(53)
    lib = S.SystemLibrary(dofs=(1, 2, 4, 16), variants=2)
    bank = S.signal_bank(rng, 5, 2000, steps_per_sample=10)
    pool = S.Pool(lanes=512, slot_config=1, steps_per_tick=300, hbm_budget=4e9)
    for interp in (None, linear_interp_matrix(10)):
        pool.set_bank(bank, 1000, interp)
        if interp is None:
            pool.register_library(lib)
        reqs = [S.draw_request(lib, rng, 300 + i, 5, tmin=200, tmax=1500, max_systems=8)
                for i in range(3)]
        jobs = [pool.submit_ids(r) for r in reqs]
        assert all(j is not None for j in jobs)
        for j, r in zip(jobs, reqs):
            assert j._kernelInfo(r.adjoint)["kernel"] == "lane", r.rid
        while not all(pool.finished(j) for j in jobs):
            pool.tick(wait=True)
        for j, r in zip(jobs, reqs):
            q, g = pool.collect(j, r)
            oq, og = _oracle(r, bank, 1000, interp)
            assert rel_err(q, oq) < TOL64, (interp is None, r.rid)
            if r.adjoint:
                assert rel_err(g, og) < TOL64, (interp is None, r.rid)
    pool.close()


def test_fp32_pool_matches_the_fp32_oracle():
    """An FP32 pool (lane modules compiled for FP32) serves registered-id
    requests and requests carrying fresh coefficients of registered
    structures within 3e-2 of the FP32 oracle (the FP32 oracle's own spread
    under a one-ULP perturbation sets that bound at the full-size
    configurations; these runs are far shorter)."""
    print("[gpu case] FP32 pool", flush=True)
    from pasif_amd.workloads import linear_interp_matrix
    rng = np.random.default_rng(54)
    lib = S.SystemLibrary(dofs=(1, 2, 4, 16), variants=2)
    same = S.FreshSystems(12, seed=10, like=lib)
    bank = S.signal_bank(rng, 5, 2000, steps_per_sample=10)
    interp = linear_interp_matrix(10)
    pool = S.Pool(lanes=512, slot_config=1, steps_per_tick=300, hbm_budget=4e9, precision=32)
    pool.set_bank(bank, 1000, interp)
    pool.register_library(lib)
    reqs = [S.draw_request(lib, rng, 400 + i, 5, tmin=200, tmax=1500, max_systems=8)
            for i in range(3)]
    fs = S.FreshRequestStream(same, 2, 84, 5, tmin=200, tmax=1500, max_systems=8,
                              adjoint_frac=1.0)
    arrays = [fs[0], fs[1]]
    jobs = [(pool.submit_ids(r), r) for r in reqs] + [(pool.submit(r), r) for r in arrays]
    assert all(j is not None for j, _ in jobs)
    for j, r in jobs:
        assert j._kernelInfo(r.adjoint)["kernel"] == "lane", r.rid
    while not all(pool.finished(j) for j, _ in jobs):
        pool.tick(wait=True)
    for j, r in jobs:
        q, g = pool.collect(j, r)
        oq, og = _oracle(r, bank, 1000, interp, dtype=np.float32)
        assert rel_err(q, oq) < 3e-2, r.rid
        if r.adjoint:
            assert rel_err(g, og) < 3e-2, r.rid
    pool.close()


def test_many_new_structures_compiled_merged_and_served():
    """Many new structures at once: the background builds split them over
    several modules and, once the queue is empty, merge those into one;
    requests admitted after that run on lanes and match the oracle."""
    print("[gpu case] background compilation of many structures, merge", flush=True)
    from pasif_amd.workloads import linear_interp_matrix
    rng = np.random.default_rng(55)
    lib = S.SystemLibrary(dofs=(1, 2), variants=1)
    new = S.FreshSystems(120, seed=11, variants=12)
    bank = S.signal_bank(rng, 5, 2000, steps_per_sample=10)
    interp = linear_interp_matrix(10)
    pool = S.Pool(lanes=512, slot_config=1, steps_per_tick=400, hbm_budget=8e9)
    pool.set_bank(bank, 1000, interp)
    pool.register_library(lib)
    pool.set_compile(True)
    small = [k for k, d in enumerate(new.dofs.tolist()) if d <= 16]
    reqs = []
    for i in range(4):
        pick = small[i::4]
        reqs.append(S.Request(950 + i, [new.systems[k] for k in pick],
                              rng.integers(0, 5, len(pick)), rng.integers(100, 600, len(pick)),
                              i % 2 == 1))
    for r in reqs:                      # queue every structure
        j = pool.submit(r)
        assert j is not None
        pool.drop(j)
    while True:
        built, failed, queued = pool.compile_wait(1)
        print(f"  compiled {built}, failed {failed}, queued {queued}", flush=True)
        assert failed == 0
        if not queued:
            break
    # these structures take several modules, then one more that merges them
    assert built >= 4, built
    jobs = [(pool.submit(r), r) for r in reqs]
    for j, r in jobs:
        assert j._kernelInfo(r.adjoint)["kernel"] == "lane", r.rid
    while not all(pool.finished(j) for j, _ in jobs):
        pool.tick(wait=True)
    for j, r in jobs:
        q, g = pool.collect(j, r)
        oq, og = _oracle(r, bank, 1000, interp)
        assert rel_err(q, oq) < TOL64, r.rid
        if r.adjoint:
            assert rel_err(g, og) < TOL64, r.rid
    pool.close()
