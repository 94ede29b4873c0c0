"""The serving loop's host logic without a GPU: SolveServer against a
stand-in pool that keeps the library pool's accounting (a tick adds
steps_per_tick steps of work to every job admitted before it started; a job
of at most T steps per system is finished once that covers T, 3T with the
adjoint; a slot cap refuses submissions until jobs are collected)."""
import numpy as np

from pasif_amd import serving as S


class _Job:
    def __init__(self, req, work):
        self.req, self.work, self.done, self.joined = req, work, 0, False


class FakePool:
    ids = None

    def __init__(self, steps_per_tick, slots):
        self.steps_per_tick = steps_per_tick
        self.slots = slots
        self.jobs = []
        self.in_flight = False
        self.finished_calls = 0
        self.collected = []
        self.dropped = []
        self.tick_s = 0.0

    def submit(self, req):
        if len(self.jobs) >= self.slots:
            return None
        j = _Job(req, int(np.max(req.steps)) * (3 if req.adjoint else 1))
        self.jobs.append(j)
        return j

    def tick(self, wait):
        assert not self.in_flight
        if self.tick_s:
            import time
            time.sleep(self.tick_s)
        for j in self.jobs:
            j.joined = True
            if j.done < j.work:
                j.done += self.steps_per_tick
        self.in_flight = True
        return len(self.jobs)

    def running(self):
        return False

    def wait(self):
        self.in_flight = False
        return 1.0

    def finished(self, job):
        self.finished_calls += 1
        return not self.in_flight and job.done >= job.work

    def drop(self, job):
        self.jobs.remove(job)
        self.dropped.append(job.req.rid)

    def collect(self, job, req):
        assert job.done >= job.work and job.req is req
        self.jobs.remove(job)
        self.collected.append(req.rid)
        return np.zeros(1), (np.zeros(1) if req.adjoint else None)


def _requests(rng, n):
    out = []
    for k in range(n):
        m = int(rng.integers(1, 5))
        out.append(S.Request(k, [None] * m, np.zeros(m, dtype=np.int64),
                             rng.integers(1, 5000, m), bool(rng.uniform() < 0.7)))
    return out


def test_server_asks_only_due_jobs_and_completes_every_request():
    rng = np.random.default_rng(3)
    reqs = _requests(rng, 60)
    pool = FakePool(steps_per_tick=700, slots=8)
    rep = S.SolveServer(pool).run(reqs, arrivals=np.zeros(len(reqs)))
    assert rep.completed == rep.admitted == len(reqs) and rep.backlog == 0
    assert sorted(pool.collected) == list(range(len(reqs)))
    # each job is asked once, at the tick that covers its work
    assert pool.finished_calls == len(reqs)
    # no job waited past that tick: total ticks are what the slot cap forces
    assert rep.ticks <= sum(-(-int(np.max(r.steps)) * (3 if r.adjoint else 1) // 700)
                            for r in reqs)


def test_second_run_starts_from_an_empty_pool():
    # the first run stops at its horizon with long jobs resident: they leave
    # the pool, and a second run on the same server serves its own requests
    # against fresh tick numbers
    rng = np.random.default_rng(5)
    pool = FakePool(steps_per_tick=100, slots=4)
    pool.tick_s = 0.005
    server = S.SolveServer(pool)
    long_reqs = [S.Request(k, [None], np.zeros(1, dtype=np.int64), np.array([100000]), False)
                 for k in range(6)]
    rep1 = server.run(long_reqs, horizon_s=0.05, drain=False, arrivals=np.zeros(6))
    assert rep1.completed == 0 and rep1.admitted == 4 and rep1.backlog == 6
    assert sorted(pool.dropped) == [0, 1, 2, 3] and pool.jobs == []
    pool.tick_s = 0.0
    reqs = _requests(rng, 20)
    for r in reqs:
        r.rid += 100
    rep2 = server.run(reqs, arrivals=np.zeros(len(reqs)))
    assert rep2.completed == rep2.admitted == len(reqs) and rep2.backlog == 0
    assert sorted(pool.collected) == [r.rid for r in reqs]


class FakeBatchPool(FakePool):
    """FakePool with registered systems: id requests are admitted in
    batches, requests carrying arrays one submit each."""
    ids = np.arange(4)

    def __init__(self, steps_per_tick, slots):
        super().__init__(steps_per_tick, slots)
        self.single, self.batched = [], []

    def submit(self, req):
        assert req.keys is None
        j = super().submit(req)
        if j is not None:
            self.single.append(req.rid)
        return j

    def submit_batch(self, reqs):
        out = []
        for r in reqs:
            assert r.keys is not None
            j = FakePool.submit(self, r)
            if j is None:
                break
            out.append(j)
            self.batched.append(r.rid)
        return out

    def collect_batch(self, jobs, reqs):
        for j, r in zip(jobs, reqs):
            self.collect(j, r)


def test_server_admits_id_batches_and_array_requests_from_one_stream():
    rng = np.random.default_rng(8)
    reqs = _requests(rng, 40)
    fresh = set(int(k) for k in rng.choice(40, 13, replace=False))
    for r in reqs:
        r.keys = None if r.rid in fresh else np.zeros(len(r.steps), dtype=np.int64)
    pool = FakeBatchPool(steps_per_tick=500, slots=6)
    rep = S.SolveServer(pool).run(reqs, arrivals=np.zeros(len(reqs)))
    assert rep.completed == rep.admitted == 40 and rep.backlog == 0
    assert sorted(pool.single) == sorted(fresh)
    assert sorted(pool.batched) == sorted(set(range(40)) - fresh)
    assert sorted(pool.collected) == list(range(40))


class FakeGroupPool(FakeBatchPool):
    """FakeBatchPool that takes requests as groups (one job per group, id
    requests and array-carrying ones never mixed): a group is refused when
    the systems resident with it would pass ``max_systems``."""

    def __init__(self, steps_per_tick, slots, max_systems):
        super().__init__(steps_per_tick, slots)
        self.max_systems = max_systems
        self.groups = []

    def submit_group(self, reqs):
        assert len({bool(r.adjoint) for r in reqs}) == 1
        assert len({r.keys is None for r in reqs}) == 1
        resident = sum(sum(len(r.steps) for r in j.req) for j in self.jobs
                       if isinstance(j.req, list))
        if len(self.jobs) >= self.slots or \
                resident + sum(len(r.steps) for r in reqs) > self.max_systems:
            return None
        work = max(int(np.max(r.steps)) for r in reqs) * (3 if reqs[0].adjoint else 1)
        j = _Job(list(reqs), work)
        self.jobs.append(j)
        self.groups.append([r.rid for r in reqs])
        return j

    def collect_groups(self, jobs, groups):
        for j, g in zip(jobs, groups):
            members = j.req if isinstance(j.req, list) else [j.req]
            assert j.done >= j.work and all(x is y for x, y in zip(members, g))
            assert len(members) == len(g)
            self.jobs.remove(j)
            for r in g:
                r.result = np.full(r.nq if hasattr(r, "nq") else 1, float(r.rid))
                self.collected.append(r.rid)


def test_server_groups_requests_by_kind_and_adjoint_flag_and_halves_refused_groups():
    rng = np.random.default_rng(11)
    reqs = _requests(rng, 120)
    fresh = set(int(k) for k in rng.choice(120, 9, replace=False))
    for r in reqs:
        r.keys = None if r.rid in fresh else np.zeros(len(r.steps), dtype=np.int64)
    pool = FakeGroupPool(steps_per_tick=500, slots=6, max_systems=40)
    rep = S.SolveServer(pool).run(reqs, arrivals=np.zeros(len(reqs)))
    assert rep.completed == rep.admitted == 120 and rep.backlog == 0
    assert pool.single == []
    grouped = [k for g in pool.groups for k in g]
    assert sorted(grouped) == list(range(120))
    assert sorted(pool.collected) == list(range(120))
    # a group never passes the system cap and holds one kind and adjoint flag
    for g in pool.groups:
        assert sum(len(reqs[k].steps) for k in g) <= 40
        assert len({reqs[k].adjoint for k in g}) == 1
        assert len({k in fresh for k in g}) == 1
    assert any(len(g) > 1 for g in pool.groups)
    for r in reqs:
        if r.rid not in fresh:
            assert r.done >= r.admitted >= r.arrival
    # group=False keeps one job per request
    pool2 = FakeGroupPool(steps_per_tick=500, slots=6, max_systems=40)
    for r in reqs:
        r.result = None
    rep2 = S.SolveServer(pool2, group=False).run(reqs, arrivals=np.zeros(len(reqs)))
    assert rep2.completed == 120 and pool2.groups == []
    assert sorted(pool2.batched) == sorted(set(range(120)) - fresh)
    assert sorted(pool2.single) == sorted(fresh)


def test_groups_stay_within_group_systems():
    rng = np.random.default_rng(12)
    reqs = _requests(rng, 80)
    for r in reqs:
        r.keys = np.zeros(len(r.steps), dtype=np.int64)
    pool = FakeGroupPool(steps_per_tick=500, slots=100, max_systems=10**6)
    rep = S.SolveServer(pool, group_systems=9).run(reqs, arrivals=np.zeros(len(reqs)))
    assert rep.completed == 80 and sorted(pool.collected) == list(range(80))
    assert len(pool.groups) > 1
    for g in pool.groups:
        assert len(g) == 1 or sum(len(reqs[k].steps) for k in g) <= 9
    # queue order is kept within each adjoint flag
    for flag in (False, True):
        order = [k for g in pool.groups for k in g if reqs[k].adjoint == flag]
        assert order == sorted(order)


class _CopyingStream:
    """A request stream that builds a new object on every access, as
    RequestStream does."""

    def __init__(self, reqs):
        self.reqs = reqs

    def __len__(self):
        return len(self.reqs)

    def __getitem__(self, k):
        import copy
        return copy.copy(self.reqs[k])


def test_latency_is_measured_from_each_arrival_with_a_copying_stream():
    rng = np.random.default_rng(13)
    base = _requests(rng, 30)
    for r in base:
        r.keys = np.zeros(len(r.steps), dtype=np.int64)
        r.steps = np.minimum(r.steps, 400)
    arrivals = 0.02 * np.arange(30)
    for group in (True, False):
        pool = FakeGroupPool(steps_per_tick=500, slots=100, max_systems=10**6)
        rep = S.SolveServer(pool, group=group).run(_CopyingStream(base), arrivals=arrivals)
        assert rep.completed == 30
        # each request is served within a few ticks of its own arrival (the
        # last arrives at 0.58 s)
        assert max(rep.latencies_ms) < 250.0, (group, max(rep.latencies_ms))


def test_group_fit_follows_free_slots_and_budget():
    # Pool.group_fit without a device: the slot and byte figures come from
    # stand-ins; the per-system estimates are what earlier groups took
    pool = object.__new__(S.Pool)
    pool._slots_per_system, pool._bytes_per_system = {}, {}
    pool.hbm_budget = 1000.0
    state = {"slots": 0, "bytes": 0.0}
    pool.slots = lambda: (state["slots"], 100)
    pool.stats = lambda: {"used_bytes": state["bytes"]}
    reqs = [S.Request(k, [None] * 5, np.zeros(5, dtype=np.int64), np.full(5, 10), False,
                      keys=np.zeros(5, dtype=np.int64)) for k in range(10)]
    assert pool.group_fit(reqs) == 10            # nothing known yet: try them all
    pool._slots_per_system[(True, False)] = 2.0  # slots per system of id groups
    state["slots"] = 50                          # 50 free: 23.8 systems -> 4 requests
    assert pool.group_fit(reqs) == 4
    state["slots"] = 100                         # full, and not empty: none
    assert pool.group_fit(reqs) == 0
    state["slots"] = 0                           # empty: 47.6 systems -> 9 requests
    assert pool.group_fit(reqs) == 9
    pool._bytes_per_system[(True, False)] = 100.0    # the budget binds: 9.5 systems
    assert pool.group_fit(reqs) == 1
    state["bytes"] = 990.0                       # no bytes left, no slots held: at least one
    assert pool.group_fit(reqs) == 1
    # array requests and adjoint groups keep estimates of their own
    arr = [S.Request(k, [None] * 5, np.zeros(5, dtype=np.int64), np.full(5, 10), True)
           for k in range(3)]
    assert pool.group_fit(arr) == 3


def test_block_concatenation_matches_a_dense_block_diagonal():
    # _cat_csr / _cat_tens against scipy's block_diag and per-part index
    # shifts written out one by one
    import scipy.sparse as sp
    rng = np.random.default_rng(14)
    mats, tens, offs = [], [], []
    o = 0
    for k in range(7):
        n = int(rng.integers(1, 6))
        m = sp.random(n, n, density=0.5, random_state=int(rng.integers(1 << 30)), format="csr")
        mats.append(m)
        nnz = int(rng.integers(0, 5))
        tens.append((rng.integers(0, n, (nnz, 3)).astype(np.int64), rng.normal(size=nnz)))
        offs.append(o)
        o += n
    data, ind, ptr = S._cat_csr([(m.data, m.indices.astype(np.int64), m.indptr.astype(np.int64))
                                 for m in mats], offs)
    got = sp.csr_matrix((data, ind, ptr), shape=(o, o)).toarray()
    np.testing.assert_array_equal(got, sp.block_diag(mats).toarray())
    offs_b = [2 * x + 1 for x in offs]
    idx, val = S._cat_tens(tens, offs, offs_b, 1)
    want_i, want_v = [], []
    for (ix, v), a, b in zip(tens, offs, offs_b):
        for row, x in zip(ix, v):
            want_i.append([row[0] + a, row[1] + b, row[2] + b])
            want_v.append(x)
    np.testing.assert_array_equal(idx, np.array(want_i, dtype=np.int64).reshape(-1, 3))
    np.testing.assert_array_equal(val, np.array(want_v))
