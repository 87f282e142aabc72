"""Host-side serving logic: lane allocation, admission by lanes and memory
budget, continuous admission while other jobs run.  A CPU stand-in pool
records what the server asks of it."""

from __future__ import annotations

import pytest
import torch

from btorch_amd.serving import (
    AdmissionError, LaneAllocator, Server, ServeStats, SimRequest, SimResult, run_plain,
)


def test_lane_allocator_first_fit_and_merge():
    a = LaneAllocator(10)
    assert a.take(4) == 0 and a.take(3) == 4 and a.take(4) is None
    assert a.free == [(7, 10)]
    a.give(0, 4)
    assert a.free == [(0, 4), (7, 10)] and a.free_lanes == 7
    assert a.take(4) == 0
    a.give(4, 3)
    a.give(0, 4)
    assert a.free == [(0, 10)]
    with pytest.raises(ValueError):
        a.give(2, 2)  # already free
    with pytest.raises(ValueError):
        a.take(0)


class FakePool:
    """Steps jobs chunk by chunk on the CPU; results are the job's rid."""

    def __init__(self, lanes, chunk, per_lane=100):
        self.lanes, self.chunk, self.per_lane = lanes, chunk, per_lane
        self.alloc = LaneAllocator(lanes)
        self.jobs = []  # [req, lane0, steps_done, admitted_at_chunk]
        self.chunks = 0
        self.log = []

    def bytes_per_lane(self):
        return self.per_lane

    def lanes_in_use(self):
        return self.lanes - self.alloc.free_lanes

    def admit(self, req, now):
        if req.batch > self.lanes:
            raise AdmissionError("too wide")
        a = self.alloc.take(req.batch)
        if a is None:
            return False
        self.jobs.append([req, a, 0, self.chunks])
        self.log.append(("admit", req.rid, self.chunks, a))
        return True

    def step_chunk(self):
        self.chunks += 1
        for j in self.jobs:
            j[2] += self.chunk

    def retire(self, now_fn, block=False):
        out = []
        for j in [j for j in self.jobs if j[2] >= j[0].T]:
            self.jobs.remove(j)
            self.alloc.give(j[1], j[0].batch)
            self.log.append(("done", j[0].rid, self.chunks))
            req = j[0]
            out.append(SimResult(req.rid, torch.zeros(req.batch, 1), torch.zeros(req.T, req.batch, 1),
                                 req.arrival, 0.0, now_fn()))
        return out

    @property
    def busy(self):
        return bool(self.jobs)

    @property
    def running(self):
        return any(j[2] < j[0].T for j in self.jobs)


def test_continuous_admission_while_jobs_run():
    pool = FakePool(lanes=8, chunk=10)
    reqs = [SimRequest(0, T=100, batch=6), SimRequest(1, T=20, batch=2),
            SimRequest(2, T=30, batch=2), SimRequest(3, T=10, batch=4)]
    st = Server(pool).serve(reqs, realtime=False)
    assert sorted(r.rid for r in st.results) == [0, 1, 2, 3]
    adm = {e[1]: e[2] for e in pool.log if e[0] == "admit"}
    done = {e[1]: e[2] for e in pool.log if e[0] == "done"}
    # 0 and 1 start together; 2 takes 1's lanes when it ends, before 0 ends
    assert adm[0] == adm[1] == 0 and adm[2] == done[1] < done[0]
    # 3 (4 lanes) waits, in order, until 6 lanes of 0 free up
    assert adm[3] >= done[0]
    assert st.lane_steps == sum(r.T * r.batch for r in reqs)


def test_budget_admission_and_rejection():
    pool = FakePool(lanes=16, chunk=10, per_lane=100)
    reqs = [SimRequest(0, T=50, batch=4), SimRequest(1, T=50, batch=4),
            SimRequest(2, T=10, batch=9), SimRequest(3, T=10, batch=0)]
    st = Server(pool, hbm_budget=800).serve(reqs, realtime=False)
    # 2 needs 900 B > 800 B: rejected outright rather than waiting forever;
    # 3 is malformed
    assert sorted(st.rejected) == [2, 3]
    assert sorted(r.rid for r in st.results) == [0, 1]
    # 0 and 1 fit together (800 B) and run side by side
    adm = {e[1]: e[2] for e in pool.log if e[0] == "admit"}
    assert adm[0] == adm[1] == 0

    pool = FakePool(lanes=16, chunk=10, per_lane=100)
    reqs = [SimRequest(0, T=50, batch=5), SimRequest(1, T=50, batch=4)]
    Server(pool, hbm_budget=800).serve(reqs, realtime=False)
    adm = {e[1]: e[2] for e in pool.log if e[0] == "admit"}
    done = {e[1]: e[2] for e in pool.log if e[0] == "done"}
    assert adm[1] >= done[0]  # over budget together: 1 waits for 0


def test_private_networks_without_factory_are_rejected():
    pool = FakePool(lanes=4, chunk=10)
    st = Server(pool).serve([SimRequest(0, T=10, batch=1, net="small")], realtime=False)
    assert st.rejected == [0] and st.results == []


def test_percentiles():
    s = ServeStats()
    s.results = [SimResult(i, None, None, 0.0, 0.0, float(i + 1)) for i in range(100)]
    assert s.pct(50) == 50.0 and s.pct(99) == 99.0 and s.pct(100) == 100.0


@pytest.mark.timeout(20)
def test_plain_path_runs_each_job_to_its_end():
    # one job at a time: every job's chunks are enqueued, then its result is
    # collected; the loop ends when the job has no chunk left to enqueue
    pools = {}

    def make(req):
        return pools.setdefault(req.batch, FakePool(lanes=req.batch, chunk=10))

    reqs = [SimRequest(0, T=35, batch=2), SimRequest(1, T=10, batch=3),
            SimRequest(2, T=21, batch=2)]
    st = run_plain(make, reqs)
    assert [r.rid for r in st.results] == [0, 1, 2]
    assert pools[2].chunks == 4 + 3 and pools[3].chunks == 1
