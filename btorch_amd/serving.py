"""Continuous-batching simulation server.

Requests (trial length T, batch, input-rate scale, seed, network) arrive over
time.  Jobs on the shared connectome are packed as contiguous lanes of one
resident lane pool: every time step of every packed job is computed by the
same launch sequence (fused step, compaction, event-driven push), and the
pool advances in chunks of ``chunk`` steps.  A request is admitted as soon as
a lane range and the HBM budget allow it -- while other jobs are running --
and its lanes are reset in place; a finished job's lanes return to the pool.
Jobs on their own (small) networks get their own pool and run on their own
HIP stream, overlapping the shared pool on the device.

A job's input is the in-kernel Poisson drive, a counter-based hash of its
seed and its own step counter, and no lane reads another lane's state, so a
job's result is the same alone and in any mix (bit for bit in the ordered
accumulation mode).  Results are spike counts per neuron and the membrane
traces of the first ``n_sample`` neurons.

The plain path for comparison (``run_plain``) runs each job alone, one at a
time, through the same engine.
"""

from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field

import torch

from .models import functional
from .models._engine import Engine, match_cell


@dataclass
class SimRequest:
    rid: int
    T: int                     # trial length in steps
    batch: int                 # lanes
    rate_scale: float = 1.0    # multiplies the base background rate
    seed: int = 0
    arrival: float = 0.0       # seconds after the server starts
    net: str = "shared"        # "shared" or the key of a private network


@dataclass
class SimResult:
    rid: int
    counts: torch.Tensor       # (batch, n) int32 spike counts over the trial
    traces: torch.Tensor       # (T, batch, n_sample) membrane of neurons 0..n_sample-1
    arrival: float
    admitted: float
    finished: float

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class AdmissionError(RuntimeError):
    """A request that can never fit (lanes or memory budget)."""


class LaneAllocator:
    """First-fit allocator of contiguous lane ranges in [0, lanes)."""

    def __init__(self, lanes: int):
        self.lanes = lanes
        self.free = [(0, lanes)]  # sorted disjoint [start, end)

    def take(self, k: int) -> int | None:
        if k <= 0:
            raise ValueError("lane count must be positive")
        for i, (a, b) in enumerate(self.free):
            if b - a >= k:
                if a + k == b:
                    del self.free[i]
                else:
                    self.free[i] = (a + k, b)
                return a
        return None

    def give(self, a: int, k: int) -> None:
        if any(not (a + k <= s or a >= e) for s, e in self.free):
            raise ValueError("returning lanes that are already free")
        self.free.append((a, a + k))
        self.free.sort()
        merged = []
        for s, e in self.free:
            if merged and merged[-1][1] == s:
                merged[-1] = (merged[-1][0], e)
            else:
                merged.append((s, e))
        self.free = merged

    @property
    def free_lanes(self) -> int:
        return sum(e - s for s, e in self.free)


@dataclass
class _Job:
    req: SimRequest
    lane0: int
    steps_done: int = 0
    admitted: float = 0.0
    traces: torch.Tensor | None = None


class LanePool:
    """Resident lanes of one network; jobs occupy contiguous lane ranges."""

    def __init__(self, model, lanes: int, chunk: int, n_sample: int = 16,
                 base_rate_hz: float = 5.0, kick: float = 12.0, dt: float = 0.1,
                 stream: torch.cuda.Stream | None = None, graph: bool = False):
        spec = match_cell(model)
        if spec is None:
            raise ValueError("the pool needs a RecurrentNN cell the fused engine runs")
        if chunk % (spec.ring_d + 1):
            raise ValueError(f"chunk ({chunk}) must be a multiple of the device ring's "
                             f"{spec.ring_d + 1} slots so the ring position repeats")
        self.model, self.spec, self.engine = model, spec, Engine(spec)
        self.n = spec.neuron.size
        self.lanes, self.chunk, self.n_sample = lanes, chunk, min(n_sample, self.n)
        self.base_p = base_rate_hz * dt * 1e-3
        self.kick, self.dt = kick, dt
        self.stream = stream
        dev = spec.neuron.v.device if spec.neuron.v is not None else torch.device("cuda")
        functional.init_net_state(model, batch_size=lanes, device=dev, dtype=torch.float32)
        self.device = dev
        self.seed = torch.zeros(lanes, dtype=torch.int64, device=dev)
        self.prob = torch.zeros(lanes, dtype=torch.float32, device=dev)
        self.t0 = torch.zeros(lanes, dtype=torch.int64, device=dev)
        self.t_end = torch.zeros(lanes, dtype=torch.int64, device=dev)
        self.counts = torch.zeros((lanes, self.n), dtype=torch.int32, device=dev)
        self.sample = torch.zeros((chunk, lanes, self.n_sample), dtype=torch.float32,
                                  device=dev)
        self.alloc = LaneAllocator(lanes)
        self.jobs: list[_Job] = []
        # chunks enqueued but not known to be finished, and finished jobs
        # whose results are still being copied out (stream order keeps every
        # later use of their lanes behind those copies)
        self.inflight: deque = deque()
        self.max_inflight = 2
        self.completing: deque = deque()
        from .models import environ

        with environ.context(dt=dt):
            self.run = self.engine.resident(
                lanes, chunk, drive={"seed": self.seed, "prob": self.prob, "t0": self.t0,
                                     "kick": kick, "t_end": self.t_end},
                counts=self.counts, sample=self.sample)
        self.graph = None
        if self.stream is not None:
            # the buffers above were allocated and filled on the current
            # stream; the pool's stream must not run ahead of those fills
            self.stream.wait_stream(torch.cuda.current_stream())
        if graph:
            self._capture()

    def _capture(self) -> None:
        """Capture one chunk (the fused run and the clock advance) in a HIP
        graph.  Every lane is at rest here (v = v_reset, no current, no
        drive), so the eager warm-up launch leaves the state unchanged."""
        from .models import environ

        stream = self.stream or torch.cuda.Stream()
        stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(stream), environ.context(dt=self.dt):
            self.run.launch()
        stream.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream), environ.context(dt=self.dt):
            self.run.launch()
            self.t0 += self.chunk
        self.graph = g

    def bytes_per_lane(self) -> int:
        """Device bytes one lane holds (state, ring, accumulators, outputs)."""
        n, d = self.n, self.spec.ring_d
        slots = self.spec.conn.sparse_config.slots
        return n * (4 * 3 + 4 * d * 2 + 4 + 8 * slots + 4) + self.chunk * (n // 8 + 4 * 16)

    @torch.no_grad()
    def admit(self, req: SimRequest, now: float) -> bool:
        if req.batch > self.lanes:
            raise AdmissionError(f"request {req.rid}: batch {req.batch} exceeds pool lanes")
        lane0 = self.alloc.take(req.batch)
        if lane0 is None:
            return False
        sl = slice(lane0, lane0 + req.batch)
        ctx = torch.cuda.stream(self.stream) if self.stream is not None else _nullctx()
        with ctx:
            self._reset_lanes(sl)
            lanes = torch.arange(req.batch, device=self.device, dtype=torch.int64)
            self.seed[sl] = (int(req.seed) * 1_000_003 + lanes) & 0x7FFFFFFFFFFFFFFF
            self.prob[sl] = float(self.base_p * req.rate_scale)
            self.t0[sl] = 0
            self.t_end[sl] = int(req.T)
            job = _Job(req, lane0, 0, now)
            job.traces = torch.empty((req.T, req.batch, self.n_sample), dtype=torch.float32,
                                     device=self.device)
        self.jobs.append(job)
        return True

    def _reset_lanes(self, sl) -> None:
        self.run.reset_lanes(sl)  # state at rest, empty delay ring
        self.counts[sl] = 0
        self.prob[sl] = 0.0
        self.t_end[sl] = 0

    @torch.no_grad()
    def step_chunk(self) -> None:
        """Advance every lane by one chunk (idle lanes have no drive)."""
        from .models import environ

        ctx = torch.cuda.stream(self.stream) if self.stream is not None else _nullctx()
        with ctx, environ.context(dt=self.dt):
            if self.graph is not None:
                self.graph.replay()
            else:
                self.run.launch()
                self.t0 += self.chunk
            for job in self.jobs:
                lo = job.steps_done
                k = min(self.chunk, job.req.T - lo)
                if k > 0:
                    job.traces[lo:lo + k] = self.sample[:k, job.lane0:job.lane0 + job.req.batch]
                job.steps_done += self.chunk
            ev = torch.cuda.Event()
            ev.record(self.stream or torch.cuda.current_stream())
        # keep the host at most max_inflight chunks ahead of the device
        self.inflight.append(ev)
        while len(self.inflight) > self.max_inflight:
            self.inflight.popleft().synchronize()

    @torch.no_grad()
    def retire(self, now_fn, block: bool = False) -> list[SimResult]:
        """Results of the jobs whose last chunk has finished on the device.
        A job whose last chunk is enqueued gives its lanes back at once: the
        copy of its counts and the reset of its lanes are enqueued on the
        pool's stream, ahead of anything a later job does with them.
        ``block``: wait for the device instead of polling."""
        done = [j for j in self.jobs if j.steps_done >= j.req.T]
        if done:
            ctx = torch.cuda.stream(self.stream) if self.stream is not None else _nullctx()
            with ctx:
                for job in done:
                    sl = slice(job.lane0, job.lane0 + job.req.batch)
                    counts = self.counts[sl].clone()
                    self._reset_lanes(sl)  # idle lanes stay quiet
                    ev = torch.cuda.Event()
                    ev.record(self.stream or torch.cuda.current_stream())
                    self.completing.append((job, counts, ev))
                    self.jobs.remove(job)
                    self.alloc.give(job.lane0, job.req.batch)
        results = []
        while self.completing:
            job, counts, ev = self.completing[0]
            if block:
                ev.synchronize()
            elif not ev.query():
                break
            self.completing.popleft()
            results.append(SimResult(job.req.rid, counts, job.traces, job.req.arrival,
                                     job.admitted, now_fn()))
        return results

    @property
    def busy(self) -> bool:
        return bool(self.jobs) or bool(self.completing)

    @property
    def running(self) -> bool:
        """Some job still has chunks to enqueue."""
        return any(j.steps_done < j.req.T for j in self.jobs)

    def lanes_in_use(self) -> int:
        return self.lanes - self.alloc.free_lanes


def _sync() -> None:
    if torch.cuda.is_available():
        torch.cuda.synchronize()


class _nullctx:
    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False


@dataclass
class ServeStats:
    results: list[SimResult] = field(default_factory=list)
    wall: float = 0.0
    lane_steps: int = 0
    rejected: list[int] = field(default_factory=list)

    def latencies(self) -> list[float]:
        return sorted(r.latency for r in self.results)

    def pct(self, q: float) -> float:
        lat = self.latencies()
        if not lat:
            return float("nan")
        k = min(len(lat) - 1, max(0, math.ceil(q / 100 * len(lat)) - 1))
        return lat[k]

    @property
    def lane_steps_per_s(self) -> float:
        return self.lane_steps / self.wall if self.wall > 0 else 0.0


class Server:
    """Admits requests into pools as they arrive; advances all busy pools."""

    def __init__(self, shared_pool: LanePool, private_factory=None, hbm_budget: int | None = None):
        self.shared = shared_pool
        self.private_factory = private_factory  # net key -> LanePool (own stream)
        self.hbm_budget = hbm_budget
        self.private: dict[str, LanePool] = {}  # pools of private nets, freed when idle

    def _fits_budget(self, req: SimRequest) -> bool:
        """Admission by HBM budget: the lanes of running jobs plus this one."""
        if self.hbm_budget is None:
            return True
        per = self.shared.bytes_per_lane()
        need = req.batch * per
        if need > self.hbm_budget:
            raise AdmissionError(f"request {req.rid}: needs {need} B, budget {self.hbm_budget} B")
        used = self.shared.lanes_in_use() * per
        used += sum(p.lanes * p.bytes_per_lane() for p in self.private.values())
        return used + need <= self.hbm_budget

    def serve(self, requests: list[SimRequest], realtime: bool = True) -> ServeStats:
        """Run until every request finished.  ``realtime``: a request becomes
        visible at its arrival time on the wall clock."""
        stats = ServeStats()
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        queue: deque[SimRequest] = deque()
        _sync()
        t_start = time.perf_counter()

        def now():
            return time.perf_counter() - t_start

        while pending or queue or self.shared.busy or any(p.busy for p in self.private.values()):
            t = now()
            while pending and (not realtime or pending[0].arrival <= t):
                queue.append(pending.popleft())
            # admission, in arrival order; a request that cannot fit yet waits
            waiting = deque()
            while queue:
                req = queue.popleft()
                try:
                    if req.T < 1 or req.batch < 1:
                        raise AdmissionError(f"request {req.rid}: T and batch must be >= 1")
                    ok = self._fits_budget(req) and self._admit(req, t)
                except AdmissionError:
                    stats.rejected.append(req.rid)
                    continue
                if not ok:
                    waiting.append(req)
            queue = waiting
            ran = False
            for pool in [self.shared, *self.private.values()]:
                if pool.running:
                    pool.step_chunk()
                    ran = True
            # with nothing left to enqueue, wait for the device rather than spin
            idle = not ran and not queue and not pending
            for key, pool in list(self.private.items()):
                stats.results += pool.retire(now, block=idle)
                if not pool.busy:
                    del self.private[key]
            stats.results += self.shared.retire(now, block=idle)
            if not ran and pending and realtime:
                time.sleep(max(0.0, min(pending[0].arrival - now(), 0.002)))
        _sync()
        stats.wall = now()
        stats.lane_steps = sum(r.traces.shape[0] * r.traces.shape[1] for r in stats.results)
        return stats

    def _admit(self, req: SimRequest, t: float) -> bool:
        if req.net == "shared":
            return self.shared.admit(req, t)
        if self.private_factory is None:
            raise AdmissionError(f"request {req.rid}: no private networks configured")
        pool = self.private.get(req.net)
        if pool is None:
            pool = self.private_factory(req.net)
            self.private[req.net] = pool
        return pool.admit(req, t)


def run_plain(make_pool, requests: list[SimRequest]) -> ServeStats:
    """One job at a time, in arrival order, each alone on its own pool sized
    to its batch (the one-job-at-a-time GPU path)."""
    stats = ServeStats()
    _sync()
    t_start = time.perf_counter()

    def now():
        return time.perf_counter() - t_start

    for req in sorted(requests, key=lambda r: r.arrival):
        wait = req.arrival - now()
        if wait > 0:
            time.sleep(wait)
        pool = make_pool(req)
        pool.admit(req, now())
        while pool.running:
            pool.step_chunk()
        stats.results += pool.retire(now, block=True)
    _sync()
    stats.wall = now()
    stats.lane_steps = sum(r.traces.shape[0] * r.traces.shape[1] for r in stats.results)
    return stats


__all__ = ["SimRequest", "SimResult", "LanePool", "LaneAllocator", "Server", "ServeStats",
           "run_plain", "AdmissionError"]
