"""Serving on the GPU: the in-kernel Poisson drive, spike counts and sampled
traces against the CPU oracle, and each job's result alone versus packed in a
continuous-batching mix (bit-exact: ordered accumulation mode)."""

import numpy as np
import pytest
import torch

from btorch_amd import workloads
from btorch_amd.models import (
    LIF, ExponentialPSC, HeterSynapsePSC, RecurrentNN, SparseConn, functional,
)
from btorch_amd.serving import LanePool, Server, SimRequest, run_plain
from oracle import rsnn

pytestmark = pytest.mark.gpu
DEV = "cuda"
DT = 0.1


def _model(n, D, seed, k_max=60):
    c = workloads.powerlaw_net_fast(n=n, k_min=5, k_max=k_max, delay_lo=1, delay_hi=D - 1,
                                    seed=seed)
    cexp = c.expanded(D)
    syn = HeterSynapsePSC(n, 1, None, SparseConn(cexp), base_psc=ExponentialPSC,
                          max_delay_steps=D, use_circular_buffer=True, tau_syn=5.0)
    return RecurrentNN(LIF(n, tau_ref=2.0), syn, step_mode="m").to(DEV), cexp


def _oracle_job(n, D, cexp, req, pool):
    """Counts and traces of one request, run alone on the CPU oracle."""
    lanes = np.arange(req.batch, dtype=np.int64)
    seeds = (int(req.seed) * 1_000_003 + lanes) & 0x7FFFFFFFFFFFFFFF
    prob = np.full(req.batch, np.float32(pool.base_p * req.rate_scale), dtype=np.float32)
    x = rsnn.poisson_drive(seeds, prob, np.zeros(req.batch, np.int64), req.T, n, pool.kick)
    net = rsnn.Net(n, cexp, rsnn.LIFParams.make(n, tau_ref=2.0), torch.tensor(5.0), DT,
                   delay_mode="hetero", max_delay_steps=D, circular=True)
    ref = rsnn.simulate(net, x)
    return ref["z"].sum(0).to(torch.int32), ref["v"][:, :, : pool.n_sample]


def test_pool_poisson_counts_and_traces_bit_exact():
    n, D = 400, 5
    model, cexp = _model(n, D, 21)
    pool = LanePool(model, lanes=6, chunk=4 * (D + 1), n_sample=7, base_rate_hz=2000.0, kick=12.0)
    req = SimRequest(0, T=53, batch=4, rate_scale=1.5, seed=99)
    assert pool.admit(req, 0.0)
    while pool.running:
        pool.step_chunk()
    (res,) = pool.retire(lambda: 0.0, block=True)
    counts, traces = _oracle_job(n, D, cexp, req, pool)
    assert int(counts.sum()) > 100
    assert torch.equal(res.counts.cpu(), counts)
    assert torch.equal(res.traces.cpu(), traces)


def test_serving_alone_vs_mix_bit_exact():
    n, D = 500, 5
    model, cexp = _model(n, D, 23)
    chunk = 6 * (D + 1)
    kw = dict(n_sample=5, base_rate_hz=1500.0, kick=12.0)
    rng = np.random.default_rng(5)
    reqs = [SimRequest(i, T=int(rng.integers(20, 140)), batch=int(rng.integers(1, 5)),
                       rate_scale=float(rng.uniform(0.5, 2.0)), seed=int(rng.integers(1 << 30)))
            for i in range(9)]
    # 8 lanes for ~20 lanes of requests: later requests are admitted into
    # lanes that earlier jobs free while the others keep running
    shared = LanePool(model, lanes=8, chunk=chunk, **kw)
    mix = Server(shared).serve(reqs, realtime=False)
    assert sorted(r.rid for r in mix.results) == list(range(9))
    admitted = sorted(r.admitted for r in mix.results)
    assert admitted[-1] > admitted[0]  # some admissions happened mid-run
    got = {r.rid: r for r in mix.results}

    alone = run_plain(lambda req: LanePool(model, lanes=req.batch, chunk=chunk, **kw), reqs)
    for r in alone.results:
        assert torch.equal(r.counts, got[r.rid].counts), r.rid
        assert torch.equal(r.traces, got[r.rid].traces), r.rid
    # and one of them against the oracle
    req = reqs[2]
    counts, traces = _oracle_job(n, D, cexp, req, shared)
    assert torch.equal(got[req.rid].counts.cpu(), counts)
    assert torch.equal(got[req.rid].traces.cpu(), traces)


def test_serving_private_networks_overlap_on_own_streams():
    n, D = 300, 5
    model, _ = _model(n, D, 25)
    chunk = 4 * (D + 1)
    kw = dict(n_sample=4, base_rate_hz=1500.0, kick=12.0)
    nets = {}

    def private(key):
        if key not in nets:
            m, _ = _model(128 + 64 * len(nets), D, 100 + len(nets))
            nets[key] = m
        return LanePool(nets[key], lanes=4, chunk=chunk, stream=torch.cuda.Stream(), **kw)

    reqs = [SimRequest(0, T=60, batch=3, seed=1), SimRequest(1, T=45, batch=2, seed=2, net="a"),
            SimRequest(2, T=70, batch=2, seed=3, net="b"), SimRequest(3, T=30, batch=2, seed=4),
            SimRequest(4, T=50, batch=3, seed=5, net="a"), SimRequest(5, T=25, batch=2, seed=6,
                                                                     net="a")]
    mix = Server(LanePool(model, lanes=6, chunk=chunk, **kw), private).serve(reqs, realtime=False)
    got = {r.rid: r for r in mix.results}
    assert sorted(got) == list(range(6))
    for req in reqs:
        if req.net == "shared":
            pool = LanePool(model, lanes=req.batch, chunk=chunk, **kw)
        else:
            pool = LanePool(nets[req.net], lanes=req.batch, chunk=chunk, **kw)
        (r,) = run_plain(lambda _r: pool, [req]).results
        assert torch.equal(r.counts, got[req.rid].counts), req.rid
        assert torch.equal(r.traces, got[req.rid].traces), req.rid


def test_serving_graph_replay_matches_eager():
    # the pool's chunk captured once in a HIP graph and replayed for every
    # chunk, with admissions, trace copies and retirements (eager work and
    # synchronisations) between replays; every job equals the eager pool's
    n, D = 400, 5
    model, cexp = _model(n, D, 27)
    chunk = 4 * (D + 1)
    kw = dict(n_sample=6, base_rate_hz=1500.0, kick=12.0)
    rng = np.random.default_rng(8)
    reqs = [SimRequest(i, T=int(rng.integers(15, 90)), batch=int(rng.integers(1, 4)),
                       rate_scale=float(rng.uniform(0.5, 2.0)), seed=int(rng.integers(1 << 30)))
            for i in range(7)]
    eager = Server(LanePool(model, lanes=6, chunk=chunk, **kw)).serve(reqs, realtime=False)
    graphed = Server(LanePool(model, lanes=6, chunk=chunk, graph=True,
                              stream=torch.cuda.Stream(), **kw)).serve(reqs, realtime=False)
    got = {r.rid: r for r in graphed.results}
    assert sorted(got) == list(range(7))
    for r in eager.results:
        assert torch.equal(r.counts, got[r.rid].counts), r.rid
        assert torch.equal(r.traces, got[r.rid].traces), r.rid
    req = reqs[0]
    counts, traces = _oracle_job(n, D, cexp, req, LanePool(model, lanes=1, chunk=chunk, **kw))
    assert torch.equal(got[req.rid].counts.cpu(), counts)
    assert torch.equal(got[req.rid].traces.cpu(), traces)


def test_serving_alpha_synapse_alone_vs_mix_bit_exact():
    # the same continuous batching with AlphaPSC kinetics (a second synapse
    # state per lane, reset when a lane is reused)
    from btorch_amd.models import AlphaPSC

    n, D = 400, 5
    c = workloads.powerlaw_net_fast(n=n, k_min=5, k_max=60, delay_lo=1, delay_hi=D - 1,
                                    seed=29)
    cexp = c.expanded(D)
    syn = HeterSynapsePSC(n, 1, None, SparseConn(cexp), base_psc=AlphaPSC,
                          max_delay_steps=D, use_circular_buffer=True, tau_syn=5.0)
    model = RecurrentNN(LIF(n, tau_ref=2.0), syn, step_mode="m").to(DEV)
    chunk = 6 * (D + 1)
    kw = dict(n_sample=5, base_rate_hz=1500.0, kick=12.0)
    rng = np.random.default_rng(8)
    reqs = [SimRequest(i, T=int(rng.integers(20, 140)), batch=int(rng.integers(1, 5)),
                       rate_scale=float(rng.uniform(0.5, 2.0)), seed=int(rng.integers(1 << 30)))
            for i in range(8)]
    shared = LanePool(model, lanes=7, chunk=chunk, **kw)
    mix = Server(shared).serve(reqs, realtime=False)
    got = {r.rid: r for r in mix.results}
    assert sorted(got) == list(range(8))
    alone = run_plain(lambda req: LanePool(model, lanes=req.batch, chunk=chunk, **kw), reqs)
    for r in alone.results:
        assert torch.equal(r.counts, got[r.rid].counts), r.rid
        assert torch.equal(r.traces, got[r.rid].traces), r.rid
    req = reqs[3]
    lanes = np.arange(req.batch, dtype=np.int64)
    seeds = (int(req.seed) * 1_000_003 + lanes) & 0x7FFFFFFFFFFFFFFF
    prob = np.full(req.batch, np.float32(shared.base_p * req.rate_scale), dtype=np.float32)
    x = rsnn.poisson_drive(seeds, prob, np.zeros(req.batch, np.int64), req.T, n, shared.kick)
    net = rsnn.Net(n, cexp, rsnn.LIFParams.make(n, tau_ref=2.0), torch.tensor(5.0), DT,
                   delay_mode="hetero", max_delay_steps=D, circular=True,
                   syn=rsnn.AlphaSyn(torch.tensor(5.0), torch.tensor(1.0)))
    ref = rsnn.simulate(net, x)
    assert int(ref["z"].sum()) > 50
    assert torch.equal(got[req.rid].counts.cpu(), ref["z"].sum(0).to(torch.int32))
    assert torch.equal(got[req.rid].traces.cpu(), ref["v"][:, :, : shared.n_sample])
