"""The workloads of the benchmark configurations at their full sizes on the
GPU, against the CPU oracle (bit-exact where the ordered product makes the
result exact, 1e-4 relative for gradients) and against properties that do
not depend on size (determinism, per-destination slices of the product).

  140k connectome   N = 140,000 LIF, E ~ 50M, D = 21, 8 lanes: teacher-forced
                    and free-running steps bit-exact at every step
  16k training      N = 16,384, ER p = 0.05 (13.4M synapses), batch 64:
                    dW, dtau, dx against the reference-style autograd
  serving mix       four requests on the shared 140k pool plus one on a
                    private 4096-neuron net: alone = mix, one job = oracle
  1M sheet          1,000,000 two-compartment neurons, ~1e9 synapses, D = 41:
                    two runs equal, psc of sampled destinations equal to the
                    oracle's ordered sum over their incoming slices
"""

import math

import numpy as np
import pytest
import torch

from btorch_amd import _kernels, workloads
from btorch_amd.models import (
    LIF, ExponentialPSC, HeterSynapsePSC, RecurrentNN, SparseConn, SuperSpike, environ,
    functional,
)
from btorch_amd.models._engine import pack_bits, unpack_bits
from oracle import rsnn

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]
DEV = "cuda"


def bits_equal(a, b) -> bool:
    a = a.detach().cpu().float()
    b = b.detach().cpu().float()
    return bool(((a == b) | (a.isnan() & b.isnan())).all())


def rel_err(a, b):
    a, b = a.detach().double().cpu(), b.detach().double().cpu()
    return float((a - b).norm() / b.norm().clamp_min(1e-30))


def _c1_model(lanes, names):
    conn, tau, nnz = workloads.connectome_140k(device=DEV)
    n, D = 140_000, 21
    syn = HeterSynapsePSC(n, 1, None, conn, base_psc=ExponentialPSC, max_delay_steps=D,
                          use_circular_buffer=True, tau_syn=5.0)
    model = RecurrentNN(LIF(n, tau=tau, tau_ref=2.0), syn, step_mode="m",
                        update_state_names=names).to(DEV)
    functional.init_net_state(model, batch_size=lanes, device=DEV, dtype=torch.float32)
    return model, conn, tau, nnz


def _oracle_net(conn, tau, n, D):
    dst = conn.indices[0].cpu().numpy()
    src = conn.indices[1].cpu().numpy()
    w = conn.magnitude.detach().cpu().numpy()
    mm = rsnn.NativeSparse.from_coalesced(dst, src, w, n * D, n)
    return rsnn.Net(n, mm, rsnn.LIFParams.make(n, tau=tau, tau_ref=2.0), torch.tensor(5.0), 0.1,
                    delay_mode="hetero", max_delay_steps=D, circular=True)


def _snapshot(model):
    h = model.synapse.history
    return {"v": model.neuron.v.cpu().clone(), "refractory": model.neuron.refractory.cpu().clone(),
            "psc": model.synapse.base_psc.psc.cpu().clone(), "history": h.history.cpu().clone(),
            "cursor": h.cursor}


def _restore(model, snap):
    model.neuron.v = snap["v"].to(DEV)
    model.neuron.refractory = snap["refractory"].to(DEV)
    model.synapse.base_psc.psc = snap["psc"].to(DEV)
    model.synapse.psc = snap["psc"].to(DEV)
    model.synapse.history.history = snap["history"].to(DEV)
    model.synapse.history.set_cursor(snap["cursor"])


def test_140k_connectome_teacher_forced_and_free_running_bit_exact():
    n, lanes, D, W, K = 140_000, 8, 21, 30, 6
    names = ["neuron.v", "neuron.refractory", "synapse.psc"]
    model, conn, tau, nnz = _c1_model(lanes, names)
    assert nnz > 45_000_000
    seeds = (np.arange(lanes, dtype=np.int64) * 7919 + 1009) & 0x7FFFFFFFFFFFFFFF
    x = _kernels.poisson_fill(torch.from_numpy(seeds).to(DEV),
                              torch.full((lanes,), 5e-4, dtype=torch.float32, device=DEV),
                              torch.zeros(lanes, dtype=torch.int64, device=DEV), 12.0, W + K, n)
    rng = np.random.default_rng(17)
    teacher = torch.from_numpy((rng.random((K, lanes, n)) < 2e-3).astype(np.float32))
    eng = model.fused_engine()
    with environ.context(dt=0.1), torch.no_grad():
        model(x[:W])                     # warm the net: delayed activity in flight
        snap = _snapshot(model)
        raster, rec = eng.run(x[W:].contiguous(), K, lanes, {"z", "v", "refractory", "psc"},
                              teacher=pack_bits(teacher.to(DEV)).contiguous())
        hist_t = model.synapse.history.history.cpu().clone()
        cur_t = model.synapse.history.cursor
        _restore(model, snap)
        z_f, st_f = model(x[W:].contiguous())
        hist_f = model.synapse.history.history.cpu().clone()
    net = _oracle_net(conn, tau, n, D)
    xc = x[W:].cpu()
    ref = rsnn.simulate(net, xc, state=snap, teacher=teacher)
    assert int(snap["history"].sum()) > 1000 and int(ref["z"].sum()) > 100
    assert torch.equal(unpack_bits(raster, n).cpu(), ref["z"])
    for k in ("z", "v", "refractory", "psc"):
        assert bits_equal(rec[k], ref[k]), k
    assert bits_equal(hist_t, ref["final"]["history"]) and cur_t == ref["final"]["cursor"]
    free = rsnn.simulate(net, xc, state=snap)
    agree = float((z_f.cpu() == free["z"]).float().mean())
    print(f"\n140k connectome: teacher-forced {K} steps bit-exact; free-running raster "
          f"agreement {agree:.6f} ({int(free['z'].sum())} spikes)", flush=True)
    assert bits_equal(z_f, free["z"])
    assert bits_equal(st_f["neuron.v"], free["v"]) and bits_equal(st_f["synapse.psc"], free["psc"])
    assert bits_equal(hist_f, free["final"]["history"])


def test_16k_training_gradients_match_reference_autograd():
    # full N, p and batch; T = 4 in checkpoint chunks of 2 keeps the CPU
    # oracle's autograd over 13M synapses x 64 lanes within a minute on a
    # slow host
    n, p, B, T = 16_384, 0.05, 64, 4
    c = workloads.er_net_fast(n=n, p=p, seed=4)
    rng = np.random.default_rng(4)
    w = rng.normal(0.0, 1.0 / math.sqrt(p * n), c.nnz).astype(np.float32)
    idx = torch.from_numpy(np.stack([c.pre, c.post]))
    conn = SparseConn(torch.sparse_coo_tensor(idx, torch.from_numpy(w), size=(n, n)),
                      enforce_dale=False)
    tau = torch.from_numpy(rng.uniform(10.0, 30.0, n).astype(np.float32))
    sur = SuperSpike(alpha=25.0 / (math.sqrt(2) - 1))
    model = RecurrentNN(LIF(n, tau=tau, tau_ref=2.0, trainable_param={"tau"},
                            surrogate_function=sur),
                        ExponentialPSC(n, 5.0, conn), step_mode="m", grad_checkpoint=True,
                        chunk_size=2, unroll=2, update_state_names=[]).to(DEV)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 400.0, 1.0, seed=4) * 2.0)
    R = torch.randn(T, B, n, generator=torch.Generator().manual_seed(1))
    q = torch.randn(B, n, generator=torch.Generator().manual_seed(2))
    xg = x.to(DEV).requires_grad_(True)
    with environ.context(dt=1.0):
        z, _ = model(xg)
    loss = (z * R.to(DEV)).sum() + (model.neuron.v * q.to(DEV)).sum() + \
        (model.synapse.psc * q.to(DEV)).sum()
    loss.backward()
    cm = model.synapse.linear
    dst, src = cm.indices[0].cpu(), cm.indices[1].cpu()
    mag = cm.magnitude.detach().cpu().clone().requires_grad_(True)
    tau_r = tau.clone().requires_grad_(True)
    xr = x.clone().requires_grad_(True)
    zr, vr, _, pr = rsnn.simulate_grad(n, dst, src, mag, xr, tau_r, 1.0, 5.0, tau_ref=2.0,
                                       surrogate=("superspike", sur.alpha, 1.0, True))
    ((zr * R).sum() + (vr * q).sum() + (pr * q).sum()).backward()
    assert torch.equal(z.detach().cpu(), zr.detach())
    assert zr.sum() > 1000
    errs = [rel_err(cm.magnitude.grad, mag.grad), rel_err(model.neuron.tau.grad, tau_r.grad),
            rel_err(xg.grad, xr.grad)]
    print(f"\n16k training: {int(zr.sum())} spikes; relative L2 error dW {errs[0]:.2e} "
          f"dtau {errs[1]:.2e} dx {errs[2]:.2e}", flush=True)
    assert max(errs) < 1e-4


def test_serving_mix_on_the_140k_pool_alone_equals_mix_and_oracle():
    from btorch_amd.serving import LanePool, Server, SimRequest, run_plain

    n, D = 140_000, 21
    model, conn, tau, _ = _c1_model(1, [])
    chunk = 2 * (D + 1)
    kw = dict(n_sample=6, base_rate_hz=5.0, kick=12.0)
    c = workloads.er_net_fast(n=4096, p=0.02, exc_frac=0.7, seed=100, w_mu=-3.7)
    idx = torch.from_numpy(np.stack([c.pre, c.post])).to(DEV)
    priv = RecurrentNN(LIF(4096, tau_ref=2.0),
                       ExponentialPSC(4096, 5.0, SparseConn(torch.sparse_coo_tensor(
                           idx, torch.from_numpy(c.w).to(DEV), size=(4096, 4096)))),
                       step_mode="m", update_state_names=[]).to(DEV)
    reqs = [SimRequest(0, T=50, batch=2, rate_scale=1.0, seed=11),
            SimRequest(1, T=30, batch=1, rate_scale=2.0, seed=12),
            SimRequest(2, T=70, batch=3, rate_scale=0.5, seed=13),
            SimRequest(3, T=40, batch=2, rate_scale=1.5, seed=14, net="p"),
            SimRequest(4, T=26, batch=1, rate_scale=1.0, seed=15)]
    private_pool = LanePool(priv, lanes=4, chunk=40, stream=torch.cuda.Stream(), **kw)
    shared = LanePool(model, lanes=5, chunk=chunk, graph=True, **kw)
    mix = Server(shared, lambda key: private_pool).serve(reqs, realtime=False)
    got = {r.rid: r for r in mix.results}
    assert sorted(got) == list(range(5))
    for req in reqs:
        if req.net == "shared":
            pool = LanePool(model, lanes=req.batch, chunk=chunk, **kw)
        else:
            pool = LanePool(priv, lanes=req.batch, chunk=40, **kw)
        (r,) = run_plain(lambda _r: pool, [req]).results
        assert torch.equal(r.counts, got[req.rid].counts), req.rid
        assert torch.equal(r.traces, got[req.rid].traces), req.rid
    # one shared-pool job against the oracle
    req = reqs[1]
    lanes = np.arange(req.batch, dtype=np.int64)
    seeds = (int(req.seed) * 1_000_003 + lanes) & 0x7FFFFFFFFFFFFFFF
    prob = np.full(req.batch, np.float32(shared.base_p * req.rate_scale), dtype=np.float32)
    x = rsnn.poisson_drive(seeds, prob, np.zeros(req.batch, np.int64), req.T, n, shared.kick)
    ref = rsnn.simulate(_oracle_net(conn, tau, n, D), x)
    assert int(ref["z"].sum()) > 50
    assert torch.equal(got[req.rid].counts.cpu(), ref["z"].sum(0).to(torch.int32))
    assert torch.equal(got[req.rid].traces.cpu(), ref["v"][:, :, : shared.n_sample])


def test_1m_sheet_deterministic_and_sampled_psc_equal_oracle_slices():
    from btorch_amd.models import MixedNeuronPopulation, TwoCompartmentGLIF

    side, D, T = 1000, 41, 20
    n = side * side
    pre, post, w, delay = workloads.sheet_net_device(side=side, seed=7, device=DEV)
    conn = SparseConn.from_edges(pre, post, w, delay, D, shape=(n, n))
    del pre, post, w, delay
    torch.cuda.empty_cache()
    assert conn.nnz > 900_000_000
    rng = np.random.default_rng(11)
    cell = TwoCompartmentGLIF(n, tau_s=torch.from_numpy(rng.uniform(10, 30, n).astype(np.float32)),
                              tau_th=torch.from_numpy(rng.uniform(100, 500, n).astype(np.float32)),
                              delta_th=1.0)
    syn = HeterSynapsePSC(n, 1, None, conn, base_psc=ExponentialPSC, max_delay_steps=D,
                          use_circular_buffer=True, tau_syn=5.0)
    model = RecurrentNN(MixedNeuronPopulation([(n, cell)]), syn, step_mode="m",
                        update_state_names=[]).to(DEV)
    teacher = (torch.rand((T, 1, n), generator=torch.Generator().manual_seed(5)) < 3e-4).float()
    words = pack_bits(teacher.to(DEV)).contiguous()
    drive = {"seed": torch.tensor([17], dtype=torch.int64, device=DEV),
             "prob": torch.full((1,), 3e-4, dtype=torch.float32, device=DEV),
             "t0": torch.zeros(1, dtype=torch.int64, device=DEV), "kick": 8000.0}
    outs = []
    for _ in range(2):
        functional.init_net_state(model, batch_size=1, device=DEV, dtype=torch.float32)
        eng = model.fused_engine()
        with environ.context(dt=0.1), torch.no_grad():
            raster, rec = eng.run(None, T, 1, {"psc", "v"}, drive=drive, teacher=words)
        outs.append((raster.cpu(), rec["psc"].cpu(), rec["v"].cpu()))
    assert torch.equal(outs[0][0], outs[1][0])
    assert bits_equal(outs[0][1], outs[1][1]) and bits_equal(outs[0][2], outs[1][2])
    # psc of 100 sampled destinations: exponential decay of the previous
    # step's psc plus the oracle's ordered sum over the destination's
    # incoming edges (virtual row pre*D + d carries the teacher spike of pre
    # d steps earlier)
    psc = outs[0][1][:, 0]
    js = torch.from_numpy(rng.choice(n, 100, replace=False))
    ptr = conn._csc_ptr[js.to(DEV)].cpu(), conn._csc_ptr[js.to(DEV) + 1].cpu()
    src_all, w_all = conn._csc_src, conn.effective_weight_f32()
    dst_l, rows, ws = [], [], []
    for k in range(100):
        a, b = int(ptr[0][k]), int(ptr[1][k])
        dst_l.append(np.full(b - a, k, dtype=np.int64))
        rows.append(src_all[a:b].cpu().numpy().astype(np.int64))
        ws.append(w_all[a:b].cpu().numpy())
    dst_l, rows, ws = np.concatenate(dst_l), np.concatenate(rows), np.concatenate(ws)
    assert np.all(np.diff(rows)[np.diff(dst_l) == 0] > 0)   # ascending per destination
    pre_n, d = rows // D, rows % D
    tz = teacher[:, 0].numpy()
    term_idx = np.arange(rows.size, dtype=np.int64)
    hits = 0
    prev = torch.zeros(100)
    for t in range(T):
        tt = t - d
        xv = np.where(tt >= 0, tz[np.clip(tt, 0, None), pre_n], 0.0).astype(np.float32)
        hits += int(xv.sum())
        s = rsnn.seq_spmm_c(dst_l, term_idx, ws, xv[None, :], 100)[0]
        ref = rsnn.exp_psc_decay(prev, torch.tensor(5.0), 0.1) + torch.from_numpy(s)
        got = psc[t, js]
        assert bits_equal(got, ref), t
        prev = got
    assert hits > 20
