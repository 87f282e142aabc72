"""GPU checks of the module-level API: the Poisson input tensor, calls that
continue from the device-resident delay ring, state broadcasting of the
standalone steps, surrogates through the fused adjoint and the per-step
loop, and the states the training path returns."""

import math

import numpy as np
import pytest
import torch

from btorch_amd import _kernels, workloads
from btorch_amd.models import (
    ATan, Erf, ExponentialPSC, HeterSynapsePSC, LIF, RecurrentNN, SparseConn, environ,
    functional,
)
from oracle import rsnn

pytestmark = pytest.mark.gpu
DEV = "cuda"


def bits_equal(a, b) -> bool:
    a = a.detach().cpu().float()
    b = b.detach().cpu().float()
    return bool(((a == b) | (a.isnan() & b.isnan())).all())


def rel_err(a, b):
    a, b = a.detach().double().cpu(), b.detach().double().cpu()
    return float((a - b).norm() / b.norm().clamp_min(1e-30))


def _delayed_model(n, D, seed, circular=True, names=None, tau_ref=2.0):
    c = workloads.powerlaw_net_fast(n=n, k_min=5, k_max=60, delay_lo=1, delay_hi=D - 1,
                                    seed=seed)
    cexp = c.expanded(D)
    syn = HeterSynapsePSC(n, 1, None, SparseConn(cexp), base_psc=ExponentialPSC,
                          max_delay_steps=D, use_circular_buffer=circular, tau_syn=5.0)
    model = RecurrentNN(LIF(n, tau_ref=tau_ref), syn, step_mode="m",
                        update_state_names=names).to(DEV)
    net = rsnn.Net(n, cexp, rsnn.LIFParams.make(n, tau_ref=tau_ref), torch.tensor(5.0), 0.1,
                   delay_mode="hetero", max_delay_steps=D, circular=circular)
    return model, net


def test_poisson_fill_equals_oracle_drive():
    lanes, n, T = 3, 1001, 6
    seed = torch.tensor([5, 2**40 + 7, 123456789], dtype=torch.int64)
    prob = torch.tensor([0.05, 0.2, 0.5], dtype=torch.float32)
    t0 = torch.tensor([0, 17, 2**33], dtype=torch.int64)
    x = _kernels.poisson_fill(seed.to(DEV), prob.to(DEV), t0.to(DEV), 3.5, T, n)
    ref = rsnn.poisson_drive(seed.numpy(), prob.numpy(), t0.numpy(), T, n, 3.5)
    assert torch.equal(x.cpu(), ref)
    assert 0.23 < float((ref != 0).float().mean()) < 0.27  # mean of the three lane rates


def test_input_tensor_run_equals_in_kernel_drive():
    # the same numbers as an input tensor or drawn inside the step kernel
    n, B, T, D = 700, 2, 60, 5
    seeds = torch.tensor([11, 29], dtype=torch.int64, device=DEV)
    prob = torch.full((B,), 0.02, dtype=torch.float32, device=DEV)
    t0 = torch.zeros(B, dtype=torch.int64, device=DEV)
    outs = []
    for use_x in (False, True):
        model, _ = _delayed_model(n, D, 3, names=["neuron.v", "synapse.psc"])
        functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
        eng = model.fused_engine()
        with environ.context(dt=0.1), torch.no_grad():
            if use_x:
                x = _kernels.poisson_fill(seeds, prob, t0, 12.0, T, n)
                raster, rec = eng.run(x, T, B, {"z", "v", "psc"})
            else:
                drive = {"seed": seeds, "prob": prob, "t0": t0, "kick": 12.0}
                raster, rec = eng.run(None, T, B, {"z", "v", "psc"}, drive=drive)
        outs.append((raster.cpu(), rec["v"].cpu(), rec["psc"].cpu()))
    assert int(outs[0][0].ne(0).sum()) > 10
    for a, b in zip(*outs):
        assert bits_equal(a.float(), b.float()) if a.dtype != torch.int64 else torch.equal(a, b)


@pytest.mark.parametrize("circular", [True, False])
@pytest.mark.parametrize("between", ["nothing", "read", "write"])
def test_consecutive_calls_continue_the_device_ring(circular, between):
    # two calls equal one call over the whole sequence; reading the history
    # between them writes it back without changing the continuation;
    # writing it makes the second call start from the written memory
    n, B, D, T1, T2 = 900, 3, 6, 37, 41
    x = torch.from_numpy(workloads.poisson_input(T1 + T2, B, n, 400.0, 0.1, seed=5) * 25.0)
    names = ["neuron.v", "synapse.psc"]
    model, net = _delayed_model(n, D, 5, circular=circular, names=names)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    with environ.context(dt=0.1), torch.no_grad():
        z1, st1 = model(x[:T1].to(DEV))
        hist = model.synapse.history
        if between == "read":
            h_mid = hist.history.clone()
        elif between == "write":
            h = hist.history.clone()
            h.view(B, D, n)[:, :, ::7] = 1.0
            hist.history = h
        snap = {"v": model.neuron.v.cpu().clone(), "refractory": model.neuron.refractory.cpu(),
                "psc": model.synapse.base_psc.psc.cpu().clone(),
                "history": hist.history.cpu().clone(), "cursor": hist.cursor}
        z2, st2 = model(x[T1:].to(DEV))
        h_end = hist.history.cpu().clone()
        cur_end = hist.cursor
    if between == "write":
        ref2 = rsnn.simulate(net, x[T1:], state=snap)
    else:
        ref = rsnn.simulate(net, x)
        assert bits_equal(z1, ref["z"][:T1])
        ref2 = {k: v[T1:] for k, v in ref.items() if k != "final"}
        ref2["final"] = ref["final"]
        if between == "read":
            r1 = rsnn.simulate(net, x[:T1])
            assert bits_equal(h_mid, r1["final"]["history"])
    assert float(ref2["z"].sum()) > 50
    assert bits_equal(z2, ref2["z"])
    assert bits_equal(st2["neuron.v"], ref2["v"]) and bits_equal(st2["synapse.psc"], ref2["psc"])
    assert bits_equal(h_end, ref2["final"]["history"])
    assert cur_end == ref2["final"]["cursor"]


def test_state_dict_and_reset_end_the_deferred_history():
    n, B, D, T = 500, 2, 4, 30
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 400.0, 0.1, seed=8) * 25.0)
    model, net = _delayed_model(n, D, 8, names=[])
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32,
                              persistent=True)
    with environ.context(dt=0.1), torch.no_grad():
        model(x.to(DEV))
    sd = model.state_dict()
    ref = rsnn.simulate(net, x)
    assert bits_equal(sd["synapse.history.history"], ref["final"]["history"])
    functional.reset_net(model)
    assert float(model.synapse.history.history.abs().sum()) == 0.0
    with environ.context(dt=0.1), torch.no_grad():
        z, _ = model(x.to(DEV))
    assert bits_equal(z, ref["z"])


def test_unbatched_state_meets_batched_input_under_autograd():
    # memories without a batch dimension and a batched input: the standalone
    # step updates a broadcast copy of every state, never past its end
    n, B = 257, 4
    lif = LIF(n, tau=15.0, tau_ref=2.0, trainable_param={"tau"}).to(DEV)
    functional.init_net_state(lif, device=DEV, dtype=torch.float32)
    assert lif.refractory.shape == (n,)
    rng = np.random.default_rng(2)
    xs = torch.from_numpy(rng.uniform(0.0, 30.0, (3, B, n)).astype(np.float32))
    v = torch.zeros(B, n)
    ref = torch.zeros(B, n)
    p = rsnn.LIFParams.make(n, tau=15.0, tau_ref=2.0)
    with environ.context(dt=0.1):
        for t in range(3):
            xg = xs[t].to(DEV).requires_grad_(True)
            z = lif(xg)
            zr, v, ref, _ = rsnn.lif_step(v, ref, xs[t], p, 0.1)
            assert bits_equal(z, zr)
            assert lif.refractory.shape == (B, n)
            assert bits_equal(lif.refractory, ref) and bits_equal(lif.v, v)
    assert float(zr.sum()) > 0


def _grad_case(surrogate, kind, n=200, B=3, T=40, seed=21):
    c = workloads.plain_lif_net(n=n, p=0.08, seed=seed)
    rng = np.random.default_rng(seed)
    tau = torch.from_numpy(rng.uniform(10, 30, n).astype(np.float32))
    lif = LIF(n, tau=tau, tau_ref=2.0, trainable_param={"tau"}, surrogate_function=surrogate)
    model = RecurrentNN(lif, ExponentialPSC(n, 5.0, SparseConn(c.matrix())), step_mode="m",
                        update_state_names=[]).to(DEV)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 400.0, 0.1, seed=seed) * 25.0)
    R = torch.randn(T, B, n, generator=torch.Generator().manual_seed(3))
    xg = x.to(DEV).requires_grad_(True)
    with environ.context(dt=0.1):
        z, _ = model(xg)
    (z * R.to(DEV)).sum().backward()
    conn = model.synapse.linear
    dst, src = conn.indices[0].cpu(), conn.indices[1].cpu()
    mag = conn.magnitude.detach().cpu().clone().requires_grad_(True)
    tau_r = tau.clone().requires_grad_(True)
    xr = x.clone().requires_grad_(True)
    zr, *_ = rsnn.simulate_grad(n, dst, src, mag, xr, tau_r, 0.1, 5.0, tau_ref=2.0,
                                surrogate=(kind, surrogate.alpha, surrogate.damping_factor,
                                           True))
    (zr * R).sum().backward()
    assert torch.equal(z.detach().cpu(), zr.detach())
    assert zr.sum() > 20
    return [rel_err(a, b) for a, b in ((conn.magnitude.grad, mag.grad),
                                       (model.neuron.tau.grad, tau_r.grad), (xg.grad, xr.grad))]


def test_erf_surrogate_through_fused_adjoint():
    # Erf (GLIF3's reference default) in the fused adjoint kernel
    from btorch_amd.models._adjoint import fused_adjoint_covers
    sur = Erf(alpha=4.0, damping_factor=0.5)
    errs = _grad_case(sur, "erf")
    print(f"\nErf: dW {errs[0]:.2e} dtau {errs[1]:.2e} dx {errs[2]:.2e}")
    assert max(errs) < 1e-4
    n = 16
    model = RecurrentNN(LIF(n, surrogate_function=sur),
                        ExponentialPSC(n, 5.0, SparseConn(torch.eye(n).to_sparse())),
                        step_mode="m").to(DEV)
    assert fused_adjoint_covers(model.fused_engine().spec, 4)


def test_atan_surrogate_trains_through_the_step_loop():
    # no adjoint kernel for ATan: the cell falls back to the per-step loop
    from btorch_amd.models._adjoint import fused_adjoint_covers
    sur = ATan(alpha=3.0)
    n = 16
    model = RecurrentNN(LIF(n, surrogate_function=sur),
                        ExponentialPSC(n, 5.0, SparseConn(torch.eye(n).to_sparse())),
                        step_mode="m").to(DEV)
    assert not fused_adjoint_covers(model.fused_engine().spec, 4)
    errs = _grad_case(sur, "atan", T=25)
    print(f"\nATan: dW {errs[0]:.2e} dtau {errs[1]:.2e} dx {errs[2]:.2e}")
    assert max(errs) < 1e-4


def test_training_path_returns_the_same_states_as_inference():
    n, B, D, T = 300, 2, 4, 20
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 400.0, 0.1, seed=4) * 25.0)
    outs = []
    for grad in (False, True):
        model, _ = _delayed_model(n, D, 4, names=None)
        functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
        xd = x.to(DEV).requires_grad_(grad)
        with environ.context(dt=0.1), torch.set_grad_enabled(grad):
            z, st = model(xd)
        outs.append((z.detach(), {k: v.detach() for k, v in st.items()}))
    assert set(outs[0][1]) == set(outs[1][1])
    assert "synapse.history.history" in outs[0][1]
    assert bits_equal(outs[0][0], outs[1][0])
    for k in outs[0][1]:
        assert bits_equal(outs[0][1][k], outs[1][1][k]), k


@pytest.mark.parametrize("circular", [True, False])
def test_cudagraph_replay_equals_eager_across_calls(circular):
    # RecurrentNN(cudagraph=True): replayed calls equal eager calls bit for
    # bit, state and delay history carried across calls, also after a reset
    n, B, D, T = 800, 2, 5, 23
    x = torch.from_numpy(workloads.poisson_input(3 * T, B, n, 400.0, 0.1, seed=9) * 25.0)
    names = ["neuron.v", "synapse.psc"]
    runs = []
    for graph in (False, True):
        model, _ = _delayed_model(n, D, 9, circular=circular, names=names)
        model.cudagraph = graph
        functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
        outs = []
        with environ.context(dt=0.1), torch.no_grad():
            for c in range(3):
                if c == 2:
                    functional.reset_net(model)
                z, st = model(x[c * T:(c + 1) * T].to(DEV))
                outs.append((z.cpu(), {k: v.cpu() for k, v in st.items()},
                             model.synapse.history.history.cpu().clone(),
                             model.synapse.history.cursor, model.neuron.v.cpu().clone()))
        if graph:
            assert len(model._graph_runner.entries) == 1
        runs.append(outs)
    assert float(runs[0][0][0].sum()) > 50
    for a, b in zip(*runs):
        assert bits_equal(a[0], b[0])
        for k in names:
            assert bits_equal(a[1][k], b[1][k]), k
        assert bits_equal(a[2], b[2]) and a[3] == b[3] and bits_equal(a[4], b[4])


def test_cudagraph_is_inference_only():
    n = 64
    model, _ = _delayed_model(n, 3, 1, names=[])
    model.cudagraph = True
    functional.init_net_state(model, batch_size=1, device=DEV, dtype=torch.float32)
    x = torch.zeros(4, 1, n, device=DEV, requires_grad=True)
    with environ.context(dt=0.1), pytest.raises(RuntimeError, match="inference-only"):
        model(x)


def test_save_grad_history_records_state_gradients():
    # save_grad_history=True: hooks on the per-step states record
    # dL/d(state after step t); the last step's is the loss weight itself
    n, B, T = 120, 2, 15
    c = workloads.plain_lif_net(n=n, p=0.1, seed=3)
    model = RecurrentNN(LIF(n, tau_ref=2.0), ExponentialPSC(n, 5.0, SparseConn(c.matrix())),
                        step_mode="m", update_state_names=["neuron.v"], save_grad_history=True,
                        grad_state_names=["neuron.v"]).to(DEV)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 400.0, 0.1, seed=3) * 25.0)
    R = torch.randn(T, B, n, generator=torch.Generator().manual_seed(0)).to(DEV)
    with environ.context(dt=0.1):
        z, st = model(x.to(DEV).requires_grad_(True))
        (st["neuron.v"] * R).sum().backward()
    hist = model.get_grad_history()["neuron.v"]
    assert len(hist) == T and all(g is not None for g in hist)
    assert bits_equal(hist[-1], R[-1])
    # earlier steps also receive what flows back from later steps
    assert float((hist[0] - R[0]).abs().max()) > 0
    model.clear_grad_history()
    assert model.get_grad_history() == {}


def test_mixed_population_glif3_two_compartment_lif_in_fused_step():
    # a three-group MixedNeuronPopulation (GLIF3, two-compartment, LIF with
    # hard reset; group sizes not multiples of 4 or 64) in ApicalRecurrentNN
    # with delayed synapses runs in btx_rsnn_run, bit for bit with the
    # oracle's per-group loop
    from btorch_amd.models import (GLIF3, ApicalRecurrentNN, MixedNeuronPopulation,
                                   TwoCompartmentGLIF)
    sizes = (333, 402, 265)
    n, B, T, D = sum(sizes), 2, 120, 5
    c = workloads.powerlaw_net_fast(n=n, k_min=5, k_max=60, delay_lo=1, delay_hi=D - 1, seed=31)
    import scipy.sparse
    cexp = scipy.sparse.coo_array(c.expanded(D) * 5.0).astype(np.float32)
    g = torch.Generator().manual_seed(3)
    gl_kw = dict(v_threshold=1.0, v_reset=0.0, c_m=1.0, tau=torch.rand(sizes[0], generator=g) * 20
                 + 10, tau_ref=2.0, k=(0.2, 0.05), asc_amps=(-0.3, 0.1))
    tc_kw = dict(v_threshold=-50.0, v_reset=-65.0, E_L=-70.0, delta_th=2.0, tau_a=50.0,
                 tau_s=torch.rand(sizes[1], generator=g) * 20 + 10, tau_th=200.0)
    lif_kw = dict(tau=15.0, tau_ref=None, hard_reset=True)
    pop = MixedNeuronPopulation({"glif": (sizes[0], GLIF3(sizes[0], **gl_kw)),
                                 "tc": (sizes[1], TwoCompartmentGLIF(sizes[1], **tc_kw)),
                                 "lif": (sizes[2], LIF(sizes[2], **lif_kw))})
    syn = HeterSynapsePSC(n, 1, None, SparseConn(cexp), base_psc=ExponentialPSC,
                          max_delay_steps=D, use_circular_buffer=True, tau_syn=5.0)
    names = ["neuron.glif.v", "neuron.tc.v", "neuron.tc.i_a", "neuron.lif.v", "synapse.psc"]
    model = ApicalRecurrentNN(pop, syn, step_mode="m", update_state_names=names).to(DEV)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    from btorch_amd.models._engine import MixedState
    assert isinstance(model.fused_engine().spec.neuron, MixedState)
    rng = np.random.default_rng(8)
    x = torch.from_numpy((rng.random((T, B, n)) < 0.05).astype(np.float32) * 1.5)
    x[..., sizes[0]:sizes[0] + sizes[1]] = x[..., sizes[0]:sizes[0] + sizes[1]] * 40.0 + 18.0
    xa = torch.randn(T, B, n, generator=torch.Generator().manual_seed(2))
    with environ.context(dt=1.0), torch.no_grad():
        z, st = model(x.to(DEV), None, xa.to(DEV))
    gp = rsnn.GLIF3Params.make(sizes[0], **gl_kw)
    mixed = rsnn.MixedParams([(0, sizes[0], gp), (sizes[0], sizes[1], rsnn.TwoCompParams.make(
        **tc_kw)), (sizes[0] + sizes[1], sizes[2], rsnn.LIFParams.make(sizes[2], **lif_kw))])
    net = rsnn.Net(n, cexp, mixed, torch.tensor(5.0), 1.0, delay_mode="hetero",
                   max_delay_steps=D, circular=True)
    ref = rsnn.simulate(net, x, x_apical=xa)
    for s0, cnt in ((0, sizes[0]), (sizes[0], sizes[1]), (sizes[0] + sizes[1], sizes[2])):
        assert float(ref["z"][..., s0:s0 + cnt].sum()) > 20, s0
    assert bits_equal(z, ref["z"])
    spans = {"glif": slice(0, sizes[0]), "tc": slice(sizes[0], sizes[0] + sizes[1]),
             "lif": slice(sizes[0] + sizes[1], n)}
    for key in ("neuron.glif.v", "neuron.tc.v", "neuron.lif.v"):
        grp = key.split(".")[1]
        assert bits_equal(st[key], ref["v"][..., spans[grp]]), key
    assert bits_equal(st["neuron.tc.i_a"], ref["i_a"][..., spans["tc"]])
    assert bits_equal(st["synapse.psc"], ref["psc"])
    fin = ref["final"]
    assert bits_equal(pop.glif.v, fin["v"][:, spans["glif"]])
    assert bits_equal(pop.glif.Iasc, fin["iasc_groups"][0])
    assert bits_equal(pop.glif.refractory, fin["refractory"][:, spans["glif"]])
    assert bits_equal(pop.tc.theta_th, fin["theta_th"][:, spans["tc"]])
    assert bits_equal(pop.lif.v, fin["v"][:, spans["lif"]])
    assert bits_equal(syn.history.history.reshape(B, D, n), fin["history"])


def _terms_per_bin(z, cexp, n, D):
    """Terms each step's push bins per (lane, 1024-neuron segment), from the
    oracle's raster: z_{t-d}[src] for every edge of virtual row src*D + d."""
    import scipy.sparse
    T, B, _ = z.shape
    A = scipy.sparse.csr_array(cexp)
    A.data[:] = 1.0
    zf = np.zeros((T, B, n * D), dtype=np.float32)
    zn = z.numpy()
    for d in range(D):
        zf[d:, :, d::D] = zn[: T - d]
    cnt = (scipy.sparse.csr_array(zf.reshape(T * B, n * D)) @ A).toarray().reshape(T, B, n)
    segs = (n + 1023) // 1024
    pad = np.zeros((T, B, segs * 1024), dtype=np.float32)
    pad[..., :n] = cnt
    return pad.reshape(T, B, segs, 1024).sum(-1)


@pytest.mark.parametrize("bin_cap,lanes,expect", [(0, 4, "big"), (0, 4, "pull"),
                                                  (64, 4, "pull"), (64, 70, "pull")])
def test_engine_high_activity_large_bins_and_dense_pull(bin_cap, lanes, expect):
    # activity at which the push bins more terms per (lane, segment) than
    # the 1024 sorted in LDS: without the dense pull the bins are sorted
    # through the global order array; with it the step is delivered as a
    # dense pull (two 64-lane groups at 70 lanes; pull-only workgroups join
    # at 4 lanes of 3 segments); all bit-exact with the oracle
    import scipy.sparse
    n, T, D = 3000, 50, 5
    c = workloads.powerlaw_net_fast(n=n, k_min=100, k_max=600, delay_lo=1, delay_hi=D - 1,
                                    seed=43)
    cexp = scipy.sparse.coo_array(c.expanded(D) * 0.05).astype(np.float32)
    cfg = {"bin_cap": bin_cap, "dense_pull": expect == "pull"}
    syn = HeterSynapsePSC(n, 1, None, SparseConn(cexp, sparse_config=cfg),
                          base_psc=ExponentialPSC, max_delay_steps=D,
                          use_circular_buffer=True, tau_syn=5.0)
    names = ["neuron.v", "synapse.psc"]
    model = RecurrentNN(LIF(n, tau_ref=2.0), syn, step_mode="m", update_state_names=names).to(DEV)
    functional.init_net_state(model, batch_size=lanes, device=DEV, dtype=torch.float32)
    x = torch.from_numpy(workloads.poisson_input(T, lanes, n, 100.0, 0.1, seed=43) * 25.0)
    with environ.context(dt=0.1), torch.no_grad():
        z, st = model(x.to(DEV))
    ws = next(iter(model.fused_engine()._ws.values())).view(torch.int32)
    pull_flags = ws[256:260].cpu().tolist()   # Ctl.push_mode of the last launches
    net = rsnn.Net(n, cexp, rsnn.LIFParams.make(n, tau_ref=2.0), torch.tensor(5.0), 0.1,
                   delay_mode="hetero", max_delay_steps=D, circular=True)
    ref = rsnn.simulate(net, x)
    terms = _terms_per_bin(ref["z"], cexp, n, D)
    print(f"\nbin_cap {bin_cap}, {lanes} lanes: terms per bin max {terms.max():.0f} mean "
          f"{terms[10:].mean():.0f}; pull flags {pull_flags}", flush=True)
    if expect == "big":
        assert 1024 < terms.max() <= 8192 and pull_flags == [0, 0, 0, 0]
    else:
        assert pull_flags == [1, 1, 1, 1]
    assert bits_equal(z, ref["z"])
    assert bits_equal(st["neuron.v"], ref["v"]) and bits_equal(st["synapse.psc"], ref["psc"])
    assert bits_equal(model.synapse.history.history.reshape(lanes, D, n), ref["final"]["history"])
