"""GPU parity of the alpha and dual-exponential synapses (synapse.py:190-418):
standalone steps and the fused engine against the oracle, bit-exact in
float32; the sequence (convolution) form and the backward to the stated
tolerances."""

import numpy as np
import pytest
import torch

from btorch_amd import workloads
from btorch_amd.models import (
    GLIF3, LIF, AlphaPSC, AlphaPSCBilleh, DelayedPSC, DualExponentialPSC, HeterSynapsePSC,
    RecurrentNN, SparseConn, environ, functional,
)
from oracle import rsnn

pytestmark = pytest.mark.gpu
DEV = "cuda"
KINDS = ["alpha", "billeh", "dual"]


def bits_equal(a, b) -> bool:
    a = a.detach().cpu().float()
    b = b.detach().cpu().float()
    return bool(((a == b) | (a.isnan() & b.isnan())).all())


def _dt(kind):
    return 1.0 if kind == "billeh" else 0.1


def _make(kind, n, linear, per_neuron=False, **kw):
    """(module, oracle synapse) of one kinetics."""
    tau = (torch.from_numpy(np.random.default_rng(2).uniform(3, 9, n).astype(np.float32))
           if per_neuron else torch.tensor(5.0))
    with environ.context(dt=_dt(kind)):
        if kind == "alpha":
            g = kw.get("g_max", 1.3)
            return AlphaPSC(n, tau, linear, g_max=g), rsnn.AlphaSyn(tau, torch.tensor(g))
        if kind == "billeh":
            return AlphaPSCBilleh(n, tau, linear), rsnn.BillehSyn(tau)
        td = tau * 3.0
        return DualExponentialPSC(n, td, tau, linear), rsnn.DualExpSyn(td, tau)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("per_neuron", [False, True])
def test_standalone_step_bit_exact(kind, per_neuron):
    n, B, T = 300, 3, 40
    m = workloads.plain_lif_net(n=n, p=0.05, seed=21).matrix()
    mod, osyn = _make(kind, n, SparseConn(m), per_neuron)
    mod = mod.to(DEV)
    functional.init_net_state(mod, batch_size=B, device=DEV, dtype=torch.float32)
    mm = rsnn.NativeSparse(m)
    st = {"psc": torch.zeros(B, n)}
    st.update({k: torch.zeros(B, n) for k in osyn.states})
    g = torch.Generator().manual_seed(5)
    dt = _dt(kind)
    for t in range(T):
        z = (torch.rand(B, n, generator=g) < 0.1).float()
        with environ.context(dt=dt), torch.no_grad():
            out = mod(z.to(DEV))
        osyn.conductance(st, dt)
        osyn.adaptation(st, mm(z), dt)
        assert bits_equal(out, st["psc"]), (kind, t)
        for k in osyn.states:
            assert bits_equal(getattr(mod, k), st[k]), (kind, k, t)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("batch_shape", [(), (2,), (2, 3)])
def test_loop_equals_convolution_form(kind, batch_shape):
    # the reference's loop-vs-conv parity (test_synapse.py:199-223), tolerance 1e-5
    n, T = 4, 16
    lin = torch.nn.Linear(n, n, bias=False)
    torch.nn.init.eye_(lin.weight)
    mod, _ = _make(kind, n, lin)
    mod = mod.to(DEV)
    z = torch.randn(T, *batch_shape, n, generator=torch.Generator().manual_seed(42)).to(DEV)
    with environ.context(dt=1.0), torch.no_grad():
        functional.init_net_state(mod, batch_size=batch_shape or None, device=DEV,
                                  dtype=torch.float32)
        loop = torch.stack([mod.single_step_forward(z[t]) for t in range(T)])
        mod.reset()
        functional.init_net_state(mod, batch_size=batch_shape or None, device=DEV,
                                  dtype=torch.float32)
        conv = mod.multi_step_forward(z, kernel_len=64)
    torch.testing.assert_close(loop, conv, atol=1e-5, rtol=1e-5)


@pytest.mark.parametrize("kind", KINDS)
def test_standalone_backward_matches_float64(kind):
    # gradient of a weighted sum of every step's psc w.r.t. the input
    # sequence, against torch autograd of the oracle in float64 (1e-4 relative)
    n, B, T = 64, 2, 25
    m = workloads.plain_lif_net(n=n, p=0.1, seed=8).matrix()
    mod, osyn = _make(kind, n, SparseConn(m))
    mod = mod.to(DEV)
    dt = _dt(kind)
    g = torch.Generator().manual_seed(9)
    x = torch.rand(T, B, n, generator=g)
    wsum = torch.randn(T, B, n, generator=g)
    xd = x.to(DEV).requires_grad_(True)
    functional.init_net_state(mod, batch_size=B, device=DEV, dtype=torch.float32)
    with environ.context(dt=dt):
        outs = torch.stack([mod.single_step_forward(xd[t]) for t in range(T)])
    (outs * wsum.to(DEV)).sum().backward()
    s64 = osyn.to(torch.float64)
    mm = rsnn.NativeSparse(m, torch.float64)
    x64 = x.double().requires_grad_(True)
    st = {"psc": torch.zeros(B, n, dtype=torch.float64)}
    st.update({k: torch.zeros(B, n, dtype=torch.float64) for k in s64.states})
    ref = []
    for t in range(T):
        s64.conductance(st, dt)
        s64.adaptation(st, mm(x64[t]), dt)
        ref.append(st["psc"])
    (torch.stack(ref) * wsum.double()).sum().backward()
    gd, gr = xd.grad.double().cpu(), x64.grad
    assert float((gd - gr).abs().max()) <= 1e-4 * float(gr.abs().max())


def _engine_case(kind, neuron, n, B, T, seed, wrap="plain", D=6, per_neuron=False):
    dt = _dt(kind)
    if wrap == "hetero":
        c = workloads.powerlaw_net_fast(n=n, k_min=5, k_max=60, delay_lo=1, delay_hi=D - 1,
                                        seed=seed)
        conn = c.expanded(D)
    else:
        conn = workloads.plain_lif_net(n=n, p=0.05, seed=seed).matrix()
    mod, osyn = _make(kind, n, SparseConn(conn), per_neuron)
    if wrap == "hetero":
        base = type(mod)
        kw = {"alpha": dict(tau_syn=5.0, g_max=1.3), "billeh": dict(tau_syn=5.0),
              "dual": dict(tau_decay=15.0, tau_rise=5.0)}[kind]
        with environ.context(dt=dt):
            syn = HeterSynapsePSC(n, 1, None, SparseConn(conn), base_psc=base,
                                  max_delay_steps=D, use_circular_buffer=True, **kw)
        prefix = "synapse.base_psc."
        net_kw = dict(delay_mode="hetero", max_delay_steps=D, circular=True)
    elif wrap == "delayed":
        syn = DelayedPSC(mod, max_delay_steps=3)
        prefix = "synapse.psc_module."
        net_kw = dict(delay_mode="fixed", max_delay_steps=3)
    else:
        syn = mod
        prefix = "synapse."
        net_kw = {}
    names = ["neuron.v", prefix + "psc"] + [prefix + k for k in osyn.states]
    model = RecurrentNN(neuron, syn, step_mode="m", update_state_names=names).to(DEV)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    scale = 25.0 if isinstance(neuron, LIF) else 8.0
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 400.0, 0.1, seed=seed) * scale)
    with environ.context(dt=dt), torch.no_grad():
        z, st = model(x.to(DEV))
    return model, conn, osyn, x, z, st, prefix, net_kw


def _oracle_net(neuron_params, n, conn, osyn, dt, **kw):
    return rsnn.Net(n, conn, neuron_params, torch.tensor(5.0), dt, syn=osyn, **kw)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("n", [700, 703])
def test_engine_lif_kinetics_bit_exact(kind, n):
    # n = 703 runs the scalar state accesses, n = 700 the 4-wide ones
    B, T = 2, 200
    model, conn, osyn, x, z, st, prefix, _ = _engine_case(
        kind, LIF(n, tau_ref=2.0), n, B, T, 31, per_neuron=(n == 703))
    assert model.fused_engine() is not None
    ref = rsnn.simulate(_oracle_net(rsnn.LIFParams.make(n, tau_ref=2.0), n, conn, osyn,
                                    _dt(kind)), x)
    assert ref["z"].sum() > 50, "test network too quiet"
    assert bits_equal(z, ref["z"])
    assert bits_equal(st["neuron.v"], ref["v"])
    assert bits_equal(st[prefix + "psc"], ref["psc"])
    for k in osyn.states:
        assert bits_equal(st[prefix + k], ref[k]), k
        assert bits_equal(getattr(model.synapse, k), ref["final"][k]), k
    assert bits_equal(model.synapse.psc, ref["final"]["psc"])
    # a second call continues from the carried synapse state
    x2 = torch.from_numpy(workloads.poisson_input(40, B, n, 400.0, 0.1, seed=32) * 25.0)
    with environ.context(dt=_dt(kind)), torch.no_grad():
        z2, _ = model(x2.to(DEV))
    ref2 = rsnn.simulate(_oracle_net(rsnn.LIFParams.make(n, tau_ref=2.0), n, conn, osyn,
                                     _dt(kind)), x2, state=ref["final"])
    assert bits_equal(z2, ref2["z"])
    assert bits_equal(model.synapse.psc, ref2["final"]["psc"])


@pytest.mark.parametrize("kind", KINDS)
def test_engine_hetero_delays_kinetics_bit_exact(kind):
    n, B, T, D = 600, 3, 150, 6
    model, conn, osyn, x, z, st, prefix, net_kw = _engine_case(
        kind, LIF(n, tau_ref=2.0), n, B, T, 41, wrap="hetero", D=D)
    base = model.synapse.base_psc
    osyn = {"alpha": rsnn.AlphaSyn(torch.tensor(5.0), torch.tensor(1.3)),
            "billeh": rsnn.BillehSyn(torch.tensor(5.0)),
            "dual": rsnn.DualExpSyn(torch.tensor(15.0), torch.tensor(5.0))}[kind]
    ref = rsnn.simulate(_oracle_net(rsnn.LIFParams.make(n, tau_ref=2.0), n, conn, osyn,
                                    _dt(kind), **net_kw), x)
    assert ref["z"].sum() > 50
    assert bits_equal(z, ref["z"])
    assert bits_equal(st[prefix + "psc"], ref["psc"])
    for k in osyn.states:
        assert bits_equal(getattr(base, k), ref["final"][k]), k
    assert bits_equal(model.synapse.psc, ref["final"]["psc"])


def test_engine_glif3_dual_exponential_delayed_bit_exact():
    n, B, T = 400, 2, 150
    rng = np.random.default_rng(3)
    tau = torch.from_numpy(rng.uniform(10.0, 30.0, n).astype(np.float32))
    kw = dict(v_threshold=-50.0, v_reset=-70.0, c_m=0.05, tau=tau, k=[0.3, 0.02],
              asc_amps=[-0.6, 0.9], tau_ref=2.0)
    model, conn, osyn, x, z, st, prefix, net_kw = _engine_case(
        "dual", GLIF3(n, **kw), n, B, T, 51, wrap="delayed")
    ref = rsnn.simulate(_oracle_net(rsnn.GLIF3Params.make(n, **kw), n, conn, osyn, 0.1,
                                    **net_kw), x)
    assert ref["z"].sum() > 50
    assert bits_equal(z, ref["z"]) and bits_equal(st["neuron.v"], ref["v"])
    assert bits_equal(st[prefix + "psc"], ref["psc"])
    assert bits_equal(st[prefix + "g_decay"], ref["g_decay"])
    assert bits_equal(model.neuron.Iasc, ref["final"]["iasc"])


def test_alpha_cell_trains_through_the_per_step_loop():
    # the fused adjoint covers the exponential synapse; an alpha-synapse cell
    # with gradients runs the per-step loop of standalone HIP steps, whose
    # rasters and input / tau gradients match a reference-style autograd of
    # the same ops on the CPU (rel. L2 1e-4)
    import math

    from btorch_amd.models import SuperSpike

    n, B, T, dt = 200, 2, 40, 0.1
    m = workloads.plain_lif_net(n=n, p=0.05, seed=3).matrix()
    rng = np.random.default_rng(3)
    tau = torch.from_numpy(rng.uniform(10, 30, n).astype(np.float32))
    sur = SuperSpike(alpha=25.0 / (math.sqrt(2) - 1))
    mod, osyn = _make("alpha", n, SparseConn(m))
    model = RecurrentNN(LIF(n, tau=tau, tau_ref=2.0, trainable_param={"tau"},
                            surrogate_function=sur), mod, step_mode="m",
                        update_state_names=[]).to(DEV)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 400.0, 0.1, seed=3) * 25.0)
    R = torch.randn(T, B, n, generator=torch.Generator().manual_seed(6))
    xg = x.to(DEV).requires_grad_(True)
    with environ.context(dt=dt):
        z, _ = model(xg)
    (z * R.to(DEV)).sum().backward()
    # reference-style autograd, float32 on the CPU
    mm = rsnn.NativeSparse(m)
    xr = x.clone().requires_grad_(True)
    tr = tau.clone().requires_grad_(True)
    v = torch.zeros(B, n)
    ref = torch.zeros(B, n)
    st = {"psc": torch.zeros(B, n), "h": torch.zeros(B, n)}
    zs = []
    for t in range(T):
        v = v + dt * (-(v - 0.0) / tr + (st["psc"] + xr[t]) / 1.0)
        s = rsnn._Spike.apply((v - 1.0) / 1.0, "superspike", sur.alpha, 1.0, True)
        s = s * (ref == 0).detach().to(s.dtype)
        v = v - 1.0 * s
        ref = torch.relu(ref + s * 2.0 - dt)
        osyn.conductance(st, dt)
        osyn.adaptation(st, mm(s), dt)
        zs.append(s)
    zr = torch.stack(zs)
    (zr * R).sum().backward()
    assert torch.equal(z.detach().cpu(), zr.detach()) and zr.sum() > 20

    def rel(a, b):
        a, b = a.detach().double().cpu(), b.detach().double().cpu()
        return float((a - b).norm() / b.norm())
    assert rel(xg.grad, xr.grad) < 1e-4
    assert rel(model.neuron.tau.grad, tr.grad) < 1e-4


def test_engine_constrained_conn_forward_and_group_gradient():
    # SparseConstrainedConn (linear.py:594-867): group-tied magnitudes; the
    # engine's rasters equal the oracle's on the effective weights, and the
    # adjoint's magnitude gradient equals the reference-style autograd summed
    # by group (rel. L2 1e-4)
    import math

    import scipy.sparse

    from btorch_amd.models import ExponentialPSC, SparseConstrainedConn, SuperSpike

    n, B, T, G = 256, 3, 60, 5
    conn = workloads.plain_lif_net(n=n, p=0.08, seed=61).matrix()
    rng = np.random.default_rng(61)
    cons = scipy.sparse.coo_array(conn)
    cons.data = rng.integers(1, G + 1, size=cons.nnz).astype(np.float64)
    sc = SparseConstrainedConn(conn, cons)
    with torch.no_grad():
        sc.magnitude.copy_(torch.from_numpy(rng.uniform(0.5, 1.5, G).astype(np.float32)))
    sur = SuperSpike(alpha=25.0 / (math.sqrt(2) - 1))
    lif = LIF(n, tau=20.0, tau_ref=2.0, surrogate_function=sur)
    model = RecurrentNN(lif, ExponentialPSC(n, 5.0, sc), step_mode="m",
                        update_state_names=[]).to(DEV)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 400.0, 0.1, seed=61) * 25.0)
    R = torch.randn(T, B, n, generator=torch.Generator().manual_seed(1))
    with environ.context(dt=0.1):
        z, _ = model(x.to(DEV))
    (z * R.to(DEV)).sum().backward()
    dst, src = sc.indices[0].cpu(), sc.indices[1].cpu()
    gid = sc._constraint_scatter_indices.cpu()
    w0 = sc.initial_weight.cpu()
    mag = sc.magnitude.detach().cpu().clone().requires_grad_(True)
    w_eff = w0 * mag[gid]
    eff = scipy.sparse.coo_array((w_eff.detach().numpy(), (src.numpy(), dst.numpy())),
                                 shape=(n, n))
    ref = rsnn.simulate(rsnn.Net(n, eff, rsnn.LIFParams.make(n, tau_ref=2.0),
                                 torch.tensor(5.0), 0.1), x)
    assert ref["z"].sum() > 20
    assert bits_equal(z, ref["z"])
    zr, *_ = rsnn.simulate_grad(n, dst, src, w_eff, x, torch.tensor(20.0), 0.1, 5.0,
                                tau_ref=2.0, surrogate=("superspike", sur.alpha, 1.0, True))
    (zr * R).sum().backward()
    a, b = sc.magnitude.grad.double().cpu(), mag.grad.double()
    assert float((a - b).norm() / b.norm()) < 1e-4
