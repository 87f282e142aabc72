"""Pins the CPU oracle against the reference's own call (torch.sparse.mm),
closed forms, invariants and the reference tests' known answers."""

import math

import numpy as np
import pytest
import scipy.sparse
import torch

from oracle import rsnn


def _rand_conn(n_src, n_dst, p, seed, signed=True):
    rng = np.random.default_rng(seed)
    dense = (rng.random((n_src, n_dst)) < p) * rng.lognormal(-1, 1.0, (n_src, n_dst))
    if signed:
        dense = dense * rng.choice([-1.0, 1.0], size=dense.shape)
    return scipy.sparse.coo_array(dense.astype(np.float32))


@pytest.mark.parametrize("values", ["spikes", "real"])
@pytest.mark.parametrize("lanes", [1, 8, 64])
def test_sparse_mm_equals_sequential_fma(values, lanes):
    # linear.py:474-483 sums each destination in ascending source order with
    # one fmaf per term; the C restatement must agree bit for bit.
    conn = _rand_conn(300, 200, 0.3, 0)
    rng = np.random.default_rng(lanes)
    if values == "spikes":
        x = (rng.random((lanes, 300)) < 0.3).astype(np.float32)
    else:
        x = (rng.random((lanes, 300)) * 3 - 1).astype(np.float32)
    mm = rsnn.NativeSparse(conn)
    out = mm(torch.from_numpy(x)).numpy()
    ref = rsnn.seq_spmm_c(mm.dst, mm.src, mm.w, x, 200)
    assert np.array_equal(out.view(np.int32), ref.view(np.int32))


def test_sequential_fma_catches_reordering():
    # A descending order differs somewhere: the check above is sensitive to order.
    conn = _rand_conn(400, 50, 0.5, 3)
    x = np.random.default_rng(1).random((4, 400)).astype(np.float32)
    mm = rsnn.NativeSparse(conn)
    out = mm(torch.from_numpy(x)).numpy()
    rev = rsnn.seq_spmm_c(mm.dst[::-1], mm.src[::-1], mm.w[::-1], x, 50)
    assert not np.array_equal(out, rev)


def test_sparse_mm_matches_dense_float64():
    # test_linear.py:24-96: sparse product equals x @ W.
    conn = _rand_conn(120, 90, 0.1, 5)
    x = np.random.default_rng(2).random((7, 120)).astype(np.float32)
    out = rsnn.NativeSparse(conn)(torch.from_numpy(x)).numpy()
    ref = x.astype(np.float64) @ conn.toarray().astype(np.float64)
    np.testing.assert_allclose(out, ref, rtol=1e-5, atol=1e-6)


def test_delay_expansion_known_answers():
    # tests/connectome/test_delay_expansion.py:34-107
    conn = scipy.sparse.coo_array(([5.0], ([0], [1])), shape=(3, 3))
    d = rsnn.expand_conn_for_delays(conn, np.array([2]), 5).toarray()
    assert d.shape == (15, 3) and d[2, 1] == 5.0 and d[:, 1].sum() == 5.0
    conn = scipy.sparse.coo_array(([1.0, 2.0, 3.0, 4.0], ([0, 0, 1, 2], [1, 2, 0, 1])),
                                  shape=(3, 3))
    d = rsnn.expand_conn_for_delays(conn, np.array([0, 1, 2, 3]), 5).toarray()
    assert d[0, 1] == 1.0 and d[1, 2] == 2.0 and d[7, 0] == 3.0 and d[13, 1] == 4.0
    conn = scipy.sparse.coo_array(([1.0, 2.0], ([0, 1], [1, 2])), shape=(3, 3))
    d = rsnn.expand_conn_for_delays(conn, np.array([10, 100]), 5).toarray()
    assert d[4, 1] == 1.0 and d[9, 2] == 2.0
    with pytest.raises(ValueError):
        rsnn.expand_conn_for_delays(conn, np.array([-1, 0]), 5)


def test_ring_and_expanded_matrix_round_trip():
    # test_delay_expansion.py:342-389: spike of neuron 0 at t=0 reaches
    # neuron 1 through a delay-2 synapse when read at t=3.
    conn = scipy.sparse.coo_array(([2.0], ([0], [1])), shape=(2, 2))
    cd = rsnn.expand_conn_for_delays(conn, np.array([2]), 3)
    for circular in (True, False):
        ring = rsnn.Ring(1, 2, 3, circular)
        ring.update(torch.tensor([[1.0, 0.0]]))
        ring.update(torch.tensor([[0.0, 0.0]]))
        ring.update(torch.tensor([[0.0, 0.0]]))
        zf = ring.get_flattened(3)
        assert zf.tolist() == [[0.0, 0.0, 1.0, 0.0, 0.0, 0.0]]
        out = rsnn.NativeSparse(cd)(zf)
        assert out.tolist() == [[0.0, 2.0]]


def test_exp_psc_closed_form_with_latency():
    # test_synapse.py:21-83: psc_{t+1} = psc_t e^{-dt/tau} + spike_{t-l}
    dt, tau, lat = 1.0, 2.0, 3
    z = torch.tensor([[1.0, 1.0, 0.5], [1.0, 0.0, 0.5], [0.0, 1.0, -0.5], [1.0, 1.0, 0.0],
                      [0.0, 0.0, 2.0]])
    ring = rsnn.Ring(1, 3, lat + 1, circular=False)
    psc = torch.zeros(1, 3)
    a = math.exp(-dt / tau)
    exp = torch.zeros(3, dtype=torch.float64)
    for t in range(5):
        ring.update(z[t][None])
        psc = rsnn.exp_psc_decay(psc, torch.tensor(tau), dt) + ring.get_delay(lat)
        exp = exp * a + (z[t - lat].double() if t >= lat else 0.0)
        np.testing.assert_allclose(psc[0].numpy(), exp.numpy(), atol=1e-6)


def test_delayed_single_neuron_exact_values():
    # test_delayed_psc.py:493-531: delay 3, tau_syn 1, dt 1
    ring = rsnn.Ring(1, 1, 4, circular=False)
    psc = torch.zeros(1, 1)
    seq = [1.0, 0.0, 0.0, 0.0, 2.0]
    outs = []
    for s in seq:
        ring.update(torch.tensor([[s]]))
        psc = rsnn.exp_psc_decay(psc, torch.tensor(1.0), 1.0) + ring.get_delay(3)
        outs.append(psc.item())
    assert outs[:3] == [0.0, 0.0, 0.0]
    assert abs(outs[3] - 1.0) < 1e-6
    assert abs(outs[4] - math.exp(-1.0)) < 1e-6


def test_lif_euler_closed_form():
    # Constant input I, no spike: v_n = I*tau*c^{-1}... Euler recursion
    # v_{n+1} = v_n (1 - dt/tau) + dt I  =>  v_n = I tau (1 - (1 - dt/tau)^n).
    p = rsnn.LIFParams.make(1, tau=20.0, v_threshold=1e9, dtype=torch.float64)
    v = torch.zeros(1, 1, dtype=torch.float64)
    ref = torch.zeros_like(v)
    I, dt = 0.7, 0.1
    for k in range(1, 200):
        _, v, ref, _ = rsnn.lif_step(v, ref, torch.full_like(v, I), p, dt)
        closed = I * 20.0 * (1 - (1 - dt / 20.0) ** k)
        assert abs(v.item() - closed) < 1e-12


def test_lif_reset_and_refractory_invariants():
    # soft reset subtracts (v_th - v_reset); hard reset lands on v_reset;
    # no spike while the refractory counter is positive, counter = tau_ref - k dt.
    for hard in (False, True):
        p = rsnn.LIFParams.make(1, tau=20.0, v_threshold=1.0, v_reset=-0.5, tau_ref=2.0,
                                hard_reset=hard, dtype=torch.float64)
        v = torch.zeros(1, 1, dtype=torch.float64)
        ref = torch.zeros_like(v)
        s, v2, ref2, v_pre = rsnn.lif_step(v, ref, torch.full_like(v, 30.0), p, 0.1)
        assert s.item() == 1.0
        expect = -0.5 if hard else v_pre.item() - 1.5
        assert abs(v2.item() - expect) < 1e-12
        assert abs(ref2.item() - 1.9) < 1e-12
        spikes = []
        for _ in range(25):
            s, v2, ref2, _ = rsnn.lif_step(v2, ref2, torch.full_like(v, 30.0), p, 0.1)
            spikes.append(s.item())
        first = spikes.index(1.0)
        assert first == 19  # 1.9 ms remaining at dt = 0.1 -> 19 silent steps


def test_float64_oracle_tracks_float32():
    from btorch_amd import workloads
    net_c = workloads.plain_lif_net(n=256, p=0.05, seed=3)
    x = torch.from_numpy(workloads.poisson_input(100, 2, 256, 200.0, 0.1, seed=3) * 12.0)
    net = rsnn.Net(256, net_c.matrix(), rsnn.LIFParams.make(256, tau_ref=2.0),
                   torch.tensor(5.0), 0.1)
    a = rsnn.simulate(net, x)
    b = rsnn.simulate(net, x, dtype=torch.float64)
    agree = (a["z"].double() == b["z"]).double().mean().item()
    assert agree > 0.999
    assert a["z"].sum() > 0


def test_mix64_splitmix64_known_answers():
    # SplitMix64 with seed 0: outputs mix64(k * golden gamma), k = 1, 2, 3
    # (Steele, Lea, Flood 2014; Vigna's reference splitmix64.c)
    g = 0x9E3779B97F4A7C15
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    got = [int(rsnn.mix64(np.uint64((g * k) % 2**64))) for k in (1, 2, 3)]
    assert got == want


def test_poisson_drive_invariants():
    n, T = 4096, 40
    p = np.float32(0.01)
    d = rsnn.poisson_drive([7, 8, 7], [p, p, p], [0, 0, 0], T, n, 12.0)
    assert set(np.unique(d.numpy())) <= {0.0, 12.0}
    # a lane's drive depends on its own seed and step only
    assert torch.equal(d[:, 0], d[:, 2]) and not torch.equal(d[:, 0], d[:, 1])
    # the step counter is the lane's own: starting at t0 skips t0 steps
    d5 = rsnn.poisson_drive([7], [p], [5], T - 5, n, 12.0)
    assert torch.equal(d5[:, 0], d[5:, 0])
    # Bernoulli(p) rate within 6 sigma; no draws at p = 0
    k = float((d[:, :2] > 0).sum())
    m = 2 * T * n
    assert abs(k / m - p) < 6 * math.sqrt(p * (1 - p) / m)
    assert float(rsnn.poisson_drive([1], [0.0], [0], 5, n, 12.0).sum()) == 0.0


def test_glif3_no_spike_matches_exponential_form():
    # Below threshold, exponential Euler with the after-spike currents held
    # over the step is v' = v_rest + (v - v_rest) e + (1 - e) tau (x + sum I)/c_m,
    # e = exp(-dt/tau), and I' = I exp(-k dt): a different formula (exp, not
    # expm1) evaluated independently in float64.
    dt, n, T = 0.1, 5, 300
    tau = np.array([5.0, 10.0, 20.0, 35.0, 80.0])
    k = np.array([0.2, 0.03])
    c_m, v_rest, x = 0.05, -70.0, 0.3
    p = rsnn.GLIF3Params.make(n, tau=torch.from_numpy(tau), v_threshold=1e6, v_reset=-70.0,
                              c_m=c_m, tau_ref=None, k=torch.from_numpy(k),
                              asc_amps=(1.0, -2.0), dtype=torch.float64)
    v = torch.full((1, n), -65.0, dtype=torch.float64)
    ia = torch.tensor([[[0.4, -0.7]] * n], dtype=torch.float64)
    ref = torch.zeros(1, n, dtype=torch.float64)
    xv = torch.full((1, n), x, dtype=torch.float64)
    ve, ie = v.numpy()[0].copy(), ia.numpy()[0].copy()
    e = np.exp(-dt / tau)
    for _ in range(T):
        z, v, ia, ref, _ = rsnn.glif3_step(v, ia, ref, xv, p, dt)
        ve = v_rest + (ve - v_rest) * e + (1 - e) * tau * (x + ie.sum(-1)) / c_m
        ie = ie * np.exp(-k * dt)
        assert float(z.sum()) == 0.0
    np.testing.assert_allclose(v.numpy()[0], ve, rtol=1e-12, atol=1e-10)
    np.testing.assert_allclose(ia.numpy()[0], ie, rtol=1e-12, atol=1e-14)


def test_glif3_spike_resets_and_asc_jump():
    dt = 0.1
    p = rsnn.GLIF3Params.make(3, tau=20.0, v_threshold=-50.0, v_reset=-70.0, c_m=0.05,
                              tau_ref=1.0, k=(0.2, 0.05), asc_amps=(1.5, -0.5),
                              dtype=torch.float64)
    v = torch.tensor([[-49.0, -60.0, -49.5]], dtype=torch.float64)
    ia = torch.zeros(1, 3, 2, dtype=torch.float64)
    ref = torch.zeros(1, 3, dtype=torch.float64)
    x = torch.zeros(1, 3, dtype=torch.float64)
    z, v1, ia1, ref1, v_pre = rsnn.glif3_step(v, ia, ref, x, p, dt)
    assert z.tolist() == [[1.0, 0.0, 1.0]]
    # soft reset subtracts the threshold gap; asc jump by the amplitudes
    np.testing.assert_allclose((v_pre - v1)[0].numpy(), [20.0, 0.0, 20.0], atol=1e-12)
    np.testing.assert_allclose(ia1[0, 0].numpy(), [1.5, -0.5], atol=1e-15)
    np.testing.assert_allclose(ia1[0, 1].numpy(), [0.0, 0.0], atol=1e-15)
    np.testing.assert_allclose(ref1[0].numpy(), [0.9, 0.0, 0.9], atol=1e-12)
    # refractory: above threshold again but no spike
    z2, *_ = rsnn.glif3_step(torch.full((1, 3), -40.0, dtype=torch.float64), ia1, ref1, x, p, dt)
    assert z2.tolist() == [[0.0, 1.0, 0.0]]


def test_delayed_autograd_restatement_matches_forward_oracle():
    # simulate_grad's delayed synapse (a spike stack per virtual row) runs the
    # same forward as simulate's SpikeHistory ring, whose layout the delay
    # expansion known answers pin
    from btorch_amd import workloads

    n, B, T, D = 60, 2, 40, 4
    c = workloads.powerlaw_net_fast(n=n, k_min=5, k_max=20, delay_lo=1, delay_hi=D - 1, seed=3)
    cexp = c.expanded(D)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 800.0, 0.1, seed=3) * 25.0)
    net = rsnn.Net(n, cexp, rsnn.LIFParams.make(n, tau_ref=2.0), torch.tensor(5.0), 0.1,
                   delay_mode="hetero", max_delay_steps=D, circular=True)
    ref = rsnn.simulate(net, x)
    dst, src, w = rsnn.coalesce(cexp)
    z, v, r, psc = rsnn.simulate_grad(n, dst, src, torch.from_numpy(w), x,
                                      torch.tensor(20.0), 0.1, 5.0, tau_ref=2.0,
                                      csr_d=D, ages=tuple(range(D)))
    assert float(ref["z"].sum()) > 20
    assert torch.equal(z, ref["z"]) and torch.equal(v, ref["final"]["v"])
    assert torch.equal(psc, ref["final"]["psc"])


# ---------------------------------------------------------------------------
# Two-compartment neuron (two_compartment.py:208-292)
# ---------------------------------------------------------------------------

def _tc_run(p, T, soma, apical, dt, dtype, v0=None, ia0=0.0, th0=0.0):
    shape = (1, 4)
    v = torch.full(shape, float(p.v_reset) if v0 is None else v0, dtype=dtype)
    ia = torch.full(shape, ia0, dtype=dtype)
    ib = torch.zeros(shape, dtype=dtype)
    th = torch.full(shape, th0, dtype=dtype)
    pp = p.to(dtype)
    out = []
    for t in range(T):
        s, v, ia, ib, th, vp = rsnn.tc_step(v, ia, ib, th, torch.full(shape, soma, dtype=dtype),
                                            torch.full(shape, apical, dtype=dtype), pp, dt)
        out.append((s, v, ia, ib, th, vp))
    return out


def test_tc_linear_regime_matches_closed_form():
    # No spike (threshold out of reach), delta_T = 0, w_Ca = 0: both
    # compartments are first-order implicit-Euler filters, so with q = 1/(1+da),
    # r = 1/(1+ds):  i_a_t = A + (i_a0 - A) q^t  and, with u = v - v*,
    # v* = E_L + R_s (I + w_as A):  u_t = r^t u_0 + (1-r) R_s w_as c q (r^t - q^t)/(r - q).
    dt, tau_a, tau_s, R, EL, w_as, I, A = 0.5, 7.0, 3.0, 1.7, -70.0, 0.8, 4.0, 2.5
    p = rsnn.TwoCompParams.make(v_threshold=1e6, tau_a=tau_a, tau_s=tau_s, R_s=R, E_L=EL,
                                w_as=w_as, dtype=torch.float64)
    ia0, v0 = 1.25, -60.0
    out = _tc_run(p, 60, I, A, dt, torch.float64, v0=v0, ia0=ia0)
    q, r = 1 / (1 + dt / tau_a), 1 / (1 + dt / tau_s)
    c = ia0 - A
    vstar = EL + R * (I + w_as * A)
    for t, (s, v, ia, ib, th, _) in enumerate(out, start=1):
        assert float(s.max()) == 0.0 and float(ib.abs().max()) == 0.0
        ia_cf = A + c * q ** t
        u_cf = r ** t * (v0 - vstar) + (1 - r) * R * w_as * c * q * (r ** t - q ** t) / (r - q)
        assert abs(float(ia[0, 0]) - ia_cf) < 1e-11
        assert abs(float(v[0, 0]) - (vstar + u_cf)) < 1e-10
    out32 = _tc_run(p, 60, I, A, dt, torch.float32, v0=v0, ia0=ia0)
    assert abs(float(out32[-1][1][0, 0]) - float(out[-1][1][0, 0])) < 1e-4


def test_tc_plateau_fixed_point():
    # with no spikes the apical current settles at the root of
    # i = A + w_Ca * sigmoid(i - theta_Ca)
    from scipy.optimize import brentq

    A, wca, thca = -0.3, 1.5, 0.4
    p = rsnn.TwoCompParams.make(v_threshold=1e6, tau_a=2.0, w_Ca=wca, theta_Ca=thca,
                                dtype=torch.float64)
    out = _tc_run(p, 400, 0.0, A, 1.0, torch.float64)
    root = brentq(lambda i: A + wca / (1 + math.exp(-(i - thca))) - i, -10, 10, xtol=1e-14)
    assert abs(float(out[-1][2][0, 0]) - root) < 1e-9
    # w_Ca = 0 removes the plateau: the current settles at A exactly
    p0 = rsnn.TwoCompParams.make(v_threshold=1e6, tau_a=2.0, w_Ca=0.0, theta_Ca=thca,
                                 dtype=torch.float64)
    assert abs(float(_tc_run(p0, 400, 0.0, A, 1.0, torch.float64)[-1][2][0, 0]) - A) < 1e-12


def test_tc_exponential_initiation_fixed_point():
    # subthreshold, the soma settles at the lower root of
    # v = E_L + R_s I + delta_T exp((v - v_th)/delta_T)   (argument below the clamp at 10)
    from scipy.optimize import brentq

    EL, R, I, dT, vth = -70.0, 1.0, 12.0, 2.0, -50.0
    p = rsnn.TwoCompParams.make(v_threshold=vth, E_L=EL, R_s=R, delta_T=dT, tau_s=5.0,
                                w_as=0.0, v_reset=-65.0, dtype=torch.float64)
    out = _tc_run(p, 600, I, 0.0, 1.0, torch.float64, v0=-70.0)
    f = lambda v: EL + R * I + dT * math.exp((v - vth) / dT) - v  # noqa: E731
    root = brentq(f, -80.0, -55.0, xtol=1e-14)
    assert all(float(o[0].max()) == 0.0 for o in out)
    assert abs(float(out[-1][1][0, 0]) - root) < 1e-9
    assert root > EL + R * I  # the exponential term raises the resting level


def test_tc_spike_reset_bap_and_threshold():
    # a spike resets v to v_reset exactly, sets i_bap = w_sa, adds delta_th to
    # the decayed threshold, and the bAP enters the next apical update exactly
    # as the same amount of apical input would (two_compartment.py:208-259)
    dt, tau_th, dth, wsa = 1.0, 20.0, 5.0, 0.7
    p = rsnn.TwoCompParams.make(v_threshold=-55.0, v_reset=-70.0, tau_s=5.0, tau_th=tau_th,
                                delta_th=dth, w_sa=wsa, tau_a=10.0, dtype=torch.float64)
    d = torch.float64
    shape = (1, 1)
    z0 = torch.zeros(shape, dtype=d)
    v0 = torch.full(shape, -70.0, dtype=d)
    s1, v1, ia1, ib1, th1, vp1 = rsnn.tc_step(v0, z0, z0, torch.full(shape, 2.0, dtype=d),
                                              torch.full(shape, 200.0, dtype=d), z0, p.to(d), dt)
    assert float(s1) == 1.0 and float(v1) == -70.0 and float(ib1) == wsa
    assert abs(float(th1) - (2.0 / (1 + dt / tau_th) + dth)) < 1e-12
    s2, *_ = rsnn.tc_step(v1, ia1, ib1, th1, z0, z0, p.to(d), dt)
    assert float(s2) == 0.0
    # bAP versus explicit apical input of the same size
    _, _, ia_a, _, _, _ = rsnn.tc_step(v1, ia1, ib1, th1, z0, z0, p.to(d), dt)
    _, _, ia_b, _, _, _ = rsnn.tc_step(v1, ia1, z0, th1, z0, ib1, p.to(d), dt)
    assert abs(float(ia_a) - float(ia_b)) < 1e-15
    # threshold decay after the spike: theta_t = theta_1 (1 + dt/tau)^-(t-1)
    th = th1
    v, ia, ib = v1, ia1, ib1
    for t in range(2, 30):
        _, v, ia, ib, th, _ = rsnn.tc_step(v, ia, ib, th, z0, z0, p.to(d), dt)
        assert abs(float(th) - float(th1) * (1 + dt / tau_th) ** -(t - 1)) < 1e-12


def test_tc_reference_test_behaviours():
    # the reference's own checks (tests/models/neurons/test_two_compartment.py:28-145)
    d, dt = torch.float32, 1.0
    p = rsnn.TwoCompParams.make(w_Ca=2.0, theta_Ca=-0.5, w_sa=0.7, delta_th=1.5)
    v = torch.full((3, 2), -65.0)
    z = torch.zeros(3, 2)
    s, v1, ia, ib, th, vp = rsnn.tc_step(v, z, z, z, torch.full((3, 2), 50.0),
                                         torch.full((3, 2), 1.5), p, dt)
    assert torch.all(ia > 0.0) and torch.all(p.v_threshold + z >= p.v_threshold)
    # exponential initiation raises the pre-spike voltage
    kw = dict(tau_s=10.0, R_s=0.2, E_L=-70.0, v_threshold=-50.0, v_reset=-65.0)
    base = rsnn.TwoCompParams.make(delta_T=0.0, **kw)
    boost = rsnn.TwoCompParams.make(delta_T=3.0, **kw)
    v = torch.full((1, 1), -51.0)
    z = torch.zeros(1, 1)
    vb = rsnn.tc_step(v, z, z, z, torch.full((1, 1), 10.0), z, base, dt)[5]
    vx = rsnn.tc_step(v, z, z, z, torch.full((1, 1), 10.0), z, boost, dt)[5]
    assert float(vx) > float(vb)
    assert d == vx.dtype


def test_tc_host_coefficients_equal_reference_expressions():
    # the host-side coefficients are the reference's own torch expressions,
    # evaluated per neuron (two_compartment.py:204-259)
    from btorch_amd import _lib
    from btorch_amd.models import TwoCompartmentGLIF
    from btorch_amd.models._plan import tc_coefficients

    n = 257
    g = torch.Generator().manual_seed(3)
    kw = dict(tau_s=torch.rand(n, generator=g) * 20 + 1e-5, tau_a=torch.rand(n, generator=g) * 90,
              tau_th=torch.rand(n, generator=g) * 400 + 100, R_s=torch.rand(n, generator=g) - 0.2,
              delta_T=torch.rand(n, generator=g) * 2 - 1.0)
    neuron = TwoCompartmentGLIF(n, **kw)
    for dt in (0.1, 1.0, 0.25):
        c = tc_coefficients(neuron, dt)
        full = lambda t: torch.broadcast_to(t, (3, n))  # noqa: E731
        da = dt / full(torch.clamp(neuron.tau_a, min=1e-4))
        ds = dt / full(torch.clamp(neuron.tau_s, min=1e-4))
        assert torch.equal(full(c[_lib.TC_DA]), da) and torch.equal(full(c[_lib.TC_DA1]), 1.0 + da)
        assert torch.equal(full(c[_lib.TC_DS]), ds) and torch.equal(full(c[_lib.TC_DS1]), 1.0 + ds)
        dth = dt / full(torch.clamp(neuron.tau_th, min=1e-4))
        assert torch.equal(full(c[_lib.TC_DTH1]), 1.0 + dth)
        dT = torch.clamp(neuron.delta_T, min=0.0)
        assert torch.equal(c[_lib.TC_DT], dT)
        assert torch.equal(c[_lib.TC_DTC], torch.clamp(dT, min=1e-4))
        assert torch.equal(c[_lib.TC_R_S], torch.clamp(neuron.R_s, min=1e-4))


def test_spikes_sparse_repr_known_answer_and_round_trip():
    # io/serialization.py:8-20: index arrays in row-major (np.nonzero) order
    z = np.zeros((2, 2, 3), np.float32)
    z[0, 1, 2] = z[1, 0, 0] = z[1, 1, 0] = z[0, 0, 1] = 1.0
    d = rsnn.spikes_sparse_repr(z)
    assert d["spikes_idx_time"].tolist() == [0, 0, 1, 1]
    assert d["spikes_idx_batch"].tolist() == [0, 1, 0, 1]
    assert d["spikes_idx_neuron"].tolist() == [1, 2, 0, 0]
    assert d["spikes"]["nnz"] == 4 and d["spikes_idx_time"].dtype == np.int32
    rng = np.random.default_rng(0)
    z = (rng.random((5, 3, 130)) < 0.1).astype(np.float32)
    d = rsnn.spikes_sparse_repr(z)
    back = np.zeros_like(z)
    back[d["spikes_idx_time"], d["spikes_idx_batch"], d["spikes_idx_neuron"]] = d["spikes_data"]
    assert np.array_equal(back, z)
    key = (d["spikes_idx_time"].astype(np.int64) * 3 + d["spikes_idx_batch"]) * 130 \
        + d["spikes_idx_neuron"]
    assert np.all(np.diff(key) > 0)  # strictly ascending flat index


# ---------------------------------------------------------------------------
# alpha and dual-exponential synapses (synapse.py:190-418)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("tau,g,dt", [(5.0, 1.0, 1.0), (2.0, 0.5, 0.1), (20.0, 3.0, 0.25)])
def test_alpha_psc_impulse_response_closed_form(tau, g, dt):
    # solving the two linear updates: k[0] = 0, k[t] = g*t*(1 - a)*a^(t-1), a = exp(-dt/tau)
    T = 80
    r = rsnn.syn_impulse_response(rsnn.AlphaSyn(torch.tensor(tau), torch.tensor(g)), T, dt)
    a = math.exp(-dt / tau)
    cf = [0.0] + [g * t * (1 - a) * a ** (t - 1) for t in range(1, T)]
    np.testing.assert_allclose(r.numpy(), cf, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("tau", [3.0, 5.0, 12.0])
def test_billeh_alpha_closed_form_and_unit_peak_at_tau(tau):
    T = 60
    r = rsnn.syn_impulse_response(rsnn.BillehSyn(torch.tensor(tau)), T, 1.0)
    c = math.exp(-1.0 / tau)
    cf = [t * (math.e / tau) * c ** t for t in range(T)]
    np.testing.assert_allclose(r.numpy(), cf, rtol=1e-12, atol=1e-14)
    # the model's normalisation: amplitude 1 at t = tau_syn (integer tau)
    assert int(torch.argmax(r)) == int(tau)
    assert abs(float(r[int(tau)]) - 1.0) < 1e-12


@pytest.mark.parametrize("td,tr,dt", [(20.0, 5.0, 1.0), (8.0, 2.0, 0.5), (3.0, 0.5, 0.1)])
def test_dual_exponential_closed_form_and_normalisation(td, tr, dt):
    T = 100
    syn = rsnn.DualExpSyn(torch.tensor(td), torch.tensor(tr))
    r = rsnn.syn_impulse_response(syn, T, dt)
    ad, ar = math.exp(-dt / td), math.exp(-dt / tr)
    a = float(syn.to(torch.float64).a)
    np.testing.assert_allclose(r.numpy(), [a * (ad ** t - ar ** t) for t in range(T)],
                               rtol=1e-12, atol=1e-14)
    # default A: the difference of exponentials peaks at 1 (continuous time,
    # t_peak = td*tr/(td - tr)*ln(td/tr)), and a = A*(td - tr)/(td*tr)
    A = a * td * tr / (td - tr)
    tp = td * tr / (td - tr) * math.log(td / tr)
    assert abs(A * (math.exp(-tp / td) - math.exp(-tp / tr)) - 1.0) < 1e-12
    grid = np.linspace(0, 10 * td, 200001)
    assert abs(A * np.max(np.exp(-grid / td) - np.exp(-grid / tr)) - 1.0) < 1e-8
    # a given A replaces the default
    a2 = float(rsnn.DualExpSyn(torch.tensor(td), torch.tensor(tr), torch.tensor(2.0))
               .to(torch.float64).a)
    assert abs(a2 - 2.0 * (td - tr) / (td * tr)) < 1e-12


@pytest.mark.parametrize("syn,dt", [
    (rsnn.AlphaSyn(torch.tensor(4.0), torch.tensor(1.5)), 0.5),
    (rsnn.BillehSyn(torch.tensor(6.0)), 1.0),
    (rsnn.DualExpSyn(torch.tensor(9.0), torch.tensor(2.0)), 0.2)])
def test_psc_step_loop_equals_convolution_with_impulse_response(syn, dt):
    # the synapse is linear and time-invariant: a step loop over any push
    # sequence equals the causal convolution with its impulse response
    # (the reference's loop-vs-conv parity, test_synapse.py:199-223)
    T = 50
    rng = np.random.default_rng(3)
    wz = torch.from_numpy(rng.normal(size=(T, 4)))
    s64 = syn.to(torch.float64)
    st = {"psc": torch.zeros(4, dtype=torch.float64)}
    st.update({k: torch.zeros(4, dtype=torch.float64) for k in s64.states})
    loop = []
    for t in range(T):
        s64.conductance(st, dt)
        s64.adaptation(st, wz[t], dt)
        loop.append(st["psc"].clone())
    loop = torch.stack(loop)
    k = rsnn.syn_impulse_response(syn, T, dt)
    conv = torch.stack([sum(k[t - j] * wz[j] for j in range(t + 1)) for t in range(T)])
    np.testing.assert_allclose(loop.numpy(), conv.numpy(), rtol=1e-10, atol=1e-12)


def test_alpha_order_mistake_is_caught():
    # the closed form separates the reference's order (psc reads h before the
    # push) from the swapped one, which would answer at step 0
    syn = rsnn.AlphaSyn(torch.tensor(5.0), torch.tensor(1.0)).to(torch.float64)
    st = {"psc": torch.zeros(1, dtype=torch.float64), "h": torch.zeros(1, dtype=torch.float64)}
    syn.adaptation(st, torch.ones(1, dtype=torch.float64), 1.0)
    syn.conductance(st, 1.0)
    assert float(st["psc"]) != 0.0
    assert float(rsnn.syn_impulse_response(syn, 2, 1.0)[0]) == 0.0


@pytest.mark.parametrize("n,nnz", [(300, 20000), (50, 20000), (3, 20000), (1, 3000)])
def test_duplicate_sum_order_equals_scipy(n, nnz):
    # the order the device builder sums duplicate edges in: scipy's
    # sum_duplicates (first value + NumPy pairwise sum of the rest), pinned
    # against scipy itself for segments of 1..3000 terms
    rng = np.random.default_rng(n)
    pre, post = rng.integers(0, n, nnz), rng.integers(0, n, nnz)
    w = rng.normal(size=nnz).astype(np.float32)
    m = scipy.sparse.coo_array((w, (pre, post)), shape=(n, n)).T.tocoo()
    m.sum_duplicates()
    groups = {}
    for p, q, x in zip(pre.tolist(), post.tolist(), w):
        groups.setdefault((q, p), []).append(x)
    want = np.array([rsnn.duplicate_sum(groups[(r, c)]) for r, c in
                     zip(m.row.tolist(), m.col.tolist())], dtype=np.float32)
    assert np.array_equal(want, m.data)
    if max(len(g) for g in groups.values()) > 1:
        # a left-to-right sum differs somewhere (the order matters)
        seq = np.array([_seq(groups[(r, c)]) for r, c in zip(m.row.tolist(), m.col.tolist())],
                       dtype=np.float32)
        assert n == 300 or not np.array_equal(seq, m.data)


def _seq(vals):
    s = np.float32(vals[0])
    for v in vals[1:]:
        s = np.float32(s + v)
    return s


def _golden_sparse_mm():
    from pathlib import Path
    return np.load(Path(__file__).with_name("golden") / "sparse_mm_order.npz")


def test_oracle_product_matches_committed_sparse_mm_fixture():
    # the committed torch-CPU sparse.mm output (tests/golden, made by
    # make_sparse_mm_order.py): the oracle's product and an ascending-source
    # sequential fp32 sum from +0.0 reproduce it bit for bit; the descending
    # order does not, so the fixture tells the orders apart
    g = _golden_sparse_mm()
    dst, src, w, x, out = g["dst"], g["src"], g["w"], g["x"], g["out"]
    n_src, n_dst = int(g["n_src"]), int(g["n_dst"])
    mm = rsnn.NativeSparse.from_coalesced(dst, src, w, n_src, n_dst)
    assert np.array_equal(mm(torch.from_numpy(x)).numpy().view(np.uint32), out.view(np.uint32))

    def seq(order):
        res = np.zeros((x.shape[0], n_dst), dtype=np.float32)
        for b in range(x.shape[0]):
            for d in range(n_dst):
                acc = np.float32(0.0)
                for e in order(np.flatnonzero(dst == d)):
                    acc = np.float32(acc + np.float32(w[e] * x[b, src[e]]))
                res[b, d] = acc
        return res
    assert np.array_equal(seq(lambda e: e), out)
    assert not np.array_equal(seq(lambda e: e[::-1]), out)


# ---------------------------------------------------------------------------
# surrogates: the reference's own invariants (tests/models/test_surrogate.py:
# 82-116) -- g(0) = damping and half width at half maximum = 1/alpha -- and
# the primitive's derivative equal to the stated derivative
# ---------------------------------------------------------------------------

SURROGATE_KINDS = ("sigmoid", "superspike", "erf", "atan")
ALPHAS = (1.0, 2.0, 4.0, 25.0 / (math.sqrt(2.0) - 1.0))


def _hwhm(g, alpha):
    """Largest u >= 0 with g(u) >= g(0)/2, by bisection on [0, 10/alpha]."""
    g0 = float(g(torch.zeros((), dtype=torch.float64)))
    lo, hi = 0.0, 10.0 / alpha
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if float(g(torch.tensor(mid, dtype=torch.float64))) >= 0.5 * g0:
            lo = mid
        else:
            hi = mid
    return lo


@pytest.mark.parametrize("kind", SURROGATE_KINDS)
@pytest.mark.parametrize("alpha", ALPHAS)
def test_oracle_surrogate_unit_peak_and_hwhm(kind, alpha):
    for damping in (1.0, 0.5):
        def g(u):
            return rsnn.surrogate_derivative(kind, u, alpha, damping)
        assert abs(float(g(torch.zeros((), dtype=torch.float64))) - damping) < 1e-12
        assert abs(_hwhm(g, alpha) - 1.0 / alpha) < 1e-9 / alpha
        # symmetric, decreasing away from the threshold
        u = torch.linspace(0.0, 5.0 / alpha, 101, dtype=torch.float64)
        assert torch.allclose(g(u), g(-u), rtol=1e-12, atol=0)
        assert bool((g(u)[1:] <= g(u)[:-1]).all())


@pytest.mark.parametrize("kind", SURROGATE_KINDS)
@pytest.mark.parametrize("alpha", ALPHAS)
def test_oracle_surrogate_primitive_integrates_to_derivative(kind, alpha):
    # the primitive (the non-spiking forward) has the derivative's shape:
    # d/du primitive(u) (autograd, float64), divided by its value at the
    # threshold, equals the unit-peak derivative (the reference normalises
    # the gradient to 1 at threshold, not the primitive's slope);
    # primitive(0) = 1/2 and it runs from 0 to 1
    u = torch.linspace(-4.0 / alpha, 4.0 / alpha, 401, dtype=torch.float64, requires_grad=True)
    p = rsnn.surrogate_primitive(kind, u, alpha)
    (gp,) = torch.autograd.grad(p.sum(), u)
    d = rsnn.surrogate_derivative(kind, u.detach(), alpha, 1.0)
    assert torch.allclose(gp / gp[200], d, rtol=1e-10, atol=1e-12)
    # its peak slope is the closed form of each primitive
    slope0 = {"sigmoid": rsnn.SIGMOID_K * alpha / 4.0,
              "superspike": 0.5 * rsnn.SUPERSPIKE_K * alpha,
              "erf": math.sqrt(math.log(2.0)) * alpha / math.sqrt(math.pi),
              "atan": alpha / math.pi}[kind]
    assert abs(float(gp[200]) - slope0) < 1e-10 * slope0
    zero = torch.zeros((), dtype=torch.float64)
    assert abs(float(rsnn.surrogate_primitive(kind, zero, alpha)) - 0.5) < 1e-12
    big = torch.tensor([-1e7 / alpha, 1e7 / alpha], dtype=torch.float64)
    ends = rsnn.surrogate_primitive(kind, big, alpha)
    assert float(ends[0]) < 1e-6 and float(ends[1]) > 1 - 1e-6


def test_oracle_surrogate_mistakes_are_caught():
    # a dropped constant (slope alpha instead of the HWHM-consistent scale)
    # moves the half width: the invariant above would fail
    alpha = 4.0
    wrong_sig = _hwhm(lambda u: 4 * torch.sigmoid(alpha * u) * (1 - torch.sigmoid(alpha * u)),
                      alpha)
    wrong_ss = _hwhm(lambda u: 1.0 / (1.0 + alpha * u.abs()) ** 2, alpha)
    wrong_erf = _hwhm(lambda u: torch.exp(-(alpha * u) ** 2), alpha)
    for hw in (wrong_sig, wrong_ss, wrong_erf):
        assert abs(hw - 1.0 / alpha) > 1e-2 / alpha


@pytest.mark.parametrize("kind", SURROGATE_KINDS)
def test_product_surrogate_classes_equal_oracle(kind):
    # the package's surrogate classes (used by the standalone steps' torch
    # backward) agree with the pinned oracle forms
    from btorch_amd.models import surrogate
    cls = {"sigmoid": surrogate.Sigmoid, "superspike": surrogate.SuperSpike,
           "erf": surrogate.Erf, "atan": surrogate.ATan}[kind]
    for alpha in ALPHAS:
        fn = cls(alpha=alpha, damping_factor=0.7)
        u = torch.linspace(-3.0 / alpha, 3.0 / alpha, 257, dtype=torch.float64)
        g = torch.ones_like(u)
        got = fn.derivative(u, g, fn.damping_factor)
        ref = rsnn.surrogate_derivative(kind, u, alpha, 0.7)
        assert torch.allclose(got, ref, rtol=1e-12, atol=0)


# ---------------------------------------------------------------------------
# step composition: hand-derived known answers of the reference's end-to-end
# test (tests/models/test_delayed_psc.py:384-450): identity weights, x = 2,
# dt = 1, LIF tau 20, threshold 1, soft reset, exponential synapse tau 2
# (a = e^-0.5 per step).  Each step: I_t = psc_{t-1} + x ; v <- v(1 - dt/tau)
# + dt*I, fire at v >= 1, v -= 1 ; psc_t = a*psc_{t-1} + z_{t-d}.
# ---------------------------------------------------------------------------

A_HALF = math.exp(-0.5)
# undelayed synapse (d = 0): psc_t = a psc_{t-1} + z_t
KNOWN_V0 = [1.0, 2.95, 4.8025 + A_HALF]
KNOWN_PSC0 = [1.0, 1.0 + A_HALF, 1.0 + A_HALF + A_HALF ** 2]
# delay of two steps (DelayedPSC(max_delay_steps=2); HeterSynapsePSC with
# every delay 2): psc_t = a psc_{t-1} + z_{t-2}
KNOWN_V2 = [1.0, 1.95, 2.8525, 4.709875, 6.47438125 + A_HALF]
KNOWN_PSC2 = [0.0, 0.0, 1.0, 1.0 + A_HALF, 1.0 + A_HALF + A_HALF ** 2]


def _eye_net(n, mode, D):
    eye = scipy.sparse.coo_array(np.eye(n, dtype=np.float32))
    if mode == "hetero":
        # every synapse delayed by D - 1 steps: row pre*D + (D - 1)
        conn = rsnn.expand_conn_for_delays(eye, np.full(n, D - 1), D)
        kw = dict(delay_mode="hetero", max_delay_steps=D, circular=True)
    elif mode == "fixed":
        conn, kw = eye, dict(delay_mode="fixed", max_delay_steps=D, circular=True)
    else:
        conn, kw = eye, {}
    lif = rsnn.LIFParams.make(n, tau=20.0)
    return rsnn.Net(n, conn, lif, torch.tensor(2.0), 1.0, **kw)


@pytest.mark.parametrize("mode,D,v_ref,psc_ref", [
    ("none", 1, KNOWN_V0, KNOWN_PSC0),
    ("fixed", 2, KNOWN_V2, KNOWN_PSC2),
    ("hetero", 3, KNOWN_V2, KNOWN_PSC2),
])
def test_step_composition_known_answers(mode, D, v_ref, psc_ref):
    n, B = 4, 2
    T = len(v_ref)
    x = torch.full((T, B, n), 2.0, dtype=torch.float64)
    out = rsnn.simulate(_eye_net(n, mode, D), x, dtype=torch.float64)
    assert bool((out["z"] == 1).all())
    v = out["v"][:, 0, 0].tolist()
    psc = out["psc"][:, 0, 0].tolist()
    assert np.allclose(v, v_ref, rtol=0, atol=1e-12), (v, v_ref)
    assert np.allclose(psc, psc_ref, rtol=0, atol=1e-12), (psc, psc_ref)
    assert torch.equal(out["v"], out["v"][:, :1, :1].expand_as(out["v"]))


def test_step_composition_mistakes_fail_the_known_answers():
    # plausible misreadings of rnn.py:517-528 give other numbers: the neuron
    # reading psc after this step's push, or the push delivering one step late
    v_ref, psc_ref = KNOWN_V0, KNOWN_PSC0
    a = A_HALF
    # neuron reads the freshly pushed psc: I_0 = 1 + 2
    v_wrong = 0.0 * 0.95 + 3.0 - 1.0
    assert abs(v_wrong - v_ref[0]) > 0.5
    # push before decay: psc_1 = a*(1 + 1) instead of a + 1
    assert abs(a * 2.0 - psc_ref[1]) > 0.1
    # off-by-one delay: the d = 2 trace would equal the d = 3 one
    assert KNOWN_PSC2[2] != 0.0


# ---------------------------------------------------------------------------
# mixed populations: a population of groups equals its groups run apart
# when no synapse crosses groups, and a population of identical LIF groups
# equals one LIF population
# ---------------------------------------------------------------------------

def _block_conn(sizes, p, seed):
    rng = np.random.default_rng(seed)
    n = sum(sizes)
    dense = np.zeros((n, n), dtype=np.float32)
    s0 = 0
    for c in sizes:
        blk = (rng.random((c, c)) < p) * rng.lognormal(-1, 0.5, (c, c))
        dense[s0:s0 + c, s0:s0 + c] = blk * rng.choice([-1.0, 1.0], size=(c, c))
        s0 += c
    return scipy.sparse.coo_array(dense)


def test_mixed_population_equals_its_groups_apart():
    sizes = (37, 29, 18)
    n, B, T = sum(sizes), 2, 60
    conn = _block_conn(sizes, 0.2, 3)
    glif = rsnn.GLIF3Params.make(sizes[0], tau=torch.linspace(10, 30, sizes[0]), v_threshold=1.0,
                                 v_reset=0.0, c_m=1.0, tau_ref=2.0, k=(0.2, 0.05),
                                 asc_amps=(-0.3, 0.1))
    tc = rsnn.TwoCompParams.make(v_threshold=1.0, v_reset=0.0, tau_s=15.0, E_L=0.0, tau_a=20.0,
                                 tau_th=50.0, delta_th=0.2)
    lif = rsnn.LIFParams.make(sizes[2], tau=12.0, tau_ref=None, hard_reset=True)
    mixed = rsnn.MixedParams([(0, sizes[0], glif), (sizes[0], sizes[1], tc),
                              (sizes[0] + sizes[1], sizes[2], lif)])
    rng = np.random.default_rng(4)
    x = torch.from_numpy((rng.random((T, B, n)) < 0.1).astype(np.float32) * 2.5)
    xa = torch.from_numpy(rng.uniform(0, 0.5, (T, B, n)).astype(np.float32))
    out = rsnn.simulate(rsnn.Net(n, conn, mixed, torch.tensor(5.0), 0.5), x, x_apical=xa)
    assert out["z"].sum() > 50
    s0 = 0
    for (c, prm) in zip(sizes, (glif, tc, lif)):
        sl = slice(s0, s0 + c)
        sub = conn.tocsr()[s0:s0 + c, s0:s0 + c].tocoo()
        ref = rsnn.simulate(rsnn.Net(c, sub, prm, torch.tensor(5.0), 0.5), x[:, :, sl],
                            x_apical=xa[:, :, sl] if prm is tc else None)
        for k in ("z", "v", "psc"):
            assert torch.equal(out[k][:, :, sl], ref[k]), (k, s0)
        s0 += c


def test_mixed_population_of_identical_lif_groups_equals_one_population():
    n, B, T = 90, 3, 80
    conn = _rand_conn(n, n, 0.1, 5)
    prm = rsnn.LIFParams.make(n, tau=20.0, tau_ref=2.0)
    mixed = rsnn.MixedParams([(0, 40, rsnn.LIFParams.make(40, tau=20.0, tau_ref=2.0)),
                              (40, 50, rsnn.LIFParams.make(50, tau=20.0, tau_ref=2.0))])
    rng = np.random.default_rng(6)
    x = torch.from_numpy((rng.random((T, B, n)) < 0.1).astype(np.float32) * 15.0)
    a = rsnn.simulate(rsnn.Net(n, conn, mixed, torch.tensor(5.0), 0.1), x)
    b = rsnn.simulate(rsnn.Net(n, conn, prm, torch.tensor(5.0), 0.1), x)
    assert a["z"].sum() > 50
    for k in ("z", "v", "refractory", "psc"):
        assert torch.equal(a[k], b[k]), k
