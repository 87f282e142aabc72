"""GPU parity of the HIP path against the CPU oracle (bit-exact for spikes,
indices and the ordered-mode currents/membranes; tolerances stated where the
order-free mode is used)."""

import numpy as np
import pytest
import scipy.sparse
import torch

from btorch_amd import _kernels, workloads
from btorch_amd.connectome import expand_conn_for_delays
from btorch_amd.models import (
    LIF, DelayedPSC, ExponentialPSC, HeterSynapsePSC, RecurrentNN, SparseConn, environ,
    functional,
)
from btorch_amd.models._engine import pack_bits, unpack_bits
from oracle import rsnn

pytestmark = pytest.mark.gpu
DEV = "cuda"


def bits_equal(a: torch.Tensor, b: torch.Tensor) -> bool:
    """Bitwise equality of float32 tensors, treating -0.0 == +0.0."""
    a = a.detach().cpu().float()
    b = b.detach().cpu().float()
    return bool(torch.equal(a, b)) or bool(((a == b) | (a.isnan() & b.isnan())).all())


def _conn(n_src, n_dst, p, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    dense = (rng.random((n_src, n_dst)) < p) * rng.lognormal(-1, 0.5, (n_src, n_dst)) * scale
    dense *= rng.choice([-1.0, 1.0], size=dense.shape)
    return scipy.sparse.coo_array(dense.astype(np.float32))


# ---------------------------------------------------------------------------
# compaction
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("lanes,n", [(1, 1), (3, 63), (4, 64), (5, 1000), (2, 70001)])
@pytest.mark.parametrize("density", [0.0, 0.01, 0.5, 1.0])
def test_compaction_dense_and_bits_bit_exact(lanes, n, density):
    rng = np.random.default_rng(lanes * 7 + n)
    x = (rng.random((lanes, n)) < density).astype(np.float32)
    xt = torch.from_numpy(x).to(DEV)
    idx, cnt = _kernels.compact_dense(xt)
    words = pack_bits(xt)
    idx2, cnt2 = _kernels.compact_bits(words.contiguous(), n)
    torch.cuda.synchronize()
    for b in range(lanes):
        ref = np.flatnonzero(x[b])
        c = int(cnt[b])
        assert c == ref.size and int(cnt2[b]) == ref.size
        assert np.array_equal(idx[b, :c].cpu().numpy(), ref)
        assert np.array_equal(idx2[b, :c].cpu().numpy(), ref)


@pytest.mark.parametrize("rows,n,offset,pad", [(3, 1, 0, 0), (5, 4093, 0, 0), (7, 4099, 1, 0),
                                               (3, 8196, 0, 0), (4, 70, 0, 2), (70001, 67, 0, 0),
                                               (66000, 68, 0, 0)])
def test_pack_unpack_rows_round_trip(rows, n, offset, pad):
    """Bits of each row against numpy's packing; unpack through the 16-byte
    store path (16-byte aligned rows, with a partial last group when rows are
    padded), the per-float path (rows one float off alignment, or a stride
    not a multiple of 4) and the row chunks above 65535 rows."""
    rng = np.random.default_rng(rows + n)
    x = (rng.random((rows, n)) < 0.3).astype(np.float32)
    words = _kernels.pack_rows(torch.from_numpy(x).to(DEV))
    ref = np.packbits(x.astype(bool), axis=1, bitorder="little")
    ref = np.pad(ref, ((0, 0), (0, words.shape[1] * 8 - ref.shape[1])))
    assert np.array_equal(words.cpu().numpy().view(np.uint8).reshape(rows, -1), ref)
    buf = torch.full((rows * (n + pad) + offset,), 7.0, device=DEV)
    out = buf[offset:].view(rows, n + pad)[:, :n]
    _kernels.unpack_rows(words, n, out=out)
    assert torch.equal(out.cpu(), torch.from_numpy(x))
    rest = buf[offset:].view(rows, n + pad)[:, n:]
    assert bool((rest == 7.0).all()) and (offset == 0 or float(buf[0]) == 7.0)


# ---------------------------------------------------------------------------
# event-driven product
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("lanes", [1, 8, 33])
@pytest.mark.parametrize("values", ["spikes", "real"])
def test_sparse_conn_ordered_bit_exact(lanes, values):
    conn = _conn(300, 200, 0.3, 0)
    rng = np.random.default_rng(lanes)
    x = (rng.random((lanes, 300)) < 0.3).astype(np.float32)
    if values == "real":
        x = x * (rng.random((lanes, 300)) * 3 - 1).astype(np.float32)
    ref = rsnn.NativeSparse(conn)(torch.from_numpy(x))
    m = SparseConn(conn, sparse_backend="native").to(DEV)
    with torch.no_grad():
        out = m(torch.from_numpy(x).to(DEV))
    assert bits_equal(out, ref)
    # a 1-D input returns a 1-D output
    with torch.no_grad():
        o1 = m(torch.from_numpy(x[0]).to(DEV))
    assert o1.shape == (200,) and bits_equal(o1, ref[0])


def test_sparse_conn_many_terms_per_destination_overflow_path():
    # dense columns force > slots terms per destination (pulled reduction)
    conn = _conn(500, 16, 0.9, 4)
    x = np.ones((3, 500), np.float32)
    ref = rsnn.NativeSparse(conn)(torch.from_numpy(x))
    m = SparseConn(conn, sparse_config={"slots": 4}).to(DEV)
    with torch.no_grad():
        out = m(torch.from_numpy(x).to(DEV))
    assert bits_equal(out, ref)


def test_sparse_conn_zero_input_and_fast_mode():
    conn = _conn(256, 128, 0.2, 1)
    m = SparseConn(conn, sparse_backend="triton").to(DEV)
    with torch.no_grad():
        z = m(torch.zeros(4, 256, device=DEV))
        assert torch.count_nonzero(z) == 0
        x = (torch.rand(4, 256, device=DEV) < 0.2).float()
        out = m(x)
    ref = rsnn.NativeSparse(conn)(x.cpu())
    # order-free float atomics: tolerance, not bits
    torch.testing.assert_close(out.cpu(), ref, rtol=1e-5, atol=1e-6)


def test_sparse_conn_backward_matches_float64():
    conn = _conn(150, 90, 0.2, 2)
    m = SparseConn(conn).to(DEV)
    x = ((torch.rand(5, 150) < 0.4).float() * torch.rand(5, 150)).to(DEV).requires_grad_(True)
    g = torch.randn(5, 90, device=DEV)
    out = m(x)
    out.backward(g)
    W = torch.from_numpy(conn.toarray()).double()
    gx_ref = g.cpu().double() @ W.T
    torch.testing.assert_close(x.grad.cpu().double(), gx_ref, rtol=1e-5, atol=1e-6)
    dst, src = m.indices[0].cpu(), m.indices[1].cpu()
    gw_ref = (x.detach().cpu().double()[:, src] * g.cpu().double()[:, dst]).sum(0)
    torch.testing.assert_close(m.magnitude.grad.cpu().double(), gw_ref, rtol=1e-5, atol=1e-6)


# ---------------------------------------------------------------------------
# neuron step
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("hard", [False, True])
def test_lif_single_step_bit_exact(hard):
    n, B = 1000, 3
    tau = torch.linspace(10, 30, n)
    lif = LIF(n, tau=tau, v_reset=-0.2, tau_ref=2.0, hard_reset=hard).to(DEV)
    lif.init_state(batch_size=B, device=DEV, dtype=torch.float32)
    p = rsnn.LIFParams.make(n, tau=tau, v_reset=-0.2, tau_ref=2.0, hard_reset=hard)
    v = torch.full((B, n), -0.2)
    ref = torch.zeros(B, n)
    g = torch.Generator().manual_seed(0)
    with environ.context(dt=0.1), torch.no_grad():
        for _ in range(60):
            x = torch.rand(B, n, generator=g) * 25
            z = lif(x.to(DEV))
            zo, v, ref, _ = rsnn.lif_step(v, ref, x, p, 0.1)
            assert bits_equal(z, zo)
            assert bits_equal(lif.v, v) and bits_equal(lif.refractory, ref)


# ---------------------------------------------------------------------------
# fused engine vs oracle
# ---------------------------------------------------------------------------

def _plain_model(n, p, seed, tau_ref=2.0, hard=False, tau=20.0, v_reset=0.0, backend="hip",
                 n_in=None):
    c = workloads.plain_lif_net(n=n, p=p, seed=seed)
    conn = c.matrix()
    inp = None
    if n_in:
        inp = SparseConn(workloads.input_weights(n_in, n, 0.05, seed), enforce_dale=False)
    model = RecurrentNN(LIF(n, tau=tau, tau_ref=tau_ref, hard_reset=hard, v_reset=v_reset),
                        ExponentialPSC(n, 5.0, SparseConn(conn, sparse_backend=backend)),
                        neuron_inp_module=inp, step_mode="m")
    return model.to(DEV), conn, inp


def _net(n, conn, tau=20.0, tau_ref=2.0, hard=False, v_reset=0.0, **kw):
    return rsnn.Net(n, conn, rsnn.LIFParams.make(n, tau=tau, tau_ref=tau_ref, hard_reset=hard,
                                                 v_reset=v_reset), torch.tensor(5.0), 0.1, **kw)


def test_engine_plain_with_input_projection_bit_exact():
    n, B, T, n_in = 512, 2, 300, 100
    model, conn, inp = _plain_model(n, 0.05, 11, n_in=n_in)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    xin = torch.from_numpy(workloads.poisson_input(T, B, n_in, 400.0, 0.1, seed=11) * 30.0)
    with environ.context(dt=0.1), torch.no_grad():
        z, st = model(xin.to(DEV))
    x_proj = rsnn.NativeSparse(workloads.input_weights(n_in, n, 0.05, 11))(
        xin.reshape(T * B, n_in)).reshape(T, B, n)
    ref = rsnn.simulate(_net(n, conn), x_proj)
    assert ref["z"].sum() > 50, "test network too quiet"
    assert bits_equal(z, ref["z"])
    assert bits_equal(st["neuron.v"], ref["v"])
    assert bits_equal(st["neuron.refractory"], ref["refractory"])
    assert bits_equal(st["synapse.psc"], ref["psc"])
    assert bits_equal(model.synapse.psc, ref["final"]["psc"])


@pytest.mark.parametrize("hard,n", [(False, 700), (True, 700), (False, 703)])
def test_engine_heterogeneous_params_bit_exact(hard, n):
    # n = 703 runs the scalar step kernel, n = 700 the 4-wide one
    B, T = 3, 250
    tau = torch.from_numpy(np.random.default_rng(5).uniform(10, 30, n).astype(np.float32))
    model, conn, _ = _plain_model(n, 0.04, 5, hard=hard, tau=tau, v_reset=-0.1)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 300.0, 0.1, seed=5) * 25.0)
    model.update_state_names = ["neuron.v", "synapse.psc"]
    with environ.context(dt=0.1), torch.no_grad():
        z, st = model(x.to(DEV))
    ref = rsnn.simulate(_net(n, conn, tau=tau, hard=hard, v_reset=-0.1), x)
    assert set(st) == {"neuron.v", "synapse.psc"}
    assert bits_equal(z, ref["z"]) and bits_equal(st["neuron.v"], ref["v"])
    assert bits_equal(st["synapse.psc"], ref["psc"])


def _delayed_net(n, D, seed, k_max=60):
    c = workloads.powerlaw_net_fast(n=n, k_min=5, k_max=k_max, delay_lo=1, delay_hi=D - 1,
                                    seed=seed)
    return c, c.expanded(D)


@pytest.mark.parametrize("circular", [True, False])
def test_engine_hetero_delays_bit_exact(circular):
    n, B, T, D = 600, 3, 200, 7
    c, cexp = _delayed_net(n, D, 7)
    syn = HeterSynapsePSC(n, 1, None, SparseConn(cexp), base_psc=ExponentialPSC,
                          max_delay_steps=D, use_circular_buffer=circular, tau_syn=5.0)
    model = RecurrentNN(LIF(n, tau_ref=2.0), syn, step_mode="m").to(DEV)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 300.0, 0.1, seed=7) * 25.0)
    model.update_state_names = ["neuron.v", "synapse.psc", "synapse.history.history"]
    with environ.context(dt=0.1), torch.no_grad():
        z, st = model(x.to(DEV))
    net = _net(n, cexp, delay_mode="hetero", max_delay_steps=D, circular=circular)
    ref = rsnn.simulate(net, x)
    assert ref["z"].sum() > 50
    assert bits_equal(z, ref["z"])
    assert bits_equal(st["neuron.v"], ref["v"]) and bits_equal(st["synapse.psc"], ref["psc"])
    assert bits_equal(model.synapse.history.history, ref["final"]["history"])
    if circular:
        assert model.synapse.history.cursor == ref["final"]["cursor"]
    # a second call continues from the carried state (ring rebuilt from history)
    x2 = torch.from_numpy(workloads.poisson_input(50, B, n, 300.0, 0.1, seed=8) * 25.0)
    with environ.context(dt=0.1), torch.no_grad():
        z2, _ = model(x2.to(DEV))
    st0 = dict(ref["final"])
    ref2 = rsnn.simulate(net, x2, state=st0)
    assert bits_equal(z2, ref2["z"])


def test_engine_recorded_history_sequence():
    n, B, T, D = 64, 2, 12, 4
    c, cexp = _delayed_net(n, D, 9, k_max=20)
    for circular in (True, False):
        syn = HeterSynapsePSC(n, 1, None, SparseConn(cexp), base_psc=ExponentialPSC,
                              max_delay_steps=D, use_circular_buffer=circular, tau_syn=5.0)
        model = RecurrentNN(LIF(n), syn, step_mode="m",
                            update_state_names=["synapse.history.history"]).to(DEV)
        functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
        x = torch.from_numpy(workloads.poisson_input(T, B, n, 3000.0, 0.1, seed=9) * 25.0)
        with environ.context(dt=0.1), torch.no_grad():
            z, st = model(x.to(DEV))
        ring = rsnn.Ring(B, n, D, circular)
        for t in range(T):
            ring.update(z[t].cpu())
            assert bits_equal(st["synapse.history.history"][t], ring.h)


def test_engine_delayed_psc_bit_exact():
    n, B, T, m = 300, 2, 150, 3
    c = workloads.plain_lif_net(n=n, p=0.05, seed=13)
    syn = DelayedPSC(ExponentialPSC(n, 5.0, SparseConn(c.matrix())), max_delay_steps=m)
    model = RecurrentNN(LIF(n, tau_ref=1.0), syn, step_mode="m",
                        update_state_names=["neuron.v", "synapse.psc_module.psc"]).to(DEV)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 400.0, 0.1, seed=13) * 25.0)
    with environ.context(dt=0.1), torch.no_grad():
        z, st = model(x.to(DEV))
    ref = rsnn.simulate(_net(n, c.matrix(), tau_ref=1.0, delay_mode="fixed", max_delay_steps=m),
                        x)
    assert bits_equal(z, ref["z"]) and bits_equal(st["synapse.psc_module.psc"], ref["psc"])


def test_engine_teacher_forced_rasters_and_fast_mode():
    n, B, T, D = 800, 2, 200, 9
    c, cexp = _delayed_net(n, D, 21)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 300.0, 0.1, seed=21) * 25.0)
    net = _net(n, cexp, delay_mode="hetero", max_delay_steps=D, circular=True)
    free = rsnn.simulate(net, x)
    teacher = free["z"]
    for backend in ("hip", "triton"):
        syn = HeterSynapsePSC(n, 1, None, SparseConn(cexp, sparse_backend=backend),
                              base_psc=ExponentialPSC, max_delay_steps=D,
                              use_circular_buffer=True, tau_syn=5.0)
        model = RecurrentNN(LIF(n, tau_ref=2.0), syn, step_mode="m").to(DEV)
        functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
        eng = model.fused_engine()
        words = pack_bits(teacher.to(DEV)).contiguous()
        with environ.context(dt=0.1), torch.no_grad():
            model.neuron.v  # state exists
            xs = x.to(DEV).contiguous()
            raster, recs = eng.run(xs, T, B, {"psc", "v"}, teacher=words)
        z = unpack_bits(raster, n).cpu()
        if backend == "hip":
            assert bits_equal(z, teacher) and bits_equal(recs["psc"], free["psc"])
        else:
            # order-free sums: currents within 1e-5 relative of the float32
            # oracle under teacher forcing; rasters agree except where v sits
            # within rounding of threshold
            psc, ref = recs["psc"].cpu(), free["psc"]
            scale = ref.abs().amax().clamp_min(1e-6)
            assert ((psc - ref).abs().max() / scale).item() < 1e-5
            assert (z == teacher).float().mean().item() > 0.999


def test_engine_float64_error_statement():
    # float32 HIP vs float64 oracle: report the gap (membranes within 1e-4 abs
    # while rasters agree) -- the bit-exact target is the float32 oracle.
    n, B, T = 400, 2, 100
    model, conn, _ = _plain_model(n, 0.05, 17)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 300.0, 0.1, seed=17) * 25.0)
    with environ.context(dt=0.1), torch.no_grad():
        z, st = model(x.to(DEV))
    r64 = rsnn.simulate(_net(n, conn), x, dtype=torch.float64, teacher=z.cpu().double())
    same = (r64["z"].float() == z.cpu())
    v_err = (st["neuron.v"].cpu().double() - r64["v"]).abs()[same.bool()].max().item()
    print(f"\nfloat32 vs float64 oracle (teacher forced): max |dv| = {v_err:.3e}")
    assert v_err < 1e-4


@pytest.mark.parametrize("circular", [True, False])
@pytest.mark.parametrize("bin_cap", [0, 96])
def test_engine_pull_path_bursts_bit_exact(circular, bin_cap):
    # slots=2 with a dense, bursty delayed net: most destinations receive > 2
    # terms per step and are reduced by the exact pull over incoming edges,
    # including pulls that read the spikes of steps before the run; a small
    # bin capacity makes whole tiles overflow and pull.
    n, B, T, D = 300, 2, 80, 5
    c = workloads.powerlaw_net_fast(n=n, k_min=40, k_max=150, delay_lo=1, delay_hi=D - 1,
                                    seed=31)
    cexp = c.expanded(D)
    syn = HeterSynapsePSC(n, 1, None, SparseConn(cexp, sparse_config={"slots": 2,
                                                                      "bin_cap": bin_cap}),
                          base_psc=ExponentialPSC, max_delay_steps=D,
                          use_circular_buffer=circular, tau_syn=5.0)
    model = RecurrentNN(LIF(n, tau_ref=0.5), syn, step_mode="m",
                        update_state_names=["synapse.psc"]).to(DEV)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 2000.0, 0.1, seed=31) * 25.0)
    net = _net(n, cexp, tau_ref=0.5, delay_mode="hetero", max_delay_steps=D, circular=circular)
    ref = rsnn.simulate(net, x)
    with environ.context(dt=0.1), torch.no_grad():
        z1, s1 = model(x[:37].to(DEV))
        z2, s2 = model(x[37:].to(DEV))
    assert ref["z"].sum() > 500
    assert bits_equal(torch.cat([z1, z2]), ref["z"])
    assert bits_equal(torch.cat([s1["synapse.psc"], s2["synapse.psc"]]), ref["psc"])


# ---------------------------------------------------------------------------
# GLIF3
# ---------------------------------------------------------------------------

def _glif_kwargs(n, seed):
    rng = np.random.default_rng(seed)
    tau = torch.from_numpy(rng.uniform(10.0, 30.0, n).astype(np.float32))
    return dict(v_threshold=-50.0, v_reset=-70.0, c_m=0.05, tau=tau, k=[0.3, 0.02],
                asc_amps=[-0.6, 0.9], tau_ref=2.0)


def test_glif3_single_step_bit_exact():
    from btorch_amd.models import GLIF3

    n, B = 1000, 3
    kw = _glif_kwargs(n, 3)
    neuron = GLIF3(n, **kw).to(DEV)
    functional.init_net_state(neuron, batch_size=B, device=DEV, dtype=torch.float32)
    p = rsnn.GLIF3Params.make(n, **kw)
    v = torch.full((B, n), -70.0)
    ia = torch.zeros(B, n, 2)
    ref = torch.zeros(B, n)
    g = torch.Generator().manual_seed(4)
    for t in range(40):
        x = (torch.rand(B, n, generator=g) < 0.2).float() * 6.0
        with environ.context(dt=0.1), torch.no_grad():
            z = neuron(x.to(DEV))
        zr, v, ia, ref, _ = rsnn.glif3_step(v, ia, ref, x, p, 0.1)
        assert bits_equal(z, zr), t
        assert bits_equal(neuron.v, v) and bits_equal(neuron.Iasc, ia)
        assert bits_equal(neuron.refractory, ref)
    assert float(zr.sum()) >= 0


@pytest.mark.parametrize("n", [600, 703])
def test_engine_glif3_hetero_delays_bit_exact(n):
    from btorch_amd.models import GLIF3

    B, T, D = 2, 150, 5
    c, cexp = _delayed_net(n, D, 17)
    kw = _glif_kwargs(n, 5)
    syn = HeterSynapsePSC(n, 1, None, SparseConn(cexp), base_psc=ExponentialPSC,
                          max_delay_steps=D, use_circular_buffer=True, tau_syn=5.0)
    model = RecurrentNN(GLIF3(n, **kw), syn, step_mode="m",
                        update_state_names=["neuron.v", "synapse.psc"]).to(DEV)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 400.0, 0.1, seed=5) * 8.0)
    with environ.context(dt=0.1), torch.no_grad():
        z, st = model(x.to(DEV))
    net = rsnn.Net(n, cexp, rsnn.GLIF3Params.make(n, **kw), torch.tensor(5.0), 0.1,
                   delay_mode="hetero", max_delay_steps=D, circular=True)
    ref = rsnn.simulate(net, x)
    assert ref["z"].sum() > 100
    assert bits_equal(z, ref["z"])
    assert bits_equal(st["neuron.v"], ref["v"]) and bits_equal(st["synapse.psc"], ref["psc"])
    assert bits_equal(model.neuron.Iasc, ref["final"]["iasc"])


# ---------------------------------------------------------------------------
# Two-compartment neuron
# ---------------------------------------------------------------------------

def _tc_kwargs(n, seed, nonlinear=False):
    g = torch.Generator().manual_seed(seed)
    kw = dict(v_threshold=-50.0, v_reset=-65.0, E_L=-70.0,
              tau_s=torch.rand(n, generator=g) * 20 + 10,      # U(10, 30)
              tau_th=torch.rand(n, generator=g) * 400 + 100,   # U(100, 500)
              delta_th=2.0, tau_a=50.0, w_sa=0.5, w_as=1.0, R_s=1.0)
    if nonlinear:
        kw.update(w_Ca=torch.rand(n, generator=g) * 2.0, theta_Ca=0.5,
                  delta_T=torch.rand(n, generator=g) * 2.0)
    return kw


def _tc_state(neuron):
    return [t.detach().cpu().clone() for t in (neuron.v, neuron.i_a, neuron.i_bap,
                                                neuron.theta_th)]


@pytest.mark.parametrize("nonlinear", [False, True])
@pytest.mark.parametrize("n", [1000, 1001])
def test_two_compartment_single_step_matches_oracle(n, nonlinear):
    # Linear coefficients (w_Ca = 0, delta_T = 0): bit for bit over 60 steps.
    # With the calcium plateau and exponential initiation the device's exp
    # and sigmoid differ from torch's by rounding: every step is compared from
    # the same entering state, states within 1e-5 relative, spikes equal
    # wherever the membrane is not within 1e-4 of the threshold.
    from btorch_amd.models import TwoCompartmentGLIF

    B, dt = 3, 1.0
    kw = _tc_kwargs(n, 11, nonlinear)
    neuron = TwoCompartmentGLIF(n, **kw).to(DEV)
    functional.init_net_state(neuron, batch_size=B, device=DEV, dtype=torch.float32)
    p = rsnn.TwoCompParams.make(**kw)
    g = torch.Generator().manual_seed(4)
    v, ia, ib, th = _tc_state(neuron)
    spikes = 0
    for t in range(60):
        soma = 19.0 + (torch.rand(B, n, generator=g) < 0.2).float() * 40.0
        apical = torch.randn(B, n, generator=g)
        if nonlinear:
            v, ia, ib, th = _tc_state(neuron)
        with environ.context(dt=dt), torch.no_grad():
            z, vd, st = neuron(soma.to(DEV), apical.to(DEV), return_state=True)
        zr, v, ia, ib, th, vp = rsnn.tc_step(v, ia, ib, th, soma, apical, p, dt)
        spikes += int(zr.sum())
        if not nonlinear:
            assert bits_equal(z, zr), t
            for a, b in ((neuron.v, v), (neuron.i_a, ia), (neuron.i_bap, ib),
                         (neuron.theta_th, th), (st["v_pre_spike"], vp)):
                assert bits_equal(a, b), t
        else:
            thr = st["v_threshold_eff"].cpu()
            clear = (vp - thr).abs() > 1e-4 * thr.abs()
            assert torch.equal(z.cpu()[clear], zr[clear]), t
            same = (z.cpu() == zr)
            for a, b in ((neuron.v, v), (neuron.i_a, ia), (neuron.theta_th, th),
                         (st["v_pre_spike"], vp)):
                torch.testing.assert_close(a.cpu()[same], b[same], rtol=1e-5, atol=1e-5)
        assert torch.equal(vd, neuron.v)
    assert spikes > 100


def test_two_compartment_module_behaviours():
    # the reference's checks (tests/models/neurons/test_two_compartment.py:28-145)
    # on the HIP step
    from btorch_amd.models import TwoCompartmentGLIF

    neuron = TwoCompartmentGLIF(2, w_Ca=2.0, theta_Ca=-0.5, w_sa=0.7, delta_th=1.5).to(DEV)
    functional.init_net_state(neuron, batch_size=3, device=DEV, dtype=torch.float32)
    with environ.context(dt=1.0), torch.no_grad():
        spike, voltage, state = neuron.single_step_forward(
            torch.full((3, 2), 50.0, device=DEV), torch.full((3, 2), 1.5, device=DEV),
            return_state=True)
    assert spike.shape == voltage.shape == state["i_a"].shape == (3, 2)
    assert torch.all(state["i_a"] > 0.0) and torch.equal(neuron.i_a, state["i_a"])
    assert torch.all(state["v_threshold_eff"] >= neuron.v_threshold)

    neuron = TwoCompartmentGLIF(1, tau_s=5.0, R_s=1.0, E_L=-70.0, v_threshold=-55.0,
                                v_reset=-70.0, tau_th=20.0, delta_th=5.0).to(DEV)
    functional.init_net_state(neuron, batch_size=1, device=DEV, dtype=torch.float32)
    with environ.context(dt=1.0), torch.no_grad():
        s1, _, st1 = neuron.single_step_forward(torch.full((1, 1), 200.0, device=DEV),
                                                torch.zeros(1, 1, device=DEV), return_state=True)
        s2, _, st2 = neuron.single_step_forward(torch.zeros(1, 1, device=DEV),
                                                torch.zeros(1, 1, device=DEV), return_state=True)
    assert float(s1) > 0.0 and float(st1["theta_th"]) > 0.0
    assert float(st2["v_threshold_eff"]) > float(neuron.v_threshold) and float(s2) == 0.0

    # a rollout repeats after a reset, and the returned tensors keep their values
    neuron = TwoCompartmentGLIF(1, step_mode="m").to(DEV)
    g = torch.Generator().manual_seed(4)
    i_soma, i_apical = torch.randn(12, 2, 1, generator=g), torch.randn(12, 2, 1, generator=g)
    functional.init_net_state(neuron, batch_size=2, device=DEV, dtype=torch.float32)
    with environ.context(dt=1.0), torch.no_grad():
        s_a, v_a = neuron(i_soma.to(DEV), i_apical.to(DEV))
        s_b, v_b = neuron(i_soma.to(DEV), i_apical.to(DEV))
        functional.reset_net(neuron, batch_size=2)
        s_c, v_c = neuron(i_soma.to(DEV), i_apical.to(DEV))
    assert not torch.allclose(v_a, v_b)
    assert torch.equal(s_a, s_c) and torch.equal(v_a, v_c)


def _tc_net(n, D, seed, scale=40.0):
    c, cexp = _delayed_net(n, D, seed)
    return scipy.sparse.coo_array(cexp * scale).astype(np.float32)


@pytest.mark.parametrize("n", [600, 703])
def test_engine_two_compartment_apical_bit_exact(n):
    # ApicalRecurrentNN over a one-group two-compartment population with
    # delayed synapses and an apical input sequence runs in the fused engine,
    # bit for bit with the oracle's ApicalRecurrentNN loop
    from btorch_amd.models import (ApicalRecurrentNN, MixedNeuronPopulation,
                                   TwoCompartmentGLIF)

    B, T, D = 2, 150, 5
    cexp = _tc_net(n, D, 19)
    kw = _tc_kwargs(n, 7)
    syn = HeterSynapsePSC(n, 1, None, SparseConn(cexp), base_psc=ExponentialPSC,
                          max_delay_steps=D, use_circular_buffer=True, tau_syn=5.0)
    tc = TwoCompartmentGLIF(n, **kw)
    model = ApicalRecurrentNN(MixedNeuronPopulation([(n, tc)]), syn, step_mode="m").to(DEV)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    x = (torch.from_numpy(workloads.poisson_input(T, B, n, 100.0, 1.0, seed=5) * 60.0)
         + 18.0).float()
    xa = torch.randn(T, B, n, generator=torch.Generator().manual_seed(2))
    assert model.fused_engine() is not None
    with environ.context(dt=1.0), torch.no_grad():
        z, st = model(x.to(DEV), None, xa.to(DEV))
    net = rsnn.Net(n, cexp, rsnn.TwoCompParams.make(**kw), torch.tensor(5.0), 1.0,
                   delay_mode="hetero", max_delay_steps=D, circular=True)
    ref = rsnn.simulate(net, x, x_apical=xa)
    assert ref["z"].sum() > 1000
    assert bits_equal(z, ref["z"])
    for key, rk in (("neuron.group_0.v", "v"), ("neuron.group_0.i_a", "i_a"),
                    ("neuron.group_0.theta_th", "theta_th"), ("synapse.psc", "psc")):
        assert bits_equal(st[key], ref[rk]), key
    fin = ref["final"]
    for a, b in ((tc.v, fin["v"]), (tc.i_a, fin["i_a"]), (tc.i_bap, fin["i_bap"]),
                 (tc.theta_th, fin["theta_th"]), (syn.psc, fin["psc"])):
        assert bits_equal(a, b)
    assert bits_equal(syn.history.history.reshape(B, D, n), fin["history"])


def test_engine_two_compartment_nonlinear_teacher_forced():
    # plateau and exponential initiation on: under teacher forcing (the
    # oracle's raster fed to both) a neuron's own spike can differ only where
    # its membrane is within 1e-4 of the threshold (device exp / sigmoid
    # rounding); until such a flip its states stay within 1e-5 relative
    from btorch_amd.models import RecurrentNN as RNN
    from btorch_amd.models import MixedNeuronPopulation, TwoCompartmentGLIF
    from btorch_amd.models._engine import Engine, match_cell

    n, B, T, D = 640, 2, 120, 5
    cexp = _tc_net(n, D, 23)
    kw = _tc_kwargs(n, 9, nonlinear=True)
    syn = HeterSynapsePSC(n, 1, None, SparseConn(cexp), base_psc=ExponentialPSC,
                          max_delay_steps=D, use_circular_buffer=True, tau_syn=5.0)
    model = RNN(MixedNeuronPopulation([(n, TwoCompartmentGLIF(n, **kw))]), syn,
                step_mode="m").to(DEV)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    x = (torch.from_numpy(workloads.poisson_input(T, B, n, 100.0, 1.0, seed=8) * 60.0)
         + 16.0).float()
    net = rsnn.Net(n, cexp, rsnn.TwoCompParams.make(**kw), torch.tensor(5.0), 1.0,
                   delay_mode="hetero", max_delay_steps=D, circular=True)
    ref = rsnn.simulate(net, x)
    assert ref["z"].sum() > 500
    teacher = ref["z"]
    ref = rsnn.simulate(net, x, teacher=teacher)
    eng = Engine(match_cell(model))
    with environ.context(dt=1.0), torch.no_grad():
        raster, recs = eng.run(x.to(DEV).contiguous(), T, B, {"z", "v", "v_pre", "i_a", "psc"},
                               teacher=pack_bits(teacher.to(DEV)).contiguous())
    thr = torch.as_tensor(-50.0) + torch.cat([torch.zeros(1, B, n), ref["theta_th"][:-1]])
    z = recs["z"].cpu()
    flip = z != ref["z"]
    first = flip & (torch.cumsum(flip.int(), 0) == 1)
    assert bool(((ref["v_pre"] - thr).abs()[first] <= 1e-4 * thr.abs()[first]).all())
    assert int(flip.any(0).sum()) <= B * n // 100
    good = torch.cumsum(flip.int(), 0) == 0
    for k in ("v_pre", "i_a"):
        torch.testing.assert_close(recs[k].cpu()[good], ref[k][good], rtol=1e-5, atol=1e-5)
    # psc sums the (forced) pushes only: bit for bit everywhere
    assert bits_equal(recs["psc"], ref["psc"])


# ---------------------------------------------------------------------------
# Recorded spikes in coordinate form
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n,p", [(1, 0.5), (63, 0.2), (64, 0.2), (65, 0.3), (700, 0.05),
                                 (700, 0.0), (129, 1.0)])
def test_raster_to_sparse_matches_nonzero_order(n, p):
    from btorch_amd import io

    g = torch.Generator().manual_seed(n)
    z = (torch.rand(7, 3, n, generator=g) < p).float()
    ref = rsnn.spikes_sparse_repr(z.numpy())
    words = pack_bits(z.to(DEV))
    got = io.raster_to_sparse(words, n)
    assert got["spikes"]["nnz"] == ref["spikes"]["nnz"]
    for k in ("spikes_idx_time", "spikes_idx_batch", "spikes_idx_neuron"):
        assert got[k].dtype == torch.int32
        assert np.array_equal(got[k].cpu().numpy(), ref[k]), k
    back = io.sparse_to_spikes(got)
    assert torch.equal(back.cpu(), z)
    dense = io.spikes_to_sparse(z.to(DEV))
    assert np.array_equal(dense["spikes_idx_neuron"].cpu().numpy(), ref["spikes_idx_neuron"])


def test_engine_raster_to_sparse():
    # the engine's own raster: the coordinates of a recorded run equal
    # np.nonzero of the run's dense spikes
    from btorch_amd import io

    n, B, T = 500, 2, 60
    model = RecurrentNN(LIF(n, tau_ref=2.0), ExponentialPSC(n, 5.0, SparseConn(
        _conn(n, n, 0.05, 3))), step_mode="m").to(DEV)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 300.0, 0.1, seed=2) * 25.0)
    eng = model.fused_engine()
    with environ.context(dt=0.1), torch.no_grad():
        raster, recs = eng.run(x.to(DEV).float().contiguous(), T, B, {"z"})
    ref = rsnn.spikes_sparse_repr(recs["z"].cpu().numpy())
    got = io.raster_to_sparse(raster, n)
    assert ref["spikes"]["nnz"] > 50
    for k in ("spikes_idx_time", "spikes_idx_batch", "spikes_idx_neuron"):
        assert np.array_equal(got[k].cpu().numpy(), ref[k]), k


@pytest.mark.parametrize("scale", [1.0, 30.0])
def test_plain_config_full_size_free_running_bit_exact(scale):
    """The reference's CPU-runnable configuration at full size: N = 4096 LIF
    (80/20 E/I, ER p = 0.02), 1000 Poisson input channels at 10 Hz through a
    fixed 0.05-density projection, B = 1, T = 1000, dt = 0.1 ms; free
    running (no teacher forcing), every step's raster, v and psc equal to the
    oracle bit for bit.  scale 30 multiplies the input so the net spikes."""
    n, n_in, B, T = 4096, 1000, 1, 1000
    model, conn, inp = _plain_model(n, 0.02, 0, n_in=n_in)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    xin = torch.from_numpy(workloads.poisson_input(T, B, n_in, 10.0, 0.1, seed=0) * scale)
    model.update_state_names = ["neuron.v", "synapse.psc"]
    with environ.context(dt=0.1), torch.no_grad():
        z, st = model(xin.to(DEV))
    x_proj = rsnn.NativeSparse(workloads.input_weights(n_in, n, 0.05, 0))(
        xin.reshape(T * B, n_in)).reshape(T, B, n)
    ref = rsnn.simulate(_net(n, conn), x_proj)
    print(f"\nconfig-0 scale {scale}: {int(ref['z'].sum())} spikes over {T} steps", flush=True)
    if scale > 1.0:
        assert ref["z"].sum() > 100
    assert bits_equal(z, ref["z"])
    assert bits_equal(st["neuron.v"], ref["v"]) and bits_equal(st["synapse.psc"], ref["psc"])


def test_engine_many_lanes_two_segments_per_push_workgroup():
    # lanes x segments >= 4096 switches the push to two segments per
    # workgroup; rasters, v and psc stay bit-exact
    n, B, T, D = 40_000, 104, 24, 5
    c, cexp = _delayed_net(n, D, 71, k_max=30)
    assert B * ((n + 1023) // 1024) >= 4096
    syn = HeterSynapsePSC(n, 1, None, SparseConn(cexp), base_psc=ExponentialPSC,
                          max_delay_steps=D, use_circular_buffer=True, tau_syn=5.0)
    model = RecurrentNN(LIF(n, tau_ref=2.0), syn, step_mode="m",
                        update_state_names=["neuron.v", "synapse.psc"]).to(DEV)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 2000.0, 0.1, seed=71) * 25.0)
    with environ.context(dt=0.1), torch.no_grad():
        z, st = model(x.to(DEV))
    ref = rsnn.simulate(_net(n, cexp, delay_mode="hetero", max_delay_steps=D, circular=True), x)
    assert ref["z"].sum() > 1000
    assert bits_equal(z, ref["z"])
    assert bits_equal(st["neuron.v"], ref["v"]) and bits_equal(st["synapse.psc"], ref["psc"])


def test_sparse_conn_matches_committed_sparse_mm_fixture():
    # the committed torch-CPU sparse.mm output (tests/golden): the HIP
    # product in ordered mode reproduces it bit for bit
    from pathlib import Path
    g = np.load(Path(__file__).with_name("golden") / "sparse_mm_order.npz")
    n_src, n_dst = int(g["n_src"]), int(g["n_dst"])
    m = scipy.sparse.coo_array((g["w"], (g["src"], g["dst"])), shape=(n_src, n_dst))
    conn = SparseConn(m, enforce_dale=False).to(DEV)
    with torch.no_grad():
        out = conn(torch.from_numpy(g["x"]).to(DEV)).cpu().numpy()
    assert np.array_equal(out.view(np.uint32), g["out"].view(np.uint32))
