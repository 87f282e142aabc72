"""GPU checks of the hand-written adjoint: against the reference's autograd
(torch-CPU restatement with the surrogate backward) and against float64
finite differences on a sigmoid-relaxed forward."""

import math

import numpy as np
import pytest
import torch

from btorch_amd import workloads
from btorch_amd.models import (
    LIF, ExponentialPSC, RecurrentNN, Sigmoid, SparseConn, SuperSpike, environ, functional,
)
from btorch_amd.models._adjoint import adjoint_chunk
from oracle import rsnn

pytestmark = pytest.mark.gpu
DEV = "cuda"


def rel_err(a, b):
    a, b = a.detach().double().cpu(), b.detach().double().cpu()
    return float((a - b).norm() / b.norm().clamp_min(1e-30))


def _model(n, seed, tau, surrogate, hard=False, tau_ref=2.0, checkpoint=False, chunk=None):
    c = workloads.plain_lif_net(n=n, p=0.08, seed=seed)
    conn = c.matrix()
    lif = LIF(n, tau=tau, tau_ref=tau_ref, hard_reset=hard, trainable_param={"tau"},
              surrogate_function=surrogate)
    model = RecurrentNN(lif, ExponentialPSC(n, 5.0, SparseConn(conn)), step_mode="m",
                        unroll=10, grad_checkpoint=checkpoint, chunk_size=chunk,
                        update_state_names=[]).to(DEV)
    return model, conn


@pytest.mark.parametrize("hard,checkpoint", [(False, False), (False, True), (True, True)])
def test_adjoint_matches_reference_autograd(hard, checkpoint):
    n, B, T = 256, 4, 60
    rng = np.random.default_rng(3)
    tau = torch.from_numpy(rng.uniform(10, 30, n).astype(np.float32))
    sur = SuperSpike(alpha=25.0 / (math.sqrt(2) - 1))  # fast sigmoid, slope 25
    model, conn = _model(n, 41, tau, sur, hard=hard, checkpoint=checkpoint, chunk=20)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 400.0, 0.1, seed=41) * 25.0)
    R = torch.randn(T, B, n, generator=torch.Generator().manual_seed(1))
    q = torch.randn(B, n, generator=torch.Generator().manual_seed(2))
    xg = x.to(DEV).requires_grad_(True)
    with environ.context(dt=0.1):
        z, _ = model(xg)
    loss = (z * R.to(DEV)).sum() + (model.neuron.v * q.to(DEV)).sum() + \
        (model.synapse.psc * q.to(DEV)).sum()
    loss.backward()
    # reference autograd on CPU, float32, same ops as the reference
    conn_m = model.synapse.linear
    dst, src = conn_m.indices[0].cpu(), conn_m.indices[1].cpu()
    mag = conn_m.magnitude.detach().cpu().clone().requires_grad_(True)
    tau_r = tau.clone().requires_grad_(True)
    xr = x.clone().requires_grad_(True)
    zr, vr, _, pr = rsnn.simulate_grad(n, dst, src, mag, xr, tau_r, 0.1, 5.0, tau_ref=2.0,
                                       hard_reset=hard,
                                       surrogate=("superspike", sur.alpha, 1.0, True))
    lr = (zr * R).sum() + (vr * q).sum() + (pr * q).sum()
    lr.backward()
    assert torch.equal(z.detach().cpu(), zr.detach()), "forward rasters differ"
    assert zr.sum() > 20
    e_w = rel_err(conn_m.magnitude.grad, mag.grad)
    e_t = rel_err(model.neuron.tau.grad, tau_r.grad)
    e_x = rel_err(xg.grad, xr.grad)
    print(f"\nrelative L2 error vs reference autograd: dW {e_w:.2e} dtau {e_t:.2e} dx {e_x:.2e}")
    assert e_w < 1e-4 and e_t < 1e-4 and e_x < 1e-4


def test_adjoint_matches_finite_differences_relaxed():
    # Sigmoid(alpha, damping = k*alpha/4, spiking=False): forward sigmoid(k alpha u),
    # backward exactly its derivative -> a smooth model finite differences can check.
    n, B, T, dt = 48, 2, 25, 0.5
    alpha = 2.0
    k = 2.0 * math.log(math.sqrt(2.0) + 1.0)
    rng = np.random.default_rng(7)
    tau = torch.from_numpy(rng.uniform(10, 30, n))
    c = workloads.plain_lif_net(n=n, p=0.15, seed=7)
    dst, src, w = rsnn.coalesce(c.matrix())
    x = torch.from_numpy(rng.uniform(0, 3, (T, B, n)))
    R = torch.from_numpy(rng.normal(size=(T, B, n)))
    q = torch.from_numpy(rng.normal(size=(B, n)))
    sur = ("sigmoid", alpha, k * alpha / 4.0, False)

    def loss64(mag, tau_):
        z, v, _, _ = rsnn.simulate_grad(n, dst, src, mag, x, tau_, dt, 5.0, surrogate=sur)
        return (z * R).sum() + (v * q).sum()

    mag64 = torch.from_numpy(w.astype(np.float64))
    traj = {}
    z, v, _, _ = rsnn.simulate_grad(n, dst, src, mag64.float(), x.float(), tau.float(), dt, 5.0,
                                    surrogate=sur, traj=traj)
    # the HIP adjoint on that trajectory
    lif = LIF(n, tau=tau.float(), trainable_param={"tau"},
              surrogate_function=Sigmoid(alpha, damping_factor=k * alpha / 4.0, spiking=False))
    conn = SparseConn(c.matrix())
    model = RecurrentNN(lif, ExponentialPSC(n, 5.0, conn), step_mode="m").to(DEV)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    spec = model.fused_engine().spec
    tr = {key: torch.stack(vals).float().to(DEV) for key, vals in traj.items()}
    tr["z"] = z.detach().float().to(DEV)
    tr["ref_prev"] = None
    with environ.context(dt=dt):
        out = adjoint_chunk(spec, tr, {"gz": R.float().to(DEV), "gv": q.float().to(DEV)},
                            conn.packed_weight().detach(), dt, valued_weights=True)
    gmag = torch.zeros(conn.nnz, dtype=torch.float64)
    gmag[conn._csr_perm.cpu()] = out["gw"].double().cpu()
    eps = 1e-6
    picks = rng.choice(len(w), size=8, replace=False)
    for e in picks:
        mp, mm = mag64.clone(), mag64.clone()
        mp[e] += eps
        mm[e] -= eps
        fd = float((loss64(mp, tau) - loss64(mm, tau)) / (2 * eps))
        assert abs(gmag[e] - fd) <= 2e-3 * max(1.0, abs(fd)), (e, float(gmag[e]), fd)
    gt = out["gtau"].double().cpu()
    for j in rng.choice(n, size=6, replace=False):
        tp, tm = tau.clone(), tau.clone()
        tp[j] += 1e-4
        tm[j] -= 1e-4
        fd = float((loss64(mag64, tp) - loss64(mag64, tm)) / 2e-4)
        assert abs(gt[j] - fd) <= 2e-3 * max(1.0, abs(fd)), (j, float(gt[j]), fd)


@pytest.mark.parametrize("kind,checkpoint", [("hetero-circular", False), ("hetero-circular", True),
                                             ("hetero-cat", True), ("fixed", True)])
def test_adjoint_delayed_matches_reference_autograd(kind, checkpoint):
    # reverse delay ring: dW, dtau and dx through HeterSynapsePSC / DelayedPSC,
    # with checkpoint chunks shorter than the delays' reach
    from btorch_amd.models import DelayedPSC, HeterSynapsePSC

    n, B, T, D = 220, 3, 48, 5
    rng = np.random.default_rng(11)
    tau = torch.from_numpy(rng.uniform(10, 30, n).astype(np.float32))
    sur = SuperSpike(alpha=25.0 / (math.sqrt(2) - 1))
    lif = LIF(n, tau=tau, tau_ref=2.0, trainable_param={"tau"}, surrogate_function=sur)
    if kind.startswith("hetero"):
        c = workloads.powerlaw_net_fast(n=n, k_min=10, k_max=60, delay_lo=1, delay_hi=D - 1,
                                        seed=12)
        conn = c.expanded(D)
        syn = HeterSynapsePSC(n, 1, None, SparseConn(conn), base_psc=ExponentialPSC,
                              max_delay_steps=D, use_circular_buffer=kind.endswith("circular"),
                              tau_syn=5.0)
        csr_d, ages = D, tuple(range(D))
    else:
        conn = workloads.plain_lif_net(n=n, p=0.08, seed=12).matrix()
        syn = DelayedPSC(ExponentialPSC(n, 5.0, SparseConn(conn)), max_delay_steps=3)
        csr_d, ages = 1, (3,)
    model = RecurrentNN(lif, syn, step_mode="m", grad_checkpoint=checkpoint, chunk_size=7,
                        unroll=7, update_state_names=[]).to(DEV)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 400.0, 0.1, seed=12) * 25.0)
    R = torch.randn(T, B, n, generator=torch.Generator().manual_seed(5))
    q = torch.randn(B, n, generator=torch.Generator().manual_seed(6))
    xg = x.to(DEV).requires_grad_(True)
    with environ.context(dt=0.1):
        z, _ = model(xg)
    eng = model.fused_engine()
    psc_mod = eng.spec.exp_psc
    loss = (z * R.to(DEV)).sum() + (model.neuron.v * q.to(DEV)).sum() + \
        (psc_mod.psc.reshape(B, n) * q.to(DEV)).sum()
    loss.backward()
    conn_m = eng.spec.conn
    dst, src = conn_m.indices[0].cpu(), conn_m.indices[1].cpu()
    mag = conn_m.magnitude.detach().cpu().clone().requires_grad_(True)
    tau_r = tau.clone().requires_grad_(True)
    xr = x.clone().requires_grad_(True)
    zr, vr, _, pr = rsnn.simulate_grad(n, dst, src, mag, xr, tau_r, 0.1, 5.0, tau_ref=2.0,
                                       surrogate=("superspike", sur.alpha, 1.0, True),
                                       csr_d=csr_d, ages=ages)
    lr = (zr * R).sum() + (vr * q).sum() + (pr * q).sum()
    lr.backward()
    assert torch.equal(z.detach().cpu(), zr.detach()), "forward rasters differ"
    assert zr.sum() > 20
    e_w = rel_err(conn_m.magnitude.grad, mag.grad)
    e_t = rel_err(model.neuron.tau.grad, tau_r.grad)
    e_x = rel_err(xg.grad, xr.grad)
    print(f"\n[{kind}, checkpoint={checkpoint}] relative L2 error vs reference autograd: "
          f"dW {e_w:.2e} dtau {e_t:.2e} dx {e_x:.2e}")
    assert e_w < 1e-4 and e_t < 1e-4 and e_x < 1e-4
    # the module's delay memory continues the run exactly like the forward path
    with torch.no_grad(), environ.context(dt=0.1):
        z2, _ = model(x[:10].to(DEV))
    zr2, *_ = rsnn.simulate_grad(n, dst, src, mag.detach(), torch.cat([x, x[:10]]),
                                 tau, 0.1, 5.0, tau_ref=2.0, csr_d=csr_d, ages=ages)
    assert torch.equal(z2.cpu(), zr2[T:].detach())


@pytest.mark.parametrize("hard", [False, True])
def test_standalone_lif_step_backward_matches_reference_autograd(hard):
    # LIF.single_step_forward with gradients (the per-step loop a cell the
    # fused engine does not match runs through): input, membrane and tau
    # gradients against the reference-style autograd (rel. L2 1e-4)
    n, B, T = 300, 3, 40
    rng = np.random.default_rng(7)
    tau = torch.from_numpy(rng.uniform(10, 30, n).astype(np.float32))
    sur = SuperSpike(alpha=25.0 / (math.sqrt(2) - 1))
    lif = LIF(n, tau=tau, tau_ref=2.0, hard_reset=hard, trainable_param={"tau"},
              surrogate_function=sur).to(DEV)
    functional.init_net_state(lif, batch_size=B, device=DEV, dtype=torch.float32)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 3000.0, 0.1, seed=7) * 25.0)
    R = torch.randn(T, B, n, generator=torch.Generator().manual_seed(3))
    q = torch.randn(B, n, generator=torch.Generator().manual_seed(4))
    xg = x.to(DEV).requires_grad_(True)
    with environ.context(dt=0.1):
        z = torch.stack([lif(xg[t]) for t in range(T)])
    ((z * R.to(DEV)).sum() + (lif.v * q.to(DEV)).sum()).backward()
    xr = x.clone().requires_grad_(True)
    tau_r = tau.clone().requires_grad_(True)
    empty = torch.zeros(0, dtype=torch.long)
    zr, vr, _, _ = rsnn.simulate_grad(n, empty, empty, torch.zeros(0), xr, tau_r, 0.1, 5.0,
                                      tau_ref=2.0, hard_reset=hard,
                                      surrogate=("superspike", sur.alpha, 1.0, True))
    ((zr * R).sum() + (vr * q).sum()).backward()
    assert torch.equal(z.detach().cpu(), zr.detach())
    assert zr.sum() > 50
    assert rel_err(xg.grad, xr.grad) < 1e-4
    assert rel_err(lif.tau.grad, tau_r.grad) < 1e-4


def test_per_step_loop_gradients_equal_fused_adjoint():
    # the same cell differentiated two ways: the fused engine's adjoint
    # (model(x)) and torch autograd through the standalone HIP steps of the
    # per-step loop (model(x, None) takes that loop); rasters equal, gradients
    # within 1e-4 relative
    n, B, T = 256, 2, 40
    rng = np.random.default_rng(11)
    tau = torch.from_numpy(rng.uniform(10, 30, n).astype(np.float32))
    sur = SuperSpike(alpha=25.0 / (math.sqrt(2) - 1))
    model, _ = _model(n, 43, tau, sur)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 400.0, 0.1, seed=43) * 25.0).to(DEV)
    R = torch.randn(T, B, n, generator=torch.Generator().manual_seed(5)).to(DEV)
    grads, rasters = [], []
    for per_step in (False, True):
        model.zero_grad(set_to_none=True)
        functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
        xg = x.clone().requires_grad_(True)
        with environ.context(dt=0.1):
            z, _ = model(xg, None) if per_step else model(xg)
        (z * R).sum().backward()
        rasters.append(z.detach().cpu())
        grads.append((xg.grad.clone(), model.synapse.linear.magnitude.grad.clone(),
                       model.neuron.tau.grad.clone()))
    assert torch.equal(rasters[0], rasters[1]) and rasters[0].sum() > 20
    for a, b in zip(*grads):
        assert rel_err(b, a) < 1e-4


def test_standalone_glif3_step_backward_matches_reference_autograd():
    # GLIF3.single_step_forward with gradients: input and tau gradients, and
    # the after-spike currents carried through time, against the
    # reference-style autograd of glif.py's ops in float32 on the CPU
    from btorch_amd.models import GLIF3

    n, B, T, dt = 300, 2, 40, 0.1
    rng = np.random.default_rng(9)
    tau = torch.from_numpy(rng.uniform(10, 30, n).astype(np.float32))
    sur = SuperSpike(alpha=25.0 / (math.sqrt(2) - 1))
    kw = dict(v_threshold=-50.0, v_reset=-70.0, c_m=0.05, k=[0.3, 0.02], asc_amps=[-0.6, 0.9],
              tau_ref=2.0)
    glif = GLIF3(n, tau=tau, trainable_param={"tau"}, surrogate_function=sur, **kw).to(DEV)
    functional.init_net_state(glif, batch_size=B, device=DEV, dtype=torch.float32)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 3000.0, 0.1, seed=9) * 8.0)
    R = torch.randn(T, B, n, generator=torch.Generator().manual_seed(2))
    xg = x.to(DEV).requires_grad_(True)
    with environ.context(dt=dt):
        z = torch.stack([glif(xg[t]) for t in range(T)])
    ((z * R.to(DEV)).sum() + glif.Iasc.sum()).backward()
    p = rsnn.GLIF3Params.make(n, tau=tau, **kw)
    xr = x.clone().requires_grad_(True)
    tr = tau.clone().requires_grad_(True)
    v = torch.full((B, n), -70.0)
    ia = torch.zeros(B, n, 2)
    ref = torch.zeros(B, n)
    zs = []
    for t in range(T):
        derivative = -(v - p.v_rest) / tr + (xr[t] + ia.sum(-1)) / p.c_m
        lin = -1.0 / tr
        v = v + torch.expm1(dt * lin) / lin * derivative
        lin_i = -p.k
        ia = ia + torch.expm1(dt * lin_i) / lin_i * (-p.k * ia)
        s = rsnn._Spike.apply((v - p.v_threshold) / (p.v_threshold - p.v_reset), "superspike",
                              sur.alpha, 1.0, True)
        s = s * (ref == 0).detach().to(s.dtype)
        v = v - (p.v_threshold - p.v_reset) * s
        ia = ia + p.asc_amps * s[..., None]
        ref = torch.relu(ref + s * p.tau_ref - dt)
        zs.append(s)
    zr = torch.stack(zs)
    ((zr * R).sum() + ia.sum()).backward()
    assert torch.equal(z.detach().cpu(), zr.detach()) and zr.sum() > 50
    assert rel_err(xg.grad, xr.grad) < 1e-4
    assert rel_err(glif.tau.grad, tr.grad) < 1e-4


def _tc_autograd_step(st, soma, apical, p, dt, sur_alpha):
    """two_compartment.py:208-306 op for op, with the surrogate backward."""
    da = dt / torch.clamp(p["tau_a"], min=1e-4)
    i_a = (st["i_a"] + da * ((apical + p["w_Ca"] * torch.sigmoid(st["i_a"] - p["theta_Ca"]))
                             + st["i_bap"])) / (1.0 + da)
    thr = p["v_threshold"] + st["theta_th"]
    d_t = torch.clamp(p["delta_T"], min=0.0)
    arg = torch.clamp((st["v"] - thr) / torch.clamp(d_t, min=1e-4), max=10.0)
    ex = torch.where(d_t > 0.0, d_t * torch.exp(arg), torch.zeros_like(st["v"]))
    ds = dt / torch.clamp(p["tau_s"], min=1e-4)
    v_pre = (st["v"] + ds * ((p["E_L"] + torch.clamp(p["R_s"], min=1e-4)
                              * (soma + p["w_as"] * i_a)) + ex)) / (1.0 + ds)
    s = rsnn._Spike.apply(v_pre - thr, "superspike", sur_alpha, 1.0, True)
    dth = dt / torch.clamp(p["tau_th"], min=1e-4)
    return s, {"v": v_pre - (v_pre - p["v_reset"]) * s, "i_a": i_a, "i_bap": p["w_sa"] * s,
               "theta_th": st["theta_th"] / (1.0 + dth) + p["delta_th"] * s}


@pytest.mark.parametrize("nonlinear", [False, True])
def test_standalone_two_compartment_backward_matches_autograd(nonlinear):
    # TwoCompartmentGLIF.single_step_forward with gradients. Linear regime:
    # spiking, bit-exact forward, gradients vs float32 CPU autograd (1e-4).
    # Plateau and exponential initiation on: a few sub-threshold steps (the
    # device's exp and sigmoid differ in the last ulp), vs float64 (1e-4).
    from btorch_amd.models import TwoCompartmentGLIF

    n, B, dt = 200, 2, 0.1
    T = 5 if nonlinear else 30
    sur = SuperSpike(alpha=25.0 / (math.sqrt(2) - 1))
    kw = dict(v_threshold=-50.0, v_reset=-65.0, tau_s=20.0, R_s=1.0, E_L=-70.0, tau_a=100.0,
              tau_th=50.0, delta_th=1.5, w_sa=0.5, w_as=1.0)
    kw.update(dict(delta_T=2.0, w_Ca=3.0, theta_Ca=1.0) if nonlinear
              else dict(delta_T=0.0, w_Ca=0.0, theta_Ca=1.0))
    cell = TwoCompartmentGLIF(n, surrogate_function=sur, **kw).to(DEV)
    functional.init_net_state(cell, batch_size=B, device=DEV, dtype=torch.float32)
    g = torch.Generator().manual_seed(12)
    scale = 50.0 if nonlinear else 3000.0
    soma = torch.rand(T, B, n, generator=g) * scale
    apical = torch.rand(T, B, n, generator=g) * 20.0
    R = torch.randn(T, B, n, generator=g)
    sg_, ag_ = soma.to(DEV).requires_grad_(True), apical.to(DEV).requires_grad_(True)
    with environ.context(dt=dt):
        outs = [cell.single_step_forward(sg_[t], ag_[t]) for t in range(T)]
    z = torch.stack([o[0] for o in outs])
    ((z * R.to(DEV)).sum() + outs[-1][1].sum() + cell.i_a.sum() + cell.theta_th.sum()).backward()
    dtype = torch.float64 if nonlinear else torch.float32
    p = {k: torch.tensor(float(v), dtype=dtype) for k, v in kw.items()}
    sr = soma.to(dtype).requires_grad_(True)
    ar = apical.to(dtype).requires_grad_(True)
    st = {"v": torch.full((B, n), -65.0, dtype=dtype), "i_a": torch.zeros(B, n, dtype=dtype),
          "i_bap": torch.zeros(B, n, dtype=dtype), "theta_th": torch.zeros(B, n, dtype=dtype)}
    zs = []
    for t in range(T):
        s, st = _tc_autograd_step(st, sr[t], ar[t], p, dt, sur.alpha)
        zs.append(s)
    zr = torch.stack(zs)
    ((zr * R.to(dtype)).sum() + st["v"].sum() + st["i_a"].sum() + st["theta_th"].sum()).backward()
    assert torch.equal(z.detach().cpu(), zr.detach().float())
    if nonlinear:
        assert zr.sum() == 0
    else:
        assert zr.sum() > 50
    assert rel_err(sg_.grad, sr.grad) < 1e-4
    assert rel_err(ag_.grad, ar.grad) < 1e-4


def test_apical_two_compartment_cell_trains_through_the_per_step_loop():
    # ApicalRecurrentNN over a two-compartment population with gradients:
    # the per-step loop of standalone HIP steps; its raster equals the fused
    # forward's and the input gradients are finite and not all zero
    import scipy.sparse

    from btorch_amd.models import (ApicalRecurrentNN, ExponentialPSC, MixedNeuronPopulation,
                                   TwoCompartmentGLIF)

    n, B, T = 300, 2, 30
    m = workloads.plain_lif_net(n=n, p=0.05, seed=13).matrix()
    m = scipy.sparse.coo_array((m.data * 40.0, (m.row, m.col)), shape=m.shape)
    sur = SuperSpike(alpha=25.0 / (math.sqrt(2) - 1))
    cell = TwoCompartmentGLIF(n, surrogate_function=sur, delta_th=1.0)
    model = ApicalRecurrentNN(MixedNeuronPopulation([(n, cell)]),
                              ExponentialPSC(n, 5.0, SparseConn(m)), step_mode="m",
                              update_state_names=[]).to(DEV)
    g = torch.Generator().manual_seed(13)
    x = (torch.rand(T, B, n, generator=g) * 3000.0).to(DEV)
    xa = (torch.rand(T, B, n, generator=g) * 20.0).to(DEV)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    with environ.context(dt=0.1), torch.no_grad():
        z0, _ = model(x, None, xa)
    functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
    xg = x.clone().requires_grad_(True)
    with environ.context(dt=0.1):
        z1, _ = model(xg, None, xa)
    z1.sum().backward()
    assert torch.equal(z0.cpu(), z1.detach().cpu()) and z0.sum() > 50
    assert torch.isfinite(xg.grad).all() and xg.grad.abs().sum() > 0
