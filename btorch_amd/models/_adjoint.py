"""Training path of the fused engine: BPTT through the hand-written adjoint.

The forward of a chunk of K steps runs the fused engine; its backward walks
the chunk in reverse with ``btx_adj_transpose`` (W dL/dpsc),
``btx_adj_wgrad`` (event-driven dL/dW over the step's spiking neurons) and
``btx_adj_lif_step`` (surrogate derivative, reset and synapse-decay adjoint,
dL/dtau).  With delayed synapses the spike of step t reaches psc at steps
t + age, so the transposed product and the weight gradient read dL/dpsc of
those later steps (``btx_adj_transpose_delayed`` / ``btx_adj_wgrad_delayed``,
the reverse delay ring); the spikes of the ring_d steps before a chunk enter
it as a differentiable input and receive their gradient from it.  With ``grad_checkpoint`` a chunk keeps only its entry state and
recomputes its forward in the backward (checkpoint every ``chunk_size``
steps, rnn.py:177-240); otherwise the forward records are kept.

Gradients flow to the packed weights (and from there to ``magnitude``), to
a trainable ``tau``, to the input sequence and to the entry state v / psc.
The refractory counter only enters through a detached comparison
(lif.py:165-167), so it carries no gradient.
"""

from __future__ import annotations

import torch

from .. import _lib
from . import _plan
from .surrogate import Erf, Sigmoid, SuperSpike


def needs_grad(rnn, x: torch.Tensor) -> bool:
    if not torch.is_grad_enabled():
        return False
    if x.requires_grad:
        return True
    if any(p.requires_grad for p in rnn.parameters()):
        return True
    for m in rnn.modules():
        v = getattr(m, "v", None)
        if torch.is_tensor(v) and v.requires_grad:
            return True
    return False


def _surrogate_code(fn, require_spiking: bool = True) -> tuple[int, float, float]:
    if isinstance(fn, SuperSpike):
        kind = _lib.SURR_SUPERSPIKE
    elif isinstance(fn, Erf):
        kind = _lib.SURR_ERF
    elif isinstance(fn, Sigmoid):
        kind = _lib.SURR_SIGMOID
    else:
        raise NotImplementedError(f"surrogate {type(fn).__name__} has no adjoint kernel")
    if require_spiking and not fn.spiking:
        raise NotImplementedError("the fused engine runs the spiking (Heaviside) forward")
    return kind, float(fn.alpha), float(fn.damping_factor)


def fused_adjoint_covers(spec, lanes: int) -> bool:
    """Whether the fused adjoint differentiates this cell; otherwise
    RecurrentNN trains through the per-step loop of standalone steps."""
    try:
        _check_supported(spec, lanes)
    except NotImplementedError:
        return False
    return True


def _check_supported(spec, lanes):
    if spec.syn_keys:
        raise NotImplementedError("the fused adjoint covers the exponential synapse (the alpha "
                                  "and dual-exponential synapses are forward-only in the "
                                  "fused engine)")
    neuron = spec.neuron
    if hasattr(neuron, "Iasc") or hasattr(neuron, "i_bap"):
        raise NotImplementedError("the fused adjoint covers LIF neurons (GLIF3 and the "
                                  "two-compartment neuron are forward-only)")
    for name in ("v_threshold", "v_reset", "c_m", "tau_ref"):
        t = getattr(neuron, name, None)
        if isinstance(t, torch.nn.Parameter) and t.requires_grad:
            raise NotImplementedError(f"gradient w.r.t. LIF.{name} is not provided")
    if lanes > 64:
        raise NotImplementedError("the fused adjoint supports at most 64 lanes")
    # a surrogate without an adjoint kernel (or a non-spiking forward)
    # trains through the per-step loop, whose steps differentiate any
    # surrogate in torch ops
    _surrogate_code(neuron.surrogate_function)


def _delayed(spec) -> bool:
    return spec.history is not None or spec.csr_d != 1 or spec.bins != [(0, 0)]


def _ages(spec, device) -> torch.Tensor:
    """Ring age of every virtual-row offset (-1 = unused), int32 on device."""
    ages = [-1] * spec.csr_d
    for age, sub in spec.bins:
        ages[sub] = age
    return torch.tensor(ages, dtype=torch.int32, device=device)


def adjoint_chunk(spec, traj: dict, grads: dict, w_csr: torch.Tensor, dt: float,
                  valued_weights: bool = False, past: torch.Tensor | None = None) -> dict:
    """Reverse a chunk.

    traj: (K, B, N) ``v_prev``, ``v_pre``, ``z`` and ``ref_prev`` (or None),
        the forward values of every step.
    grads: ``gz`` (K, B, N) or None, ``gv``/``gpsc`` (B, N) or None -- the
        gradients of the outputs and of the state after the chunk.
    valued_weights: z holds real values (relaxed forward) -- the weight
        gradient then uses the valued product instead of spike masks.
    past: delayed synapses only -- (D, B, N) spikes of the D = ring_d steps
        before the chunk, oldest first.
    Returns ``gx`` (K,B,N), ``gw`` (E,) in CSR order, ``gtau`` ((N,) or (1,)),
    ``gv``/``gpsc`` (B, N) of the state entering the chunk, and with delays
    ``gpast`` (D, B, N), the gradient of the spikes before the chunk.
    """
    neuron, syn, conn = spec.neuron, spec.exp_psc, spec.conn
    z = traj["z"]
    K, B, n = z.shape
    dev = z.device
    lib = _lib.load()
    stream = _lib.stream_ptr()

    def nb(t):  # (B, N) -> (N, B)
        return t.transpose(-1, -2).contiguous()

    v_prev, v_pre, zz = nb(traj["v_prev"]), nb(traj["v_pre"]), nb(z)
    ref_prev = nb(traj["ref_prev"]) if traj.get("ref_prev") is not None else None
    gz = nb(grads["gz"]) if grads.get("gz") is not None else torch.zeros_like(zz)
    gv = nb(grads["gv"]) if grads.get("gv") is not None else torch.zeros(n, B, device=dev)
    gpsc = nb(grads["gpsc"]) if grads.get("gpsc") is not None else torch.zeros(n, B,
                                                                              device=dev)
    gsyn = torch.empty(n, B, device=dev)
    gx = torch.empty(K, n, B, device=dev)
    gw = torch.zeros(conn.nnz, device=dev)
    keep: list = []
    plan = _plan.lif_plan(neuron, B, dev, keep)
    tau_vec, tau_s = keep[0], plan.tau_stride
    gtau = torch.zeros(tau_vec.numel(), device=dev)
    ts, tss, k, ks = syn.decay_coefficients(dev)
    kind, alpha, damping = _surrogate_code(neuron.surrogate_function, require_spiking=False)
    a = _lib.AdjLifArgs()
    a.n, a.lanes, a.dt = n, B, float(dt)
    a.hard_reset = int(bool(neuron.hard_reset))
    a.detach_reset = int(bool(neuron.detach_reset))
    a.surrogate, a.alpha, a.damping = kind, alpha, damping
    a.tau, a.tau_stride = plan.tau, plan.tau_stride
    a.v_th, a.v_th_stride = plan.v_th, plan.v_th_stride
    a.v_reset, a.v_reset_stride = plan.v_reset, plan.v_reset_stride
    a.c_m, a.c_m_stride = plan.c_m, plan.c_m_stride
    a.tau_syn, a.tau_syn_stride = ts.data_ptr(), tss
    a.k_syn, a.k_syn_stride = k.data_ptr(), ks
    a.gsyn = gsyn.data_ptr()
    a.gv = gv.data_ptr()
    a.gpsc = gpsc.data_ptr()
    a.gtau = gtau.data_ptr()
    layout = conn._layout
    w = w_csr.detach().to(torch.float32).contiguous()
    shifts = torch.arange(B, device=dev, dtype=torch.int64)
    if not valued_weights:
        masks = ((zz != 0).to(torch.int64) << shifts).sum(-1).contiguous()  # (K, N)
    if _delayed(spec):
        if valued_weights:
            raise NotImplementedError("the relaxed weight gradient covers undelayed synapses")
        return _adjoint_delayed(spec, a, lib, stream, layout, w, masks, zz, v_prev, v_pre,
                                ref_prev, gz, gv, gpsc, gsyn, gx, gw, gtau, past, shifts)
    if not valued_weights:
        # the whole reverse sweep in one call
        a.v_prev, a.v_pre, a.z = v_prev.data_ptr(), v_pre.data_ptr(), zz.data_ptr()
        a.ref_prev = ref_prev.data_ptr() if ref_prev is not None else None
        a.gz, a.gx = gz.data_ptr(), gx.data_ptr()
        _lib.check(lib.btx_adj_lif_chunk(a, K, layout.row_ptr.data_ptr(),
                                         layout.col.data_ptr(), w.data_ptr(), masks.data_ptr(),
                                         gw.data_ptr(), stream), "btx_adj_lif_chunk")
        return {"gx": gx.permute(0, 2, 1).contiguous(), "gw": gw, "gtau": gtau,
                "gv": gv.t().contiguous(), "gpsc": gpsc.t().contiguous()}
    for t in range(K - 1, -1, -1):
        _lib.check(lib.btx_adj_transpose(layout.row_ptr.data_ptr(), layout.col.data_ptr(),
                                         w.data_ptr(), n, B, gpsc.data_ptr(),
                                         gsyn.data_ptr(), stream), "btx_adj_transpose")
        if valued_weights:
            zt = traj["z"][t].contiguous()
            from .._kernels import compact_dense, spmv_weight_grad
            act, cnt = compact_dense(zt)
            spmv_weight_grad(layout, zt, act, cnt, gpsc.t().contiguous(), gw)
        else:
            _lib.check(lib.btx_adj_wgrad(layout.row_ptr.data_ptr(), layout.col.data_ptr(), n,
                                         B, masks[t].data_ptr(), gpsc.data_ptr(),
                                         gw.data_ptr(), stream), "btx_adj_wgrad")
        a.v_prev = v_prev[t].data_ptr()
        a.v_pre = v_pre[t].data_ptr()
        a.z = zz[t].data_ptr()
        a.ref_prev = ref_prev[t].data_ptr() if ref_prev is not None else None
        a.gz = gz[t].data_ptr()
        a.gx = gx[t].data_ptr()
        _lib.check(lib.btx_adj_lif_step(a, stream), "btx_adj_lif_step")
    return {"gx": gx.permute(0, 2, 1).contiguous(), "gw": gw, "gtau": gtau,
            "gv": gv.t().contiguous(), "gpsc": gpsc.t().contiguous()}


def _adjoint_delayed(spec, a, lib, stream, layout, w, masks, zz, v_prev, v_pre, ref_prev,
                     gz, gv, gpsc, gsyn, gx, gw, gtau, past, shifts):
    """Reverse sweep of a chunk with delayed synapses (see adjoint_chunk)."""
    K, n, B = zz.shape
    dev = zz.device
    csr_d = spec.csr_d
    ages = _ages(spec, dev)
    G = torch.empty(K, n, B, device=dev)  # dL/dpsc of every step of the chunk
    rp, col = layout.row_ptr.data_ptr(), layout.col.data_ptr()
    for t in range(K - 1, -1, -1):
        G[t].copy_(gpsc)
        _lib.check(lib.btx_adj_transpose_delayed(rp, col, w.data_ptr(), n, csr_d, B,
                                                 G.data_ptr(), K, t, ages.data_ptr(),
                                                 gsyn.data_ptr(), 0, stream),
                   "btx_adj_transpose_delayed")
        _lib.check(lib.btx_adj_wgrad_delayed(rp, col, n, csr_d, B, masks[t].data_ptr(),
                                             G.data_ptr(), K, t, ages.data_ptr(), gw.data_ptr(),
                                             stream), "btx_adj_wgrad_delayed")
        a.v_prev = v_prev[t].data_ptr()
        a.v_pre = v_pre[t].data_ptr()
        a.z = zz[t].data_ptr()
        a.ref_prev = ref_prev[t].data_ptr() if ref_prev is not None else None
        a.gz = gz[t].data_ptr()
        a.gx = gx[t].data_ptr()
        _lib.check(lib.btx_adj_lif_step(a, stream), "btx_adj_lif_step")
    # the spikes before the chunk: their share of this chunk's psc
    D = past.shape[0]
    pz = past.detach().transpose(-1, -2).contiguous()  # (D, N, B)
    pmasks = ((pz != 0).to(torch.int64) << shifts).sum(-1).contiguous()
    gpast = torch.empty(D, n, B, device=dev)
    for p in range(D):
        tau = p - D
        _lib.check(lib.btx_adj_transpose_delayed(rp, col, w.data_ptr(), n, csr_d, B,
                                                 G.data_ptr(), K, tau, ages.data_ptr(),
                                                 gpast[p].data_ptr(), 0, stream),
                   "btx_adj_transpose_delayed")
        _lib.check(lib.btx_adj_wgrad_delayed(rp, col, n, csr_d, B, pmasks[p].data_ptr(),
                                             G.data_ptr(), K, tau, ages.data_ptr(),
                                             gw.data_ptr(), stream), "btx_adj_wgrad_delayed")
    return {"gx": gx.permute(0, 2, 1).contiguous(), "gw": gw, "gtau": gtau,
            "gv": gv.t().contiguous(), "gpsc": gpsc.t().contiguous(),
            "gpast": gpast.permute(0, 2, 1).contiguous()}


def _history_from_past(spec, past: torch.Tensor, cursor: int) -> None:
    """Module SpikeHistory memory holding `past` ((D, B, N), oldest first: the
    last entry is the most recent step) with the given cursor."""
    hist = spec.history
    D = past.shape[0]
    p = past.detach().to(torch.float32)
    if hist.use_circular_buffer:
        # delay d lives in slot (cursor - 1 - d) mod D and is past[D - 1 - d]
        order = [(D - 1 - ((cursor - 1 - j) % D)) for j in range(D)]
        h = p[order].permute(1, 0, 2)
    else:
        h = p.flip(0).permute(1, 0, 2)
    hist.history = h.reshape(hist.history.shape).contiguous()
    hist.set_cursor(cursor)


def _past_from_history(spec) -> torch.Tensor:
    """(D, B, N) spikes held by the module's SpikeHistory, oldest first."""
    hist = spec.history
    D = spec.ring_d
    h = hist.history.reshape(hist.history.shape[0], D, -1)
    if hist.use_circular_buffer:
        c = hist.cursor
        return torch.stack([h[:, (c - 1 - (D - 1 - p)) % D] for p in range(D)]).detach()
    return torch.stack([h[:, D - 1 - p] for p in range(D)]).detach()


def _set_state(spec, v, ref, psc):
    neuron, syn = spec.neuron, spec.exp_psc
    neuron.v = v.detach().clone().contiguous()
    if neuron._use_refractory:
        neuron.refractory = ref.detach().clone().contiguous()
    syn.psc = psc.detach().clone().contiguous()


def _forward_chunk(eng, x, v0, ref0, psc0, record, past0=None):
    spec = eng.spec
    _set_state(spec, v0, ref0, psc0)
    if past0 is not None:
        _history_from_past(spec, past0, 0)
    K, B = x.shape[0], x.shape[1]
    raster, recs = eng.run(x.detach().contiguous(), K, B, record)
    n = spec.neuron
    ref_end = n.refractory if n._use_refractory else torch.zeros_like(n.v)
    return recs, n.v, ref_end, spec.exp_psc.psc


class _FusedChunk(torch.autograd.Function):
    """K fused steps; differentiable w.r.t. x, packed weights, tau, v0, psc0."""

    @staticmethod
    def forward(ctx, eng, cfg, x, w_csr, tau, v0, ref0, psc0, past0):
        checkpoint, extra, sink = cfg
        record = {"z"} | set(extra)
        if not checkpoint:
            record |= {"v_prev", "v_pre", "refractory"}
        delayed = past0.numel() > 0
        recs, v1, ref1, psc1 = _forward_chunk(eng, x, v0, ref0, psc0, record,
                                              past0 if delayed else None)
        ctx.eng, ctx.checkpoint, ctx.tau_shape = eng, checkpoint, tau.shape
        ctx.dt = eng.spec_dt
        ctx.delayed = delayed
        if checkpoint:
            ctx.save_for_backward(x, w_csr, v0, ref0, psc0, past0)
        else:
            ctx.save_for_backward(x, w_csr, v0, ref0, psc0, past0, recs["v_prev"],
                                  recs["v_pre"], recs["z"], recs["refractory"])
        for k in extra:
            sink.setdefault(k, []).append(recs[k])
        return recs["z"], v1.clone(), ref1.clone(), psc1.clone()

    @staticmethod
    def backward(ctx, gz, gv, gref, gpsc):
        from . import environ

        # the backward runs outside the caller's environment: restore its dt
        with environ.context(dt=ctx.dt):
            return _FusedChunk._backward(ctx, gz, gv, gpsc)

    @staticmethod
    def _backward(ctx, gz, gv, gpsc):
        eng = ctx.eng
        saved = ctx.saved_tensors
        x, w_csr, v0, ref0, psc0, past0 = saved[:6]
        if ctx.checkpoint:
            # recompute the chunk from its entry state, with records
            with torch.no_grad():
                snapshot = _snapshot(eng.spec)
                recs, _, _, _ = _forward_chunk(eng, x, v0, ref0, psc0,
                                               {"z", "v_prev", "v_pre", "refractory"},
                                               past0 if ctx.delayed else None)
                _restore(eng.spec, snapshot)
            v_prev, v_pre, z, ref = recs["v_prev"], recs["v_pre"], recs["z"], recs["refractory"]
        else:
            v_prev, v_pre, z, ref = saved[6:]
        ref_prev = (torch.cat([ref0.detach().unsqueeze(0), ref[:-1]])
                    if eng.spec.neuron._use_refractory else None)
        traj = {"v_prev": v_prev, "v_pre": v_pre, "z": z, "ref_prev": ref_prev}
        out = adjoint_chunk(eng.spec, traj, {"gz": gz, "gv": gv, "gpsc": gpsc}, w_csr, ctx.dt,
                            past=past0 if ctx.delayed else None)
        gtau = out["gtau"]
        if len(ctx.tau_shape) == 0 or int(torch.Size(ctx.tau_shape).numel()) == 1:
            gtau = gtau.sum().reshape(ctx.tau_shape)
        else:
            gtau = gtau.reshape(ctx.tau_shape)
        gpast = out.get("gpast") if ctx.delayed else None
        return (None, None, out["gx"], out["gw"], gtau, out["gv"], None, out["gpsc"], gpast)


def _snapshot(spec):
    n = spec.neuron
    h = spec.history
    return (n.v, n.refractory if n._use_refractory else None, spec.exp_psc.psc,
            None if h is None else (h.history, h.cursor))


def _restore(spec, snap):
    n = spec.neuron
    n.v = snap[0]
    if n._use_refractory:
        n.refractory = snap[1]
    spec.exp_psc.psc = snap[2]
    if snap[3] is not None:
        spec.history.history = snap[3][0]
        spec.history.set_cursor(snap[3][1])


def run_with_grad(rnn, eng, x):
    """RecurrentNN.multi_step_forward under autograd through the adjoint."""
    from . import environ

    spec = eng.spec
    neuron, syn, conn = spec.neuron, spec.exp_psc, spec.conn
    n = neuron.size
    xs, _ = rnn._input_sequence(x, n)
    T, lanes = xs.shape[0], xs.shape[1]
    _check_supported(spec, lanes)
    _surrogate_code(neuron.surrogate_function)
    eng.spec_dt = float(environ.get("dt"))
    names = rnn._record_names()
    # every requested state maps onto an engine record (NotImplementedError
    # otherwise, as the no-grad path); the history memory is rebuilt from the
    # produced spikes
    want_all = rnn._wanted_records(names, spec)
    want = want_all - {"z", "history"}
    hist0 = (spec.history.history.detach().clone()
             if spec.history is not None and "history" in want_all else None)
    cursor_rec = spec.history.cursor if spec.history is not None else 0
    chunk = T
    if rnn.grad_checkpoint:
        chunk = rnn.chunk_size or (T if rnn.unroll is False else int(rnn.unroll))
    w_csr = conn.packed_weight()
    tau = neuron.tau
    v = neuron.v
    ref = neuron.refractory if neuron._use_refractory else torch.zeros_like(v)
    psc = syn.psc
    zs, sink = [], {}
    cfg = (bool(rnn.grad_checkpoint), tuple(sorted(want)), sink)
    delayed = _delayed(spec)
    if delayed:
        D = spec.ring_d
        cursor0 = spec.history.cursor
        past = _past_from_history(spec)  # constants: no gradient to the history memory
    else:
        past = torch.empty(0, device=v.device)
    for c0 in range(0, T, chunk):
        z, v, ref, psc = _FusedChunk.apply(eng, cfg, xs[c0:c0 + chunk], w_csr, tau, v, ref, psc,
                                           past)
        zs.append(z)
        if delayed:
            past = torch.cat([past, z])[-D:]
    if delayed:
        _history_from_past(spec, past, (cursor0 + T) % D
                           if spec.history.use_circular_buffer else 0)
    neuron.v = v
    if neuron._use_refractory:
        neuron.refractory = ref.detach()
    syn.psc = psc
    for m in spec.psc_modules:
        if m is not syn:
            m.psc = psc.reshape(m.psc.shape)
    shape = tuple(v.shape)
    z = torch.cat(zs).reshape(T, *shape)
    # recorded hidden states come back without gradient history
    recs = {k: torch.cat(vals) for k, vals in sink.items()}
    states = {}
    for key in names:
        rec = rnn._record_of(key, spec)
        if rec == "history":
            from ._engine import pack_bits
            words = pack_bits(z.detach().reshape(T, lanes, n))
            seq = eng.history_sequence(words, hist0, T, lanes, n, cursor_rec)
            states[key] = seq.reshape(T, *spec.history.history.shape)
        else:
            states[key] = recs[rec].reshape(T, *shape)
    return z, states
