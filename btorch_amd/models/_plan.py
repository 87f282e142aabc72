"""Host-side assembly of ``btx_rsnn_plan`` fields from module parameters."""

from __future__ import annotations

import torch

from .. import _lib
from . import environ


def param_vec(t, n: int, shape: tuple[int, ...], device) -> tuple[torch.Tensor, int]:
    """(fp32 contiguous tensor, stride) for a scalar or per-neuron parameter."""
    t = torch.as_tensor(t).detach()
    if t.device.type == "cpu" and torch.device(device).type == "cuda":
        # through pinned memory: an asynchronous copy, no wait on the stream
        t = t.to(torch.float32).contiguous().pin_memory().to(device, non_blocking=True)
    else:
        t = t.to(device=device, dtype=torch.float32)
    if t.numel() == 1:
        return t.reshape(1).contiguous(), 0
    if t.numel() == n:
        return t.reshape(n).contiguous(), 1
    return torch.broadcast_to(t, shape).reshape(n).contiguous(), 1


def _to_device(t: torch.Tensor, device) -> torch.Tensor:
    """A host tensor on ``device`` (through pinned memory for a HIP device:
    an asynchronous copy, no wait on the stream)."""
    if torch.device(device).type == "cuda":
        return t.pin_memory().to(device, non_blocking=True)
    return t.to(device)


def lane_state(module, name: str, shape: tuple[int, ...]) -> torch.Tensor:
    """The memory ``name`` of ``module`` as a contiguous fp32 tensor of
    exactly ``shape`` (the lanes x neurons a launch covers), stored back on
    the module.  A memory with fewer batch dimensions (an unbatched state
    meeting a batched input) is broadcast into a fresh copy; one that cannot
    broadcast, or does not hold ``prod(shape)`` elements afterwards, raises
    before any kernel sees it."""
    t = orig = getattr(module, name, None)
    if t is None:
        raise RuntimeError(f"{type(module).__name__}.{name} is not initialised "
                           "(call init_net_state)")
    if tuple(t.shape) != tuple(shape):
        try:
            t = torch.broadcast_to(t, shape)
        except RuntimeError as e:
            raise RuntimeError(f"{type(module).__name__}.{name} of shape {tuple(t.shape)} does "
                               f"not match the launch shape {tuple(shape)}") from e
    t = t.contiguous()
    n = 1
    for d in shape:
        n *= int(d)
    if t.numel() != n:
        raise RuntimeError(f"{type(module).__name__}.{name} holds {t.numel()} values, the "
                           f"launch covers {n}")
    if t is not orig:
        setattr(module, name, t)
    return t


def exp_euler_coeff(tau_syn: torch.Tensor, dt: float) -> torch.Tensor:
    """k = expm1(dt*l)/l with l = -1/tau_s, evaluated by torch on the CPU in
    fp32 exactly as exp_euler_step does (btorch/models/ode.py:50 with the
    linear term of synapse.py:171-174), so the GPU never evaluates expm1 on
    the parity path."""
    ts = torch.as_tensor(tau_syn).detach().to("cpu", torch.float32)
    lin = -1.0 / ts
    return torch.expm1(dt * lin) / lin


def lif_plan(neuron, lanes: int, device, keep: list) -> _lib.RsnnPlan:
    """Neuron fields of a plan for a LIF module (state must be initialised)."""
    p = _lib.RsnnPlan()
    n = neuron.size
    shape = tuple(neuron.n_neuron)
    p.n = n
    p.lanes = lanes
    p.neuron_model = _lib.NEURON_LIF
    p.hard_reset = int(bool(neuron.hard_reset))
    p.use_refractory = int(bool(neuron._use_refractory))
    p.dt = float(environ.get("dt"))
    for field, value in (("tau", neuron.tau), ("v_th", neuron.v_threshold),
                         ("v_reset", neuron.v_reset), ("c_m", neuron.c_m)):
        vec, stride = param_vec(value, n, shape, device)
        keep.append(vec)
        setattr(p, field, vec.data_ptr())
        setattr(p, field + "_stride", stride)
    if neuron._use_refractory:
        vec, stride = param_vec(neuron.tau_ref, n, shape, device)
        keep.append(vec)
        p.tau_ref = vec.data_ptr()
        p.tau_ref_stride = stride
    if hasattr(neuron, "Iasc"):
        glif3_fields(p, neuron, lanes, device, keep)
    return p


def glif3_fields(p, neuron, lanes: int, device, keep: list) -> None:
    """GLIF3 fields: v_rest, the membrane's exponential-Euler coefficient, the
    after-spike current rates, coefficients and amplitudes as (n, n_asc), and
    the Iasc state (lanes, n, n_asc)."""
    n, shape, m = neuron.size, tuple(neuron.n_neuron), neuron.n_Iasc
    dt = float(environ.get("dt"))
    p.neuron_model = _lib.NEURON_GLIF3
    p.n_asc = m
    vec, stride = param_vec(neuron.v_rest, n, shape, device)
    keep.append(vec)
    p.v_rest, p.v_rest_stride = vec.data_ptr(), stride
    vec, stride = param_vec(exp_euler_coeff(neuron.tau, dt), n, shape, device)
    keep.append(vec)
    p.k_v, p.k_v_stride = vec.data_ptr(), stride
    k_cpu = torch.as_tensor(neuron.k).detach().to("cpu", torch.float32)
    lin = -k_cpu
    kk = torch.expm1(dt * lin) / lin
    for field, t in (("asc_k", k_cpu), ("asc_kk", kk), ("asc_amp", neuron.asc_amps)):
        t = torch.as_tensor(t).detach().to(device=device, dtype=torch.float32)
        t = torch.broadcast_to(t, shape + (m,)).reshape(n, m).contiguous()
        keep.append(t)
        setattr(p, field, t.data_ptr())
    ia = neuron.Iasc
    if ia is None or ia.numel() != lanes * n * m or ia.dtype != torch.float32:
        raise RuntimeError(f"GLIF3.Iasc is not initialised for {lanes} lanes x {n} neurons")
    if not ia.is_contiguous():
        neuron.Iasc = ia = ia.contiguous()
    keep.append(ia)
    p.iasc = ia.data_ptr()


def _cpu32(t) -> torch.Tensor:
    return torch.as_tensor(t).detach().to("cpu", torch.float32)


def tc_coefficients(neuron, dt: float) -> list[torch.Tensor]:
    """Coefficients BTX_TC_* of a TwoCompartmentGLIF, evaluated by torch on
    the CPU in fp32 with the reference's ops (two_compartment.py:204-259:
    _positive clamps, ``dt / tau``, ``1.0 + decay``)."""
    c = [None] * _lib.TC_NPARAM
    da = dt / torch.clamp(_cpu32(neuron.tau_a), min=1e-4)
    ds = dt / torch.clamp(_cpu32(neuron.tau_s), min=1e-4)
    dth = dt / torch.clamp(_cpu32(neuron.tau_th), min=1e-4)
    delta_T = torch.clamp(_cpu32(neuron.delta_T), min=0.0)
    c[_lib.TC_DA], c[_lib.TC_DA1] = da, 1.0 + da
    c[_lib.TC_DS], c[_lib.TC_DS1] = ds, 1.0 + ds
    c[_lib.TC_DTH1] = 1.0 + dth
    c[_lib.TC_R_S] = torch.clamp(_cpu32(neuron.R_s), min=1e-4)
    c[_lib.TC_E_L] = _cpu32(neuron.E_L)
    c[_lib.TC_DT] = delta_T
    c[_lib.TC_DTC] = torch.clamp(delta_T, min=1e-4)
    c[_lib.TC_W_CA] = _cpu32(neuron.w_Ca)
    c[_lib.TC_THETA_CA] = _cpu32(neuron.theta_Ca)
    c[_lib.TC_W_SA] = _cpu32(neuron.w_sa)
    c[_lib.TC_W_AS] = _cpu32(neuron.w_as)
    c[_lib.TC_DELTA_TH] = _cpu32(neuron.delta_th)
    return c


def tc_plan(neuron, lanes: int, device, keep: list) -> _lib.RsnnPlan:
    """Neuron fields of a plan for a TwoCompartmentGLIF (state initialised)."""
    p = _lib.RsnnPlan()
    n, shape = neuron.size, tuple(neuron.n_neuron)
    p.n = n
    p.lanes = lanes
    p.neuron_model = _lib.NEURON_TWO_COMPARTMENT
    p.hard_reset = 1
    p.use_refractory = 0
    p.dt = float(environ.get("dt"))
    for field, value in (("v_th", neuron.v_threshold), ("v_reset", neuron.v_reset)):
        vec, stride = param_vec(value, n, shape, device)
        keep.append(vec)
        setattr(p, field, vec.data_ptr())
        setattr(p, field + "_stride", stride)
    for k, t in enumerate(tc_coefficients(neuron, p.dt)):
        vec, stride = param_vec(t, n, shape, device)
        keep.append(vec)
        p.tc[k] = vec.data_ptr()
        p.tc_stride[k] = stride
    for name in ("i_a", "i_bap", "theta_th"):
        t = getattr(neuron, name)
        if t is None or t.numel() != lanes * n or t.dtype != torch.float32:
            raise RuntimeError(f"TwoCompartmentGLIF.{name} is not initialised for "
                               f"{lanes} lanes x {n} neurons")
        if not t.is_contiguous():
            setattr(neuron, name, t.contiguous())
            t = getattr(neuron, name)
        keep.append(t)
        setattr(p, name, t.data_ptr())
    return p


def neuron_plan(neuron, lanes: int, device, keep: list) -> _lib.RsnnPlan:
    """Neuron fields of a plan for any neuron the HIP step runs."""
    if hasattr(neuron, "i_bap"):
        return tc_plan(neuron, lanes, device, keep)
    return lif_plan(neuron, lanes, device, keep)


def mixed_plan(state, lanes: int, device, keep: list) -> _lib.RsnnPlan:
    """Neuron fields of a plan for a multi-group MixedNeuronPopulation (the
    engine's ``MixedState``): every parameter laid out per neuron along the
    population axis (a group's scalar repeated over its range, values the
    group's model does not use left at a neutral 0 or 1), a model code per
    neuron, and the population-wide state arrays.  Exponential-Euler and
    two-compartment coefficients come from torch on the CPU, with the ops of
    the single-model plans above."""
    from .neurons.glif import GLIF3
    from .neurons.two_compartment import TwoCompartmentGLIF

    pop = state.pop
    lay = pop.layout
    n = pop.size
    dt = float(environ.get("dt"))
    m = max(1, state.n_asc)
    vec = {k: torch.zeros(n, dtype=torch.float32) for k in
           ("tau", "v_th", "v_reset", "c_m", "tau_ref", "v_rest", "k_v")}
    vec["tau"].fill_(1.0)
    vec["c_m"].fill_(1.0)
    asc = {k: torch.zeros(n, m, dtype=torch.float32) for k in ("asc_k", "asc_kk", "asc_amp")}
    tcv = [torch.zeros(n, dtype=torch.float32) for _ in range(_lib.TC_NPARAM)]
    kinds = torch.zeros(n, dtype=torch.uint8)
    models = 0

    def put(dst, sl, val):
        c = sl.stop - sl.start
        dst[sl] = torch.as_tensor(val).detach().to("cpu", torch.float32).reshape(-1).expand(c)

    for g, mod in enumerate(pop.groups()):
        sl = lay.span(g)
        c = sl.stop - sl.start
        put(vec["v_th"], sl, mod.v_threshold)
        put(vec["v_reset"], sl, mod.v_reset)
        if isinstance(mod, TwoCompartmentGLIF):
            code = _lib.NEURON_TWO_COMPARTMENT
            for k, t in enumerate(tc_coefficients(mod, dt)):
                put(tcv[k], sl, t)
        else:
            code = _lib.NEURON_GLIF3 if isinstance(mod, GLIF3) else _lib.NEURON_LIF
            put(vec["tau"], sl, mod.tau)
            put(vec["c_m"], sl, mod.c_m)
            if getattr(mod, "_use_refractory", False):
                put(vec["tau_ref"], sl, mod.tau_ref)
            if mod.hard_reset:
                code |= _lib.MIXED_HARD_RESET
            if isinstance(mod, GLIF3):
                put(vec["v_rest"], sl, mod.v_rest)
                put(vec["k_v"], sl, exp_euler_coeff(mod.tau, dt))
                k_cpu = _cpu32(mod.k)
                lin = -k_cpu
                kk = torch.expm1(dt * lin) / lin
                for key, t in (("asc_k", k_cpu), ("asc_kk", kk), ("asc_amp", mod.asc_amps)):
                    asc[key][sl] = torch.broadcast_to(_cpu32(t), (c, m))
        kinds[sl] = code
        models |= 1 << (code & 3)
    p = _lib.RsnnPlan()
    p.n = n
    p.lanes = lanes
    p.neuron_model = _lib.NEURON_MIXED
    p.mixed_models = models
    p.hard_reset = 0
    p.use_refractory = int(bool(state._use_refractory))
    p.dt = dt

    def dev_vec(t):
        d, stride = param_vec(t, n, (n,), device)
        keep.append(d)
        return d.data_ptr(), stride

    for field in ("tau", "v_th", "v_reset", "c_m"):
        ptr, stride = dev_vec(vec[field])
        setattr(p, field, ptr)
        setattr(p, field + "_stride", stride)
    if p.use_refractory:
        p.tau_ref, p.tau_ref_stride = dev_vec(vec["tau_ref"])
    if models & (1 << _lib.NEURON_GLIF3):
        p.n_asc = state.n_asc
        p.v_rest, p.v_rest_stride = dev_vec(vec["v_rest"])
        p.k_v, p.k_v_stride = dev_vec(vec["k_v"])
        for key in ("asc_k", "asc_kk", "asc_amp"):
            d = _to_device(asc[key][:, : state.n_asc].contiguous(), device)
            keep.append(d)
            setattr(p, key, d.data_ptr())
    if models & (1 << _lib.NEURON_TWO_COMPARTMENT):
        for k in range(_lib.TC_NPARAM):
            p.tc[k], p.tc_stride[k] = dev_vec(tcv[k])
    kd = _to_device(kinds, device)
    keep.append(kd)
    p.neuron_kinds = kd.data_ptr()
    return p
