"""Host-side assembly of ``btx_rsnn_plan`` fields from module parameters."""

from __future__ import annotations

import torch

from .. import _lib
from . import environ


def param_vec(t, n: int, shape: tuple[int, ...], device) -> tuple[torch.Tensor, int]:
    """(fp32 contiguous tensor, stride) for a scalar or per-neuron parameter."""
    t = torch.as_tensor(t)
    t = t.detach().to(device=device, dtype=torch.float32)
    if t.numel() == 1:
        return t.reshape(1).contiguous(), 0
    if t.numel() == n:
        return t.reshape(n).contiguous(), 1
    return torch.broadcast_to(t, shape).reshape(n).contiguous(), 1


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
