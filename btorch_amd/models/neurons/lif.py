"""Leaky integrate-and-fire neuron (btorch/models/neurons/lif.py:29-187).

dV/dt = -(V - V_reset)/tau + I/c_m, explicit Euler with the environment dt;
fire when (V - V_th)/(V_th - V_reset) >= 0; soft or hard reset; optional
refractory counter in ms.  ``single_step_forward`` runs one step in the HIP
library (``btx_lif_step``); inside ``RecurrentNN`` the step is fused with
the synapse and the sparse product instead.  A standalone step that needs
gradients returns new state tensors and differentiates through the
surrogate (for the input, the membrane and tau).
"""

from __future__ import annotations

from collections.abc import Callable, Sequence

import torch

from ... import _lib
from .. import _plan, environ
from ..base import BaseNode
from ..surrogate import Sigmoid


def _reduce_to(g: torch.Tensor, like: torch.Tensor) -> torch.Tensor:
    """Sum a gradient over the dimensions a parameter was broadcast along."""
    if like.numel() == 1:
        return g.sum().reshape(like.shape)
    lead = g.dim() - like.dim()
    return g.sum(dim=tuple(range(lead))).reshape(like.shape)


class _LIFStep(torch.autograd.Function):
    """One LIF step by btx_lif_step on copies of the state, with the
    surrogate-gradient backward of lif.py:144-187 in torch ops:
        v1 = v + dt*((-(v - vr))/tau + x/cm),  s = H((v1 - vth)/(vth - vr))*mask
        v' = v1 - (vth - vr)*s  (hard: v1 - (v1 - vr)*s)
    the refractory mask is detached (lif.py:165-167) and the reset carries
    gradient unless detach_reset."""

    @staticmethod
    def forward(ctx, x, v, tau, neuron):
        z, v_new, ref_in, v_pre = neuron._launch(x, v.detach().clone())
        ctx.neuron = neuron
        ctx.dt = float(environ.get("dt"))
        ctx.save_for_backward(v.detach(), v_pre, z, tau)
        ctx.ref_in = ref_in
        ctx.x_shape = x.shape
        ctx.mark_non_differentiable(v_pre)
        return z, v_new, v_pre

    @staticmethod
    def backward(ctx, gz, gv_new, _gv_pre):
        v, v1, z, tau = ctx.saved_tensors
        nrn, dt = ctx.neuron, ctx.dt
        vth = nrn.v_threshold.detach()
        vr = nrn.v_reset.detach()
        cm = nrn.c_m.detach()
        gz = torch.zeros_like(z) if gz is None else gz
        gvn = torch.zeros_like(v) if gv_new is None else gv_new
        gs = gz
        if nrn.hard_reset:
            dvdv1 = 1.0 - z
            if not nrn.detach_reset:
                gs = gs + gvn * (-(v1 - vr))
        else:
            dvdv1 = 1.0
            if not nrn.detach_reset:
                gs = gs + gvn * (-(vth - vr))
        if ctx.ref_in is not None:
            gs = gs * (ctx.ref_in == 0).to(gs.dtype)
        sg = nrn.surrogate_function
        gu = sg.derivative((v1 - vth) / (vth - vr), gs, sg.damping_factor)
        gv1 = gvn * dvdv1 + gu / (vth - vr)
        td = tau.detach()
        gv = gv1 * (1.0 - dt / td)
        gx = (gv1 * (dt / cm)).sum_to_size(ctx.x_shape)
        gtau = None
        if ctx.needs_input_grad[2]:
            gtau = _reduce_to(gv1 * dt * (v - vr) / (td * td), tau)
        return gx, gv, gtau, None


class LIF(BaseNode):
    """LIF neuron; constructor and buffers as in the reference (see module doc)."""

    def __init__(self, n_neuron: int | Sequence[int], v_threshold=1.0, v_reset=0.0, c_m=1.0,
                 tau=20.0, tau_ref=None, trainable_param=frozenset(),
                 surrogate_function: Callable | None = None, detach_reset: bool = False,
                 hard_reset: bool = False, pre_spike_v: bool = False, step_mode: str = "s",
                 backend: str = "torch", device=None, dtype=None):
        super().__init__(n_neuron=n_neuron, v_threshold=v_threshold, v_reset=v_reset,
                         trainable_param=trainable_param,
                         surrogate_function=surrogate_function or Sigmoid(),
                         detach_reset=detach_reset, hard_reset=hard_reset,
                         pre_spike_v=pre_spike_v, step_mode=step_mode, backend=backend,
                         device=device, dtype=dtype)
        fk = {"device": device, "dtype": dtype}
        self.def_param("c_m", c_m, sizes=self.n_neuron, trainable_param=self.trainable_param, **fk)
        self.def_param("tau", tau, sizes=self.n_neuron, trainable_param=self.trainable_param, **fk)
        self._use_refractory = tau_ref is not None
        if self._use_refractory:
            self.def_param("tau_ref", tau_ref, sizes=self.n_neuron,
                           trainable_param=self.trainable_param, **fk)
            self.register_memory("refractory", 0.0, self.n_neuron)
        else:
            self.tau_ref = None

    def _needs_grad(self, x: torch.Tensor) -> bool:
        if not torch.is_grad_enabled():
            return False
        if x.requires_grad or (self.v is not None and self.v.requires_grad):
            return True
        return any(p.requires_grad for p in self.parameters(recurse=False))

    def _launch(self, x: torch.Tensor, v: torch.Tensor, want_pre: bool = True):
        """btx_lif_step on state ``v`` (updated in place) and the module's
        refractory counter; returns (z, v, refractory before the step,
        membrane before reset)."""
        lead = tuple(v.shape[: v.ndim - len(self.n_neuron)])
        x = torch.broadcast_to(x.detach(), v.shape).contiguous()
        lanes = 1
        for d in lead:
            lanes *= d
        keep: list = []
        plan = _plan.lif_plan(self, lanes, x.device, keep)
        plan.v = v.data_ptr()
        ref_in = None
        if self._use_refractory:
            # the counter covers exactly the lanes the launch updates (an
            # unbatched counter meeting a batched input gets its own copy)
            ref = _plan.lane_state(self, "refractory", tuple(v.shape))
            ref_in = ref.clone()
            plan.refractory = ref.data_ptr()
        z = torch.empty_like(x)
        v_pre = torch.empty_like(x) if want_pre else None
        _lib.check(_lib.load().btx_lif_step(plan, x.data_ptr(), z.data_ptr(),
                                            None if v_pre is None else v_pre.data_ptr(),
                                            _lib.stream_ptr()), "btx_lif_step")
        return z, v, ref_in, v_pre

    def single_step_forward(self, x: torch.Tensor) -> torch.Tensor:
        _lib.require_hip(x, self.v)
        if x.dtype != torch.float32 or self.v.dtype != torch.float32:
            raise TypeError("LIF on the HIP backend supports float32 state only")
        if self._needs_grad(x):
            if not getattr(self.surrogate_function, "spiking", True):
                raise NotImplementedError("the HIP step runs the spiking (Heaviside) forward")
            for name in ("v_threshold", "v_reset", "c_m", "tau_ref"):
                t = getattr(self, name, None)
                if isinstance(t, torch.Tensor) and t.requires_grad:
                    raise NotImplementedError(f"gradient w.r.t. LIF.{name} is not provided")
            v_in = torch.broadcast_to(self.v, torch.broadcast_shapes(self.v.shape, x.shape))
            z, v_new, v_pre = _LIFStep.apply(x, v_in.contiguous(), self.tau, self)
            self.v = v_new
            if self.pre_spike_v:
                self.v_pre_spike = v_pre
            return z
        v = self.v.contiguous()
        self.v = v
        z, _, _, v_pre = self._launch(x, v, want_pre=self.pre_spike_v)
        if self.pre_spike_v:
            self.v_pre_spike = v_pre
        return z

    def extra_repr(self):
        ref = self._format_repr_value(self.tau_ref) if self._use_refractory else "None"
        return (f"c_m={self._format_repr_value(self.c_m)}, tau={self._format_repr_value(self.tau)}, "
                f"tau_ref={ref}, " + super().extra_repr())


class IF(LIF):
    """Integrate-and-fire: LIF without leak (tau -> infinity)."""

    def __init__(self, *args, **kwargs):
        kwargs["tau"] = float("inf")
        super().__init__(*args, **kwargs)
