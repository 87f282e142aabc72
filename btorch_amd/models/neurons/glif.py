"""GLIF3: leaky integrate-and-fire with after-spike currents
(btorch/models/neurons/glif.py:66-342; Teeter et al. 2018).

Per step, charge -> adaptation -> fire -> reset (base.py:1108-1116):
  v    <- v + expm1(dt*l)/l * ((-(v - v_rest))/tau + (x + sum_j Iasc_j)/c_m),  l = -1/tau
  Iasc <- Iasc + expm1(dt*(-k))/(-k) * ((-k) * Iasc)
  s    = H((v - v_th)/(v_th - v_reset)) * (refractory == 0)
  v    <- soft/hard reset ; Iasc <- Iasc + asc_amps * s ; refractory <- relu(ref + s*tau_ref - dt)
The exponential-Euler coefficients are evaluated by torch on the CPU in fp32
(the same ops as ode.py:50); the step runs in the HIP library.  Inside
``RecurrentNN`` the step is fused with the synapse and the sparse product.
A standalone step that needs gradients returns new state tensors and
differentiates through the surrogate (input, membrane, Iasc and tau).
"""

from __future__ import annotations

from collections.abc import Callable, Sequence
from numbers import Number

import torch

from ... import _lib
from .. import _plan, environ
from ..base import BaseNode
from ..surrogate import Erf


class _GLIF3Step(torch.autograd.Function):
    """One GLIF3 step by btx_lif_step on copies of v and Iasc, with the
    surrogate-gradient backward of the step above in torch ops (the
    refractory mask detached, the reset carrying gradient unless
    detach_reset, tau entering through 1/tau and the exponential-Euler
    coefficient k_v = expm1(dt*l)/l, l = -1/tau)."""

    @staticmethod
    def forward(ctx, x, v, iasc, tau, neuron):
        xb = torch.broadcast_to(x.detach(), v.shape).contiguous()
        v_new = v.detach().clone()
        ia_new = iasc.detach().clone()
        lanes = v_new.numel() // neuron.size
        keep: list = []
        plan = _plan.lif_plan(neuron, lanes, xb.device, keep)
        plan.v = v_new.data_ptr()
        plan.iasc = ia_new.data_ptr()
        ref_in = None
        if neuron._use_refractory:
            ref = _plan.lane_state(neuron, "refractory", tuple(v_new.shape))
            ref_in = ref.clone()
            plan.refractory = ref.data_ptr()
        z = torch.empty_like(xb)
        v_pre = torch.empty_like(xb)
        _lib.check(_lib.load().btx_lif_step(plan, xb.data_ptr(), z.data_ptr(), v_pre.data_ptr(),
                                            _lib.stream_ptr()), "btx_lif_step")
        ctx.neuron, ctx.dt, ctx.ref_in, ctx.x_shape = neuron, float(environ.get("dt")), ref_in, \
            x.shape
        ctx.save_for_backward(xb, v.detach(), iasc.detach(), v_pre, z, tau)
        ctx.mark_non_differentiable(v_pre)
        return z, v_new, ia_new, v_pre

    @staticmethod
    def backward(ctx, gz, gv_new, gia_new, _gv_pre):
        x, v, ia, v1, z, tau = ctx.saved_tensors
        nrn, dt = ctx.neuron, ctx.dt
        vth, vr = nrn.v_threshold.detach(), nrn.v_reset.detach()
        vrest, cm = nrn.v_rest.detach(), nrn.c_m.detach()
        k, amp = nrn.k.detach(), nrn.asc_amps.detach()
        td = tau.detach()
        gz = torch.zeros_like(z) if gz is None else gz
        gvn = torch.zeros_like(v) if gv_new is None else gv_new
        gia_n = torch.zeros_like(ia) if gia_new is None else gia_new
        # fire and reset: Iasc += amp*s also feeds the spike's gradient
        gs = gz + (gia_n * amp).sum(-1)
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
        # charge: v1 = v + kv*dv, dv = (-(v - vrest))/tau + (x + sum Iasc)/cm
        lin = -1.0 / td
        kv = torch.expm1(dt * lin) / lin
        gdv = gv1 * kv
        gv = gv1 + gdv * lin
        gI = gdv / cm
        gx = gI.sum_to_size(ctx.x_shape)
        # adaptation: Iasc' = Iasc + kk*((-k)*Iasc), kk = expm1(dt*(-k))/(-k)
        kk = torch.expm1(dt * (-k)) / (-k)
        gia = gia_n * (1.0 + kk * (-k)) + gI.unsqueeze(-1)
        gtau = None
        if ctx.needs_input_grad[3]:
            dv = (-(v - vrest)) / td + (x + ia.sum(-1)) / cm
            dkv_dl = (dt * torch.exp(dt * lin) * lin - torch.expm1(dt * lin)) / (lin * lin)
            dv1_dtau = dkv_dl / (td * td) * dv + kv * (v - vrest) / (td * td)
            g = gv1 * dv1_dtau
            gtau = g.sum().reshape(tau.shape) if tau.numel() == 1 else \
                g.sum(dim=tuple(range(g.dim() - tau.dim()))).reshape(tau.shape)
        return gx, gv, gia, gtau, None


class GLIF3(BaseNode):
    """GLIF3 neuron; constructor, parameters and memories as in the reference."""

    def __init__(self, n_neuron: int | Sequence[int], v_threshold=-50.0, v_reset=-70.0,
                 v_rest=None, c_m=0.05, tau=20.0, k=(0.2,), asc_amps=(0.0,), tau_ref=0.0,
                 trainable_param=frozenset(), surrogate_function: Callable | None = None,
                 detach_reset: bool = False, hard_reset: bool = False,
                 pre_spike_v: bool = False, step_mode: str = "s", backend: str = "torch",
                 device=None, dtype=None):
        super().__init__(n_neuron=n_neuron, v_threshold=v_threshold, v_reset=v_reset,
                         trainable_param=trainable_param,
                         surrogate_function=surrogate_function or Erf(alpha=4.0,
                                                                      damping_factor=0.5),
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
        if v_rest is not None:
            self.def_param("_v_rest", v_rest, sizes=self.n_neuron,
                           trainable_param=self.trainable_param, **fk)
        else:
            self._v_rest = None
        k = [k] if isinstance(k, Number) else k
        asc_amps = [asc_amps] if isinstance(asc_amps, Number) else asc_amps
        n_asc = int(torch.as_tensor(k).shape[-1]) if torch.as_tensor(k).ndim else 1
        n_asc = max(n_asc, int(torch.as_tensor(asc_amps).shape[-1])
                    if torch.as_tensor(asc_amps).ndim else 1)
        self.n_Iasc = n_asc
        sizes = tuple(self.n_neuron) + (n_asc,)
        self.def_param("k", k, sizes=sizes, trainable_param=self.trainable_param,
                       normalize_to_sizes=True, **fk)
        self.def_param("asc_amps", asc_amps, sizes=sizes, trainable_param=self.trainable_param,
                       normalize_to_sizes=True, **fk)
        self.register_memory("Iasc", [0.0] * n_asc, sizes)

    @property
    def v_rest(self) -> torch.Tensor:
        """Resting potential; v_reset unless given (glif.py:239-262)."""
        return self.v_reset if self._v_rest is None else self._v_rest

    def single_step_forward(self, x: torch.Tensor) -> torch.Tensor:
        _lib.require_hip(x, self.v)
        if x.dtype != torch.float32 or self.v.dtype != torch.float32:
            raise TypeError("GLIF3 on the HIP backend supports float32 state only")
        if torch.is_grad_enabled() and (
                x.requires_grad or self.v.requires_grad or self.Iasc.requires_grad
                or any(p.requires_grad for p in self.parameters(recurse=False))):
            for name, p in self.named_parameters(recurse=False):
                if p.requires_grad and name != "tau":
                    raise NotImplementedError(f"gradient w.r.t. GLIF3.{name} is not provided")
            if not getattr(self.surrogate_function, "spiking", True):
                raise NotImplementedError("the HIP step runs the spiking (Heaviside) forward")
            shape = torch.broadcast_shapes(self.v.shape, x.shape)
            v_in = torch.broadcast_to(self.v, shape).contiguous()
            ia_in = torch.broadcast_to(self.Iasc, shape + (self.n_Iasc,)).contiguous()
            z, v_new, ia_new, v_pre = _GLIF3Step.apply(x, v_in, ia_in, self.tau, self)
            self.v, self.Iasc = v_new, ia_new
            if self.pre_spike_v:
                self.v_pre_spike = v_pre
            return z
        lead = tuple(self.v.shape[: self.v.ndim - len(self.n_neuron)])
        x = torch.broadcast_to(x, self.v.shape).contiguous()
        lanes = 1
        for d in lead:
            lanes *= d
        keep: list = []
        plan = _plan.lif_plan(self, lanes, x.device, keep)
        self.v = self.v.contiguous()
        plan.v = self.v.data_ptr()
        if self._use_refractory:
            plan.refractory = _plan.lane_state(self, "refractory", tuple(self.v.shape)).data_ptr()
        z = torch.empty_like(x)
        v_pre = torch.empty_like(x) if self.pre_spike_v else None
        _lib.check(_lib.load().btx_lif_step(plan, x.data_ptr(), z.data_ptr(),
                                            None if v_pre is None else v_pre.data_ptr(),
                                            _lib.stream_ptr()), "btx_lif_step")
        if self.pre_spike_v:
            self.v_pre_spike = v_pre
        return z

    def extra_repr(self):
        return (f"n_Iasc={self.n_Iasc}, c_m={self._format_repr_value(self.c_m)}, "
                f"tau={self._format_repr_value(self.tau)}, " + super().extra_repr())
