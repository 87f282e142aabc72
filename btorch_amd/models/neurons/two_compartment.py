"""Two-compartment current-based GLIF neuron with apical plateaus
(btorch/models/neurons/two_compartment.py:27-407).

Per step, with A the apical input and I the somatic input:
  i_a   <- (i_a + d_a*((A + w_Ca*sigmoid(i_a - theta_Ca)) + i_bap)) / (1 + d_a)
  thr    = v_threshold + theta_th
  v_pre  = (v + d_s*((E_L + R_s*(I + w_as*i_a)) + ex)) / (1 + d_s)
           ex = delta_T*exp(min((v - thr)/delta_T, 10)) when delta_T > 0, else 0
  s      = H(v_pre - thr) ; v <- v_pre - (v_pre - v_reset)*s
  i_bap <- w_sa*s ; theta_th <- theta_th/(1 + d_th) + delta_th*s
with d_x = dt / max(tau_x, 1e-4), R_s and delta_T clamped as the reference
does.  The coefficients are evaluated by torch on the CPU; the step runs in
the HIP library (``btx_lif_step`` with the two-compartment neuron model) and,
inside a recurrent cell, fused with the synapse and the sparse product.
"""

from __future__ import annotations

from collections.abc import Callable, Sequence

import torch

from ... import _lib
from .. import _plan, environ
from ..base import MemoryModule, ParamBufferMixin, _broadcastable, normalize_n_neuron
from ..surrogate import ATan

_PARAMS = ("v_threshold", "v_reset", "tau_s", "R_s", "E_L", "tau_a", "tau_th", "delta_th",
           "delta_T", "w_Ca", "theta_Ca", "w_sa", "w_as")


class _TCStep(torch.autograd.Function):
    """One two-compartment step by btx_lif_step, with the surrogate-gradient
    backward of two_compartment.py:208-306 in torch ops (inputs and the
    four states; the plateau sigmoid and the clamped exponential initiation
    differentiated as written, detach_reset as in the reference)."""

    @staticmethod
    def forward(ctx, soma, apical, v, i_a, i_bap, th, neuron):
        sc = soma.detach().to(torch.float32).contiguous()
        ap = apical.detach().to(torch.float32).contiguous()
        z, v_pre = neuron._launch(sc, ap, True)
        ctx.neuron, ctx.dt = neuron, float(environ.get("dt"))
        ctx.shapes = (soma.shape, apical.shape)
        ctx.save_for_backward(v.detach(), i_a.detach(), th.detach(), v_pre, z)
        ctx.mark_non_differentiable(v_pre)
        return z, neuron.v, neuron.i_a, neuron.i_bap, neuron.theta_th, v_pre

    @staticmethod
    def backward(ctx, gz, gv_n, gia_n, gbap_n, gth_n, _gv_pre):
        v, i_a, th, v_pre, z = ctx.saved_tensors
        nrn, dev = ctx.neuron, v.device
        c = [t.to(dev) for t in _plan.tc_coefficients(nrn, ctx.dt)]
        da, da1, ds, ds1, dth1, r_s = (c[_lib.TC_DA], c[_lib.TC_DA1], c[_lib.TC_DS],
                                       c[_lib.TC_DS1], c[_lib.TC_DTH1], c[_lib.TC_R_S])
        d_t, d_tc, w_ca, th_ca = c[_lib.TC_DT], c[_lib.TC_DTC], c[_lib.TC_W_CA], \
            c[_lib.TC_THETA_CA]
        w_sa, w_as, delta_th = c[_lib.TC_W_SA], c[_lib.TC_W_AS], c[_lib.TC_DELTA_TH]
        vth, vr = nrn.v_threshold.detach(), nrn.v_reset.detach()
        zero = torch.zeros_like(v)
        gz, gv_n, gia_n, gbap_n, gth_n = (zero if g is None else g
                                          for g in (gz, gv_n, gia_n, gbap_n, gth_n))
        thr = vth + th
        # fire, reset, bAP and threshold updates
        gs = gz + gbap_n * w_sa
        if not nrn.detach_reset:
            gs = gs + gv_n * (-(v_pre - vr)) + gth_n * delta_th
        sg = nrn.surrogate_function
        gu = sg.derivative(v_pre - thr, gs, sg.damping_factor)
        gv_pre = gv_n * (1.0 - z) + gu
        g_thr = -gu
        # soma: v_pre = (v + DS*((E_L + R_s*(I + w_as*i_a')) + ex))/DS1
        g_num = gv_pre / ds1
        g_total = g_num * ds * r_s
        arg = (v - thr) / d_tc
        dex = torch.where((d_t > 0.0) & (arg <= 10.0),
                          d_t * torch.exp(torch.clamp(arg, max=10.0)) / d_tc, zero)
        gv = g_num + g_num * ds * dex
        g_thr = g_thr - g_num * ds * dex
        # apical: i_a' = (i_a + DA*((A + w_Ca*sigmoid(i_a - theta_Ca)) + i_bap))/DA1
        g_d = (gia_n + g_total * w_as) / da1
        sgm = torch.sigmoid(i_a - th_ca)
        g_ia = g_d * (1.0 + da * w_ca * sgm * (1.0 - sgm))
        g_ap = g_d * da
        g_th = g_thr + gth_n / dth1
        soma_shape, ap_shape = ctx.shapes
        return (g_total.sum_to_size(soma_shape), g_ap.sum_to_size(ap_shape), gv, g_ia, g_ap,
                g_th, None)


class TwoCompartmentGLIF(ParamBufferMixin, MemoryModule):
    """Soma + apical current neuron; constructor, parameters, memories
    (``v``, ``i_a``, ``i_bap``, ``theta_th``) and return values as in the
    reference (two_compartment.py:127-182, 261-307)."""

    def __init__(self, n_neuron: int | Sequence[int], *, v_threshold=-50.0, v_reset=-65.0,
                 tau_s=20.0, R_s=1.0, E_L=-70.0, tau_a=100.0, tau_th=50.0, delta_th=0.0,
                 delta_T=0.0, w_Ca=0.0, theta_Ca=1.0, w_sa=0.5, w_as=1.0,
                 trainable_param=frozenset(), surrogate_function: Callable | None = None,
                 detach_reset: bool = False, step_mode: str = "s", backend: str = "torch",
                 device=None, dtype=None):
        super().__init__()
        self.n_neuron, self.size = normalize_n_neuron(n_neuron)
        self.trainable_param = set(trainable_param)
        self.surrogate_function = surrogate_function if surrogate_function is not None \
            else ATan()
        self.detach_reset = detach_reset
        self.step_mode = step_mode
        self.backend = backend
        self.register_memory("v", v_reset, self.n_neuron)
        self.register_memory("i_a", 0.0, self.n_neuron)
        self.register_memory("i_bap", 0.0, self.n_neuron)
        self.register_memory("theta_th", 0.0, self.n_neuron)
        values = (v_threshold, v_reset, tau_s, R_s, E_L, tau_a, tau_th, delta_th, delta_T,
                  w_Ca, theta_Ca, w_sa, w_as)
        for name, value in zip(_PARAMS, values):
            self.def_param(name, value, sizes=self.n_neuron,
                           trainable_param=self.trainable_param, device=device, dtype=dtype)

    # -- inputs (two_compartment.py:184-201) ----------------------------------
    def _prepare_step_tensor(self, x, *, name: str) -> torch.Tensor:
        target = self.v
        x = torch.as_tensor(x, device=target.device, dtype=target.dtype)
        if x.shape == target.shape:
            return x
        if x.numel() == 1:
            return torch.full_like(target, x.reshape(()))
        if _broadcastable(x.shape, target.shape):
            return torch.broadcast_to(x, target.shape)
        raise RuntimeError(
            f"{name} shape mismatch. Expected a tensor broadcastable to "
            f"{tuple(target.shape)}, got {tuple(x.shape)}.")

    def _lanes(self) -> int:
        lead = tuple(self.v.shape[: self.v.ndim - len(self.n_neuron)])
        lanes = 1
        for d in lead:
            lanes *= d
        return lanes

    def _launch(self, soma, apical, want_pre: bool):
        """btx_lif_step on fresh copies of the state (assigned to the module);
        returns (z, v_pre or None)."""
        for name in ("v", "i_a", "i_bap", "theta_th"):
            setattr(self, name, getattr(self, name).detach().clone(
                memory_format=torch.contiguous_format))
        keep: list = []
        plan = _plan.tc_plan(self, self._lanes(), soma.device, keep)
        plan.v = self.v.data_ptr()
        plan.x_apical = apical.data_ptr()
        z = torch.empty_like(soma)
        v_pre = torch.empty_like(soma) if want_pre else None
        _lib.check(_lib.load().btx_lif_step(plan, soma.data_ptr(), z.data_ptr(),
                                            None if v_pre is None else v_pre.data_ptr(),
                                            _lib.stream_ptr()), "btx_lif_step")
        return z, v_pre

    def single_step_forward(self, i_soma, i_apical=None, *, return_state: bool = False):
        """(spike, v), or (spike, v, states) with ``return_state``; the states
        are v, v_pre_spike, i_a, i_bap, theta_th and v_threshold_eff."""
        _lib.require_hip(self.v)
        soma = self._prepare_step_tensor(i_soma, name="i_soma")
        apical = (torch.zeros_like(soma) if i_apical is None
                  else self._prepare_step_tensor(i_apical, name="i_apical"))
        if self.v.dtype != torch.float32:
            raise TypeError("TwoCompartmentGLIF on the HIP backend supports float32 state only")
        threshold_eff = self.v_threshold + self.theta_th if return_state else None
        states = (self.v, self.i_a, self.i_bap, self.theta_th)
        if torch.is_grad_enabled() and (soma.requires_grad or apical.requires_grad or any(
                t.requires_grad for t in states)):
            for name, p in self.named_parameters(recurse=False):
                if p.requires_grad:
                    raise NotImplementedError(
                        f"gradient w.r.t. TwoCompartmentGLIF.{name} is not provided")
            z, v, i_a, i_bap, th, v_pre = _TCStep.apply(soma, apical, *states, self)
            self.v, self.i_a, self.i_bap, self.theta_th = v, i_a, i_bap, th
        else:
            # the reference assigns new state tensors each step; the kernel
            # updates copies in place, so tensors returned earlier keep their values
            z, v_pre = self._launch(soma.to(torch.float32).contiguous(),
                                    apical.to(torch.float32).contiguous(), return_state)
        if not return_state:
            return z, self.v
        return z, self.v, {"v": self.v, "v_pre_spike": v_pre, "i_a": self.i_a,
                           "i_bap": self.i_bap, "theta_th": self.theta_th,
                           "v_threshold_eff": threshold_eff}

    def multi_step_forward(self, i_soma_seq, i_apical_seq=None, *, return_state: bool = False):
        """Loop of single steps with stacked outputs (two_compartment.py:309-371)."""
        spikes, vs, states = [], [], []
        for t in range(i_soma_seq.shape[0]):
            step_apical = None if i_apical_seq is None else i_apical_seq[t]
            out = self.single_step_forward(i_soma_seq[t], step_apical,
                                           return_state=return_state)
            spikes.append(out[0])
            vs.append(out[1])
            if return_state:
                states.append(out[2])
        spike_seq, v_seq = torch.stack(spikes), torch.stack(vs)
        if not return_state:
            return spike_seq, v_seq
        stacked = {k: torch.stack([s[k] for s in states]) for k in states[0]}
        return spike_seq, v_seq, stacked

    def forward(self, i_soma, i_apical=None, *, return_state: bool = False):
        if self.step_mode == "m":
            return self.multi_step_forward(i_soma, i_apical, return_state=return_state)
        return self.single_step_forward(i_soma, i_apical, return_state=return_state)

    def extra_repr(self) -> str:
        return (f"n_neuron={self.n_neuron}, step_mode={self.step_mode}, "
                f"backend={self.backend}, "
                f"surrogate={self.surrogate_function.__class__.__name__}")


__all__ = ["TwoCompartmentGLIF"]
