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
from .. import _plan
from ..base import MemoryModule, ParamBufferMixin, _broadcastable, normalize_n_neuron
from ..surrogate import ATan

_PARAMS = ("v_threshold", "v_reset", "tau_s", "R_s", "E_L", "tau_a", "tau_th", "delta_th",
           "delta_T", "w_Ca", "theta_Ca", "w_sa", "w_as")


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

    def single_step_forward(self, i_soma, i_apical=None, *, return_state: bool = False):
        """(spike, v), or (spike, v, states) with ``return_state``; the states
        are v, v_pre_spike, i_a, i_bap, theta_th and v_threshold_eff."""
        _lib.require_hip(self.v)
        soma = self._prepare_step_tensor(i_soma, name="i_soma")
        apical = (torch.zeros_like(soma) if i_apical is None
                  else self._prepare_step_tensor(i_apical, name="i_apical"))
        if torch.is_grad_enabled() and (soma.requires_grad or apical.requires_grad or any(
                p.requires_grad for p in self.parameters(recurse=False))):
            raise NotImplementedError(
                "the two-compartment step on the HIP backend is forward-only")
        if self.v.dtype != torch.float32:
            raise TypeError("TwoCompartmentGLIF on the HIP backend supports float32 state only")
        soma = soma.to(torch.float32).contiguous()
        apical = apical.to(torch.float32).contiguous()
        threshold_eff = self.v_threshold + self.theta_th if return_state else None
        # the reference assigns new state tensors each step; the kernel
        # updates copies in place, so tensors returned earlier keep their values
        for name in ("v", "i_a", "i_bap", "theta_th"):
            setattr(self, name, getattr(self, name).clone(memory_format=torch.contiguous_format))
        keep: list = []
        plan = _plan.tc_plan(self, self._lanes(), soma.device, keep)
        plan.v = self.v.data_ptr()
        plan.x_apical = apical.data_ptr()
        z = torch.empty_like(soma)
        v_pre = torch.empty_like(soma) if return_state else None
        _lib.check(_lib.load().btx_lif_step(plan, soma.data_ptr(), z.data_ptr(),
                                            None if v_pre is None else v_pre.data_ptr(),
                                            _lib.stream_ptr()), "btx_lif_step")
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
