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
"""

from __future__ import annotations

from collections.abc import Callable, Sequence
from numbers import Number

import torch

from ... import _lib
from .. import _plan
from ..base import BaseNode
from ..surrogate import Erf


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
        if torch.is_grad_enabled() and (x.requires_grad or any(
                p.requires_grad for p in self.parameters(recurse=False))):
            raise NotImplementedError("the GLIF3 step on the HIP backend is forward-only")
        if x.dtype != torch.float32 or self.v.dtype != torch.float32:
            raise TypeError("GLIF3 on the HIP backend supports float32 state only")
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
            self.refractory = self.refractory.contiguous()
            plan.refractory = self.refractory.data_ptr()
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
