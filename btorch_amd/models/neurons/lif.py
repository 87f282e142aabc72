"""Leaky integrate-and-fire neuron (btorch/models/neurons/lif.py:29-187).

dV/dt = -(V - V_reset)/tau + I/c_m, explicit Euler with the environment dt;
fire when (V - V_th)/(V_th - V_reset) >= 0; soft or hard reset; optional
refractory counter in ms.  ``single_step_forward`` runs one step in the HIP
library (``btx_lif_step``); inside ``RecurrentNN`` the step is fused with
the synapse and the sparse product instead.
"""

from __future__ import annotations

from collections.abc import Callable, Sequence

import torch

from ... import _lib
from .. import _plan
from ..base import BaseNode
from ..surrogate import Sigmoid


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

    def single_step_forward(self, x: torch.Tensor) -> torch.Tensor:
        _lib.require_hip(x, self.v)
        if self._needs_grad(x):
            raise NotImplementedError(
                "btorch_amd differentiates neurons inside RecurrentNN (fused adjoint); "
                "a standalone LIF step is forward-only."
            )
        if x.dtype != torch.float32 or self.v.dtype != torch.float32:
            raise TypeError("LIF on the HIP backend supports float32 state only")
        lead = tuple(self.v.shape[: self.v.ndim - len(self.n_neuron)])
        x = torch.broadcast_to(x, self.v.shape).contiguous()
        lanes = 1
        for d in lead:
            lanes *= d
        keep: list = []
        plan = _plan.lif_plan(self, lanes, x.device, keep)
        v = self.v.contiguous()
        self.v = v
        plan.v = v.data_ptr()
        if self._use_refractory:
            ref = self.refractory.contiguous()
            self.refractory = ref
            plan.refractory = ref.data_ptr()
        z = torch.empty_like(x)
        v_pre = torch.empty_like(x) if self.pre_spike_v else None
        _lib.check(_lib.load().btx_lif_step(plan, x.data_ptr(), z.data_ptr(),
                                            None if v_pre is None else v_pre.data_ptr(),
                                            _lib.stream_ptr()), "btx_lif_step")
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
