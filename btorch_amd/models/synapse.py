"""Current-based synapses (btorch/models/synapse.py).

``ExponentialPSC`` (synapse.py:145-187): psc decays by exponential Euler
``psc + k*((-psc)/tau_syn)`` with ``k = expm1(dt*l)/l``, ``l = -1/tau_syn``,
then ``psc = psc + linear(z)``.  ``DelayedPSC`` (synapse.py:537-623) feeds
the spike of ``max_delay_steps`` steps ago; ``HeterSynapsePSC``
(synapse.py:626-761) feeds the delay-flattened history to a delay-expanded
connection and sums receptors.  Standalone steps run ``btx_exp_decay`` and
the event-driven ``SparseConn``; inside ``RecurrentNN`` the whole step is one
fused launch sequence.
"""

from __future__ import annotations

from collections.abc import Sequence

import torch
import torch.nn.functional as F

from .. import _lib
from . import _plan, environ
from .base import MemoryModule, flatten_neuron, normalize_n_neuron, unflatten_neuron
from .history import SpikeHistory


class _ExpDecay(torch.autograd.Function):
    """psc + k*((-psc)/tau) computed by btx_exp_decay; linear backward."""

    @staticmethod
    def forward(ctx, psc, tau_vec, tau_stride, k_vec, k_stride, shape_n):
        out = psc.contiguous().clone()
        lanes = out.numel() // shape_n
        _lib.check(_lib.load().btx_exp_decay(out.data_ptr(), lanes, shape_n, tau_vec.data_ptr(),
                                             tau_stride, k_vec.data_ptr(), k_stride,
                                             _lib.stream_ptr()), "btx_exp_decay")
        ctx.save_for_backward(tau_vec, k_vec)
        ctx.meta = (tau_stride, k_stride, shape_n)
        return out

    @staticmethod
    def backward(ctx, g):
        tau_vec, k_vec = ctx.saved_tensors
        tau_stride, k_stride, n = ctx.meta
        tau = tau_vec if tau_stride else tau_vec.expand(n)
        k = k_vec if k_stride else k_vec.expand(n)
        g2 = g.reshape(-1, n)
        return (g2 + k * ((-g2) / tau)).reshape(g.shape), None, None, None, None, None


class BasePSC(MemoryModule):
    """Post-synaptic current with memory ``psc`` and a weight module ``linear``."""

    def __init__(self, n_neuron, linear: torch.nn.Module, step_mode="s", backend="torch"):
        super().__init__()
        self.n_neuron, self.size = normalize_n_neuron(n_neuron)
        self.step_mode = step_mode
        self.backend = backend
        self.linear = linear
        self.register_memory("psc", 0.0, self.n_neuron)

    def conductance_charge(self):
        raise NotImplementedError

    def adaptation_charge(self, z: torch.Tensor):
        z_flat, lead = flatten_neuron(z, self.n_neuron, self.size)
        wz = unflatten_neuron(self.linear(z_flat), lead, self.n_neuron)
        self.psc = self.psc + wz

    def current_charge(self, v=None):
        if v is not None:
            raise NotImplementedError("Only current-based PSC is supported.")
        return self.psc

    def single_step_forward(self, z: torch.Tensor):
        self.conductance_charge()
        self.adaptation_charge(z)
        return self.current_charge()

    def multi_step_forward(self, z_seq: torch.Tensor, kernel_len: int = 64):
        """Sequence form as a causal depthwise convolution with the truncated
        impulse response (synapse.py:105-142)."""
        dt = environ.get("dt")
        z_flat, lead = flatten_neuron(z_seq, self.n_neuron, self.size)
        wz = self.linear(z_flat)
        kernel = self.get_kernel(dt, kernel_len).to(wz.device, wz.dtype)
        ch = wz.reshape(wz.shape[0], -1).transpose(0, 1).unsqueeze(0)
        w = kernel.flip(0).view(1, 1, kernel_len).repeat(ch.shape[1], 1, 1)
        out = F.conv1d(F.pad(ch, (kernel_len - 1, 0)), w, groups=ch.shape[1])
        out = out.squeeze(0).transpose(0, 1).reshape(*lead, self.size)
        return unflatten_neuron(out, lead, self.n_neuron)

    def get_kernel(self, dt, kernel_len):
        raise NotImplementedError


class ExponentialPSC(BasePSC):
    """d psc/dt = -psc/tau_syn, plus ``linear(z)`` each step."""

    def __init__(self, n_neuron, tau_syn, linear, step_mode="s", backend="torch"):
        super().__init__(n_neuron, linear, step_mode=step_mode, backend=backend)
        self.register_buffer("tau_syn", torch.as_tensor(tau_syn, dtype=torch.float32),
                             persistent=False)

    def decay_coefficients(self, device):
        """(tau vec, stride, k vec, stride) for the kernels; k from CPU torch."""
        dt = environ.get("dt")
        tau, ts = _plan.param_vec(self.tau_syn, self.size, tuple(self.n_neuron), device)
        k, ks = _plan.param_vec(_plan.exp_euler_coeff(self.tau_syn, dt), self.size,
                                tuple(self.n_neuron), device)
        return tau, ts, k, ks

    def conductance_charge(self):
        _lib.require_hip(self.psc)
        tau, ts, k, ks = self.decay_coefficients(self.psc.device)
        self.psc = _ExpDecay.apply(self.psc, tau, ts, k, ks, self.size)
        return self.psc

    def get_kernel(self, dt, kernel_len):
        a = torch.exp(-dt / self.tau_syn)
        return a ** torch.arange(kernel_len, dtype=a.dtype, device=a.device)


class DelayedPSC(MemoryModule):
    """Feeds ``psc`` the spikes of ``max_delay_steps`` steps ago."""

    def __init__(self, psc: BasePSC, max_delay_steps: int = 1, use_circular_buffer: bool = False):
        super().__init__()
        self.psc_module = psc
        self.max_delay_steps = max_delay_steps
        self.use_circular_buffer = use_circular_buffer
        if max_delay_steps > 1:
            self.history = SpikeHistory(n_neuron=psc.n_neuron, max_delay_steps=max_delay_steps + 1,
                                        use_circular_buffer=use_circular_buffer)
        else:
            self.history = None

    @property
    def n_neuron(self):
        return self.psc_module.n_neuron

    @property
    def size(self):
        return self.psc_module.size

    @property
    def step_mode(self):
        return self.psc_module.step_mode

    @step_mode.setter
    def step_mode(self, value):
        if hasattr(self, "_modules") and "psc_module" in self._modules:
            self.psc_module.step_mode = value

    @property
    def psc(self):
        return self.psc_module.psc

    def single_step_forward(self, z: torch.Tensor):
        if self.history is not None:
            self.history.update(z)
            z = self.history.get_delay(self.max_delay_steps)
        return self.psc_module.single_step_forward(z)

    def multi_step_forward(self, z_seq: torch.Tensor, kernel_len: int = 64):
        return torch.stack([self.single_step_forward(z_seq[t]) for t in range(z_seq.shape[0])])


class HeterSynapsePSC(BasePSC):
    """Receptor-typed synapse over a delay-expanded connection.

    ``base_psc`` must be given (the reference defaults to AlphaPSC, whose
    kinetics this build does not provide); ``ExponentialPSC`` is the kinetics
    the fused engine runs.
    """

    def __init__(self, n_neuron, n_receptor: int, receptor_type_index, linear,
                 base_psc: type[BasePSC] | None = None, max_delay_steps: int = 1,
                 use_circular_buffer: bool = False, step_mode="s", backend="torch", **kwargs):
        super().__init__(n_neuron, linear, step_mode=step_mode, backend=backend)
        if base_psc is None:
            raise ValueError("HeterSynapsePSC needs base_psc (e.g. ExponentialPSC) in this build.")
        self.max_delay_steps = max_delay_steps
        self.use_circular_buffer = use_circular_buffer
        self.history = (SpikeHistory(n_neuron=self.size, max_delay_steps=max_delay_steps,
                                     use_circular_buffer=use_circular_buffer)
                        if max_delay_steps > 1 else None)
        self.base_psc = base_psc(n_neuron=self.size * n_receptor, linear=linear,
                                 step_mode=step_mode, backend=backend, **kwargs)
        self.n_receptor = n_receptor
        self.receptor_type_index = receptor_type_index

    def single_step_forward(self, z: torch.Tensor):
        lead = z.shape[: z.ndim - len(self.n_neuron)]
        z_flat = z.reshape(*lead, self.size)
        if self.history is not None:
            self.history.update(z_flat)
            z_flat = self.history.get_flattened(self.max_delay_steps)
        psc = self.base_psc.single_step_forward(z_flat)
        self.psc = psc.reshape(*lead, *self.n_neuron, self.n_receptor).sum(-1)
        return self.psc

    def get_psc(self, receptor_type: int | None = None, psc: torch.Tensor | None = None):
        psc = psc if psc is not None else self.base_psc.psc
        if receptor_type is None:
            return psc
        lead = psc.shape[:-1]
        return psc.reshape(*lead, self.size, self.n_receptor)[..., int(receptor_type)]


__all__ = ["BasePSC", "ExponentialPSC", "DelayedPSC", "HeterSynapsePSC"]
