"""Current-based synapses (btorch/models/synapse.py).

``ExponentialPSC`` (synapse.py:145-187): psc decays by exponential Euler
``psc + k*((-psc)/tau_syn)`` with ``k = expm1(dt*l)/l``, ``l = -1/tau_syn``,
then ``psc = psc + linear(z)``.  ``DelayedPSC`` (synapse.py:537-623) feeds
the spike of ``max_delay_steps`` steps ago; ``HeterSynapsePSC``
(synapse.py:626-761) feeds the delay-flattened history to a delay-expanded
connection and sums receptors.  Standalone steps run ``btx_exp_decay`` and
the event-driven ``SparseConn``; inside ``RecurrentNN`` the whole step is one
fused launch sequence.

``AlphaPSC`` (synapse.py:274-334), ``AlphaPSCBilleh`` (synapse.py:201-271)
and ``DualExponentialPSC`` (synapse.py:337-418) carry a second state (h) or
two conductances (g_rise, g_decay).  The fused step runs them too; a
standalone step decays every state in one ``btx_syn_decay`` launch (with a
linear backward) and then adds the push.
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


class _SynDecay(torch.autograd.Function):
    """Decay of an alpha or dual-exponential synapse by btx_syn_decay
    (psc, h) or (psc, g_rise, g_decay) -> new tensors; linear backward."""

    @staticmethod
    def forward(ctx, mod, psc, *states):
        psc = psc.contiguous().clone()
        states = [t.contiguous().clone() for t in states]
        p, coeffs = mod.coefficient_vectors(psc.device)
        p.lanes = psc.numel() // mod.size
        p.psc = psc.data_ptr()
        p.syn_h1 = states[0].data_ptr()
        p.syn_h2 = states[1].data_ptr() if len(states) > 1 else None
        _lib.check(_lib.load().btx_syn_decay(p, _lib.stream_ptr()), "btx_syn_decay")
        ctx.kind, ctx.n, ctx.coeffs = mod._KIND, mod.size, coeffs
        return (psc, *states)

    @staticmethod
    def backward(ctx, *grads):
        n, c = ctx.n, ctx.coeffs
        shape = grads[0].shape
        k, ts = c["k"], c["tau"]
        g = [x.reshape(-1, n) for x in grads]
        if ctx.kind == _lib.SYN_ALPHA:
            gp, gh = g
            out = [gp + k * ((-gp) / ts), k * (gp / ts) + (gh + k * ((-gh) / ts))]
        elif ctx.kind == _lib.SYN_ALPHA_BILLEH:
            gp, gh = g
            out = [k * gp, k * gp + k * gh]
        else:
            gp, g1, g2 = g
            out = [gp, g1 + k * ((-g1) / ts), g2 + c["k2"] * ((-g2) / c["tau2"])]
        return (None, *[x.reshape(shape) for x in out])


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

    def synapse_plan(self, p, device, keep: list) -> None:
        """Kinetics fields of a fused-step plan (BTX_SYN_EXP)."""
        tau, ts, k, ks = self.decay_coefficients(device)
        keep += [tau, k]
        p.syn_kind = _lib.SYN_EXP
        p.tau_syn, p.tau_syn_stride = tau.data_ptr(), ts
        p.k_syn, p.k_syn_stride = k.data_ptr(), ks


class _KineticsPSC(BasePSC):
    """Synapse with states besides psc, decayed on the device by
    ``btx_syn_decay`` (the fused step's own update)."""

    _KIND = -1
    _STATES: tuple[str, ...] = ()   # memories in syn_h1, syn_h2 order

    def _vec(self, t, device, keep):
        vec, stride = _plan.param_vec(t, self.size, tuple(self.n_neuron), device)
        keep.append(vec)
        return vec.data_ptr(), stride

    def synapse_plan(self, p, device, keep: list) -> None:
        raise NotImplementedError

    def state_tensors(self) -> list[torch.Tensor]:
        return [getattr(self, k) for k in self._STATES]

    def conductance_charge(self):
        """Decays psc and the other states together: psc reads h before h
        decays, as the reference's conductance_charge precedes the decay in
        adaptation_charge.  ``adaptation_charge`` then adds the push."""
        _lib.require_hip(self.psc)
        for k in ("psc",) + self._STATES:
            t = getattr(self, k)
            if t is None or t.dtype != torch.float32 or t.shape != self.psc.shape:
                raise RuntimeError(f"{type(self).__name__}.{k} is not initialised "
                                   "(call init_state)")
        st = self.state_tensors()
        out = _SynDecay.apply(self, self.psc, *st)
        self.psc = out[0]
        for k, t in zip(self._STATES, out[1:]):
            setattr(self, k, t)

    def coefficient_vectors(self, device):
        """A plan holding this synapse's coefficients (n set, lanes not), and
        the same coefficients as (n,) device tensors (which keep the plan's
        pointers alive)."""
        keep: list = []
        p = _lib.RsnnPlan()
        p.n = self.size
        self.synapse_plan(p, device, keep)
        n = self.size
        by_ptr = {t.data_ptr(): (t if t.numel() == n else t.expand(n)) for t in keep}

        def get(ptr):
            return by_ptr[ptr] if ptr else None
        return p, {"k": get(p.k_syn), "tau": get(p.tau_syn), "k2": get(p.k_syn2),
                   "tau2": get(p.tau_syn2), "_keep": keep}

    def _push(self, z: torch.Tensor) -> torch.Tensor:
        if len(self.n_neuron) > 1 and z.shape[-len(self.n_neuron):] == self.n_neuron:
            z_flat, lead = flatten_neuron(z, self.n_neuron, self.size)
            return unflatten_neuron(self.linear(z_flat), lead, self.n_neuron)
        return self.linear(z)


class _Adaptive2VarPSC(_KineticsPSC):
    _STATES = ("h",)

    def __init__(self, n_neuron, linear, step_mode="s", backend="torch"):
        super().__init__(n_neuron, linear, step_mode=step_mode, backend=backend)
        self.register_memory("h", 0.0, self.n_neuron)


class AlphaPSCBilleh(_Adaptive2VarPSC):
    """Current-based alpha synapse of Billeh et al. (2019), peak 1 at
    t = tau_syn for W = 1; needs dt == 1:
    psc = c*psc + c*h ; h = c*h + (e/tau_syn)*W z with c = exp(-1/tau_syn)."""

    _KIND = _lib.SYN_ALPHA_BILLEH

    def __init__(self, n_neuron, tau_syn, linear, step_mode="s", backend="torch"):
        super().__init__(n_neuron, linear, step_mode, backend)
        dt = environ.get("dt", None)
        if dt is not None and dt != 1.0:
            raise ValueError(f"dt must be 1.0 for this model, got {dt}")
        self.register_buffer("tau_syn", torch.as_tensor(tau_syn, dtype=torch.float32),
                             persistent=False)
        self.register_buffer("syn_decay", torch.exp(-1.0 / self.tau_syn), persistent=False)
        # e/tau_syn in the reference's ops (python scalar over the tensor)
        self.register_buffer("syn_gain", torch.e / self.tau_syn, persistent=False)

    def gain(self) -> torch.Tensor:
        return self.syn_gain

    def synapse_plan(self, p, device, keep: list) -> None:
        p.syn_kind = self._KIND
        p.tau_syn, p.tau_syn_stride = self._vec(self.tau_syn, device, keep)
        p.k_syn, p.k_syn_stride = self._vec(self.syn_decay, device, keep)
        p.syn_gain, p.syn_gain_stride = self._vec(self.gain(), device, keep)

    def adaptation_charge(self, z: torch.Tensor):
        self.h = self.h + self.syn_gain * self._push(z)

    def get_kernel(self, dt, kernel_len):
        """k[0] = 0, k[t] = t*(e/tau_syn)*c^t (synapse.py:255-271)."""
        a = self.syn_decay
        t = torch.arange(kernel_len, dtype=a.dtype, device=a.device)
        kernel = torch.zeros_like(t)
        if kernel_len > 1:
            kernel[1:] = t[1:] * (torch.e / self.tau_syn) * (a ** t[1:])
        return kernel


class AlphaPSC(_Adaptive2VarPSC):
    """Current-based alpha synapse (Brainpy form):
    psc += k*((-psc)/tau + h/tau) ; h = (h + k*((-h)/tau)) + g_max*W z."""

    _KIND = _lib.SYN_ALPHA

    def __init__(self, n_neuron, tau_syn, linear, g_max=1.0, step_mode="s", backend="torch"):
        super().__init__(n_neuron, linear, step_mode, backend)
        self.register_buffer("tau_syn", torch.as_tensor(tau_syn, dtype=torch.float32),
                             persistent=False)
        self.register_buffer("g_max", torch.as_tensor(g_max, dtype=torch.float32),
                             persistent=False)

    def synapse_plan(self, p, device, keep: list) -> None:
        dt = environ.get("dt")
        p.syn_kind = self._KIND
        p.tau_syn, p.tau_syn_stride = self._vec(self.tau_syn, device, keep)
        p.k_syn, p.k_syn_stride = self._vec(_plan.exp_euler_coeff(self.tau_syn, dt), device,
                                            keep)
        p.syn_gain, p.syn_gain_stride = self._vec(self.g_max, device, keep)

    def adaptation_charge(self, z: torch.Tensor):
        self.h = self.h + self.g_max.to(self.h.device) * self._push(z)

    def get_kernel(self, dt, kernel_len):
        """k[0] = 0, k[t] = t*(1 - a)*a^(t-1), a = exp(-dt/tau_syn), times g_max
        (synapse.py:319-334)."""
        a = torch.exp(-dt / self.tau_syn)
        t = torch.arange(kernel_len, dtype=a.dtype, device=a.device)
        kernel = torch.zeros_like(t)
        if kernel_len > 1:
            kernel[1:] = t[1:] * (1 - a) * (a ** (t[1:] - 1))
        return self.g_max.to(kernel.dtype) * kernel


class DualExponentialPSC(_KineticsPSC):
    """Difference of a rise and a decay conductance:
    g_r, g_d decay by exponential Euler, both add W z, psc = a*(g_d - g_r),
    a = (tau_decay - tau_rise)/tau_rise/tau_decay*A with A normalising the
    peak of exp(-t/tau_decay) - exp(-t/tau_rise) to 1 by default."""

    _KIND = _lib.SYN_DUAL_EXP
    _STATES = ("g_rise", "g_decay")

    def __init__(self, n_neuron, tau_decay, tau_rise, linear, A=None, step_mode="s",
                 backend="torch"):
        super().__init__(n_neuron, linear, step_mode=step_mode, backend=backend)
        self.register_buffer("tau_decay", torch.as_tensor(tau_decay, dtype=torch.float32),
                             persistent=False)
        self.register_buffer("tau_rise", torch.as_tensor(tau_rise, dtype=torch.float32),
                             persistent=False)
        if A is None:
            A = (self.tau_decay / (self.tau_decay - self.tau_rise)
                 * (self.tau_rise / self.tau_decay)
                 ** (self.tau_rise / (self.tau_rise - self.tau_decay)))
        A = torch.as_tensor(A)
        a = (self.tau_decay - self.tau_rise) / self.tau_rise / self.tau_decay * A
        self.register_buffer("a", a, persistent=False)
        self.register_memory("g_rise", 0.0, self.n_neuron)
        self.register_memory("g_decay", 0.0, self.n_neuron)

    def synapse_plan(self, p, device, keep: list) -> None:
        dt = environ.get("dt")
        p.syn_kind = self._KIND
        p.tau_syn, p.tau_syn_stride = self._vec(self.tau_rise, device, keep)
        p.k_syn, p.k_syn_stride = self._vec(_plan.exp_euler_coeff(self.tau_rise, dt), device,
                                            keep)
        p.tau_syn2, p.tau_syn2_stride = self._vec(self.tau_decay, device, keep)
        p.k_syn2, p.k_syn2_stride = self._vec(_plan.exp_euler_coeff(self.tau_decay, dt),
                                              device, keep)
        p.syn_a, p.syn_a_stride = self._vec(self.a, device, keep)

    def adaptation_charge(self, z: torch.Tensor):
        wz = self._push(z)
        self.g_rise = self.g_rise + wz
        self.g_decay = self.g_decay + wz
        self.psc = self.a.to(wz.device) * (self.g_decay - self.g_rise)

    def get_kernel(self, dt, kernel_len):
        """k[t] = a*(a_d^t - a_r^t), a_x = exp(-dt/tau_x) (synapse.py:408-418)."""
        a_r = torch.exp(-dt / self.tau_rise)
        a_d = torch.exp(-dt / self.tau_decay)
        t = torch.arange(kernel_len, dtype=self.a.dtype, device=self.a.device)
        return self.a * (a_d ** t - a_r ** t)


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

    ``base_psc`` defaults to AlphaPSC, as in the reference.
    """

    def __init__(self, n_neuron, n_receptor: int, receptor_type_index, linear,
                 base_psc: type[BasePSC] | None = None, max_delay_steps: int = 1,
                 use_circular_buffer: bool = False, step_mode="s", backend="torch", **kwargs):
        super().__init__(n_neuron, linear, step_mode=step_mode, backend=backend)
        if base_psc is None:
            base_psc = AlphaPSC
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


__all__ = ["BasePSC", "ExponentialPSC", "AlphaPSC", "AlphaPSCBilleh", "DualExponentialPSC",
           "DelayedPSC", "HeterSynapsePSC"]
