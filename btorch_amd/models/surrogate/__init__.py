"""Surrogate-gradient spike functions (Heaviside forward, smooth backward).

Same classes and parameters as btorch/models/surrogate: ``alpha`` is the
inverse half-width of the surrogate derivative (its peak is ``damping``),
``spiking=False`` swaps the Heaviside forward for the smooth primitive.
The fused HIP kernels evaluate the same derivative per neuron; these modules
carry the parameters and give the standalone elementwise form.
"""

from __future__ import annotations

import math

import torch

SIGMOID_K = 2.0 * math.log(math.sqrt(2.0) + 1.0)
SUPERSPIKE_K = math.sqrt(2.0) - 1.0
LN2 = math.log(2.0)


class _SurrogateFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, module):
        ctx.module = module
        ctx.save_for_backward(x)
        if module.spiking:
            return (x >= 0).to(x)
        return module.primitive(x)

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        m = ctx.module
        return m.derivative(x, g, m.damping_factor), None


class SurrogateFunctionBase(torch.nn.Module):
    """Heaviside spike with a surrogate derivative; see module docstring."""

    kernel_id = -1

    def __init__(self, alpha: float = 1.0, damping_factor: float = 1.0, spiking: bool = True):
        super().__init__()
        self.alpha = alpha
        self.damping_factor = damping_factor
        self.spiking = spiking

    def primitive(self, x):
        raise NotImplementedError

    def derivative(self, x, grad_output, damping_factor: float = 1.0):
        raise NotImplementedError

    def forward(self, x):
        return _SurrogateFn.apply(x, self)


class Sigmoid(SurrogateFunctionBase):
    """g(v) = 4 s (1 - s), s = sigmoid(k alpha v), k = 2 ln(sqrt2 + 1)
    (btorch/models/surrogate/sigmoid.py:16-29)."""

    kernel_id = 0

    def __init__(self, alpha: float = 2.0, damping_factor: float = 1.0, spiking: bool = True):
        super().__init__(alpha, damping_factor, spiking)

    def primitive(self, x):
        return torch.sigmoid(SIGMOID_K * self.alpha * x)

    def derivative(self, x, grad_output, damping_factor: float = 1.0):
        s = torch.sigmoid(SIGMOID_K * self.alpha * x)
        return grad_output * (damping_factor * 4.0 * s * (1.0 - s))


class SuperSpike(SurrogateFunctionBase):
    """g(v) = 1/(1 + (sqrt2 - 1) alpha |v|)^2, the fast sigmoid
    (btorch/models/surrogate/superspike.py:16-31)."""

    kernel_id = 1

    def __init__(self, alpha: float = 2.0, damping_factor: float = 1.0, spiking: bool = True):
        super().__init__(alpha, damping_factor, spiking)

    def primitive(self, x):
        ka = SUPERSPIKE_K * self.alpha
        return torch.where(x >= 0.0, 1.0 - 0.5 / (1.0 + ka * x), 0.5 / (1.0 - ka * x))

    def derivative(self, x, grad_output, damping_factor: float = 1.0):
        return grad_output * (damping_factor / (1.0 + SUPERSPIKE_K * self.alpha * x.abs()) ** 2)


class Erf(SurrogateFunctionBase):
    """g(v) = 2^{-(alpha v)^2} (btorch/models/surrogate/erf.py:22-28)."""

    kernel_id = 2

    def __init__(self, alpha: float = 4.0, variance: float | None = None,
                 damping_factor: float = 1.0, spiking: bool = True):
        if variance is not None:
            alpha = 1.0 / math.sqrt(variance)
        super().__init__(alpha, damping_factor, spiking)

    def primitive(self, x):
        return torch.special.erfc(-math.sqrt(LN2) * self.alpha * x) / 2.0

    def derivative(self, x, grad_output, damping_factor: float = 1.0):
        return grad_output * (damping_factor * torch.exp(-LN2 * (self.alpha * x) ** 2))


class ATan(SurrogateFunctionBase):
    """g(v) = 1/(1 + (alpha v)^2), the Cauchy kernel (btorch/models/surrogate/atan.py:49-72)."""

    kernel_id = 3

    def __init__(self, alpha: float = 2.0, damping_factor: float = 1.0, spiking: bool = True):
        super().__init__(alpha, damping_factor, spiking)

    def primitive(self, x):
        return 0.5 + torch.atan(self.alpha * x) / math.pi

    def derivative(self, x, grad_output, damping_factor: float = 1.0):
        return grad_output * (damping_factor / (1.0 + (self.alpha * x) ** 2))


__all__ = ["SurrogateFunctionBase", "Sigmoid", "SuperSpike", "Erf", "ATan"]
