"""Training path of the fused engine (placeholder until the adjoint lands)."""

from __future__ import annotations

import torch


def needs_grad(rnn, x: torch.Tensor) -> bool:
    if not torch.is_grad_enabled():
        return False
    if x.requires_grad:
        return True
    return any(p.requires_grad for p in rnn.parameters())


def run_with_grad(rnn, eng, x):
    raise NotImplementedError("gradients through the fused engine are not available yet")
