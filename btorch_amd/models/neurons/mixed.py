"""Population of several neuron models behind one neuron axis
(btorch/models/neurons/mixed.py:16-193).

The input current is sliced along the last axis, each group steps with its
own HIP neuron step, and the group spikes are concatenated.  A population of
one two-compartment group is what ``ApicalRecurrentNN`` hands to the fused
engine.
"""

from __future__ import annotations

from collections.abc import Mapping, Sequence

import torch
from torch import nn

from .two_compartment import TwoCompartmentGLIF


class MixedNeuronPopulation(nn.Module):
    """Groups given as ``[(count, neuron), ...]`` (named ``group_<i>``) or a
    mapping ``name -> (count, neuron)``; raises ValueError for no groups or
    a count below 1, as the reference does."""

    def __init__(self, groups: Sequence[tuple[int, nn.Module]] | Mapping[str, tuple[int, nn.Module]],
                 step_mode: str = "s"):
        super().__init__()
        if not groups:
            raise ValueError("groups must contain at least one population.")
        if isinstance(groups, Mapping):
            items = [(name, count, neuron) for name, (count, neuron) in groups.items()]
        else:
            items = [(f"group_{i}", count, neuron) for i, (count, neuron) in enumerate(groups)]
        counts, names, total = [], [], 0
        for name, count, neuron in items:
            if count <= 0:
                raise ValueError(f"Group {name!r} count must be positive, got {count}.")
            self.add_module(name, neuron)
            names.append(name)
            counts.append(count)
            total += int(neuron.size) if hasattr(neuron, "size") else count
        self.counts = counts
        self._group_names = names
        self._cumsum = [0]
        for c in counts:
            self._cumsum.append(self._cumsum[-1] + c)
        self.n_neuron = (total,)
        self.size = total
        self.step_mode = step_mode

    def groups(self) -> list[nn.Module]:
        return [getattr(self, name) for name in self._group_names]

    def _indices(self, idx: int) -> torch.Tensor:
        return torch.arange(self._cumsum[idx], self._cumsum[idx + 1], dtype=torch.long)

    def _concat_attr(self, attr: str) -> torch.Tensor:
        """``attr`` of every group along the neuron axis, scalars broadcast to
        the group's count (mixed.py:85-112)."""
        parts = []
        for idx, neuron in enumerate(self.groups()):
            val = getattr(neuron, attr, None)
            if val is None:
                raise AttributeError(f"{neuron.__class__.__name__} has no attribute {attr!r}")
            if isinstance(val, nn.Parameter):
                val = val.data
            if val.dim() == 0:
                val = val.expand(self._cumsum[idx + 1] - self._cumsum[idx])
            parts.append(val)
        return torch.cat(parts, dim=-1)

    @property
    def v_threshold(self) -> torch.Tensor:
        return self._concat_attr("v_threshold")

    @property
    def v_reset(self) -> torch.Tensor:
        return self._concat_attr("v_reset")

    def _slice(self, x: torch.Tensor, idx: int) -> torch.Tensor:
        return x[..., self._cumsum[idx]:self._cumsum[idx + 1]]

    def single_step_forward(self, x: torch.Tensor, i_apical: torch.Tensor | None = None):
        spikes = []
        for idx, neuron in enumerate(self.groups()):
            soma = self._slice(x, idx)
            if isinstance(neuron, TwoCompartmentGLIF):
                apical = self._slice(i_apical, idx) if i_apical is not None else None
                out = neuron(soma, apical)
            else:
                out = neuron(soma)
            spikes.append(out[0] if isinstance(out, tuple) else out)
        return torch.cat(spikes, dim=-1)

    def multi_step_forward(self, x: torch.Tensor, i_apical: torch.Tensor | None = None):
        return torch.stack([self.single_step_forward(x[t], None if i_apical is None
                                                     else i_apical[t])
                            for t in range(x.shape[0])])

    def forward(self, x: torch.Tensor, i_apical: torch.Tensor | None = None):
        if self.step_mode == "m":
            return self.multi_step_forward(x, i_apical)
        return self.single_step_forward(x, i_apical)

    def extra_repr(self) -> str:
        parts = [f"n_neuron={self.n_neuron}"]
        for i, (name, neuron) in enumerate(self.named_children()):
            parts.append(f"{name}={neuron.__class__.__name__}(n={self.counts[i]})")
        return ", ".join(parts)


__all__ = ["MixedNeuronPopulation"]
