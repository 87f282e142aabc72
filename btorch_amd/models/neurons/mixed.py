"""One neuron axis made of several neuron models (the interface of
btorch/models/neurons/mixed.py:16-193: ``MixedNeuronPopulation(groups,
step_mode)`` with groups ``[(count, neuron), ...]`` or ``{name: (count,
neuron)}``, submodules named ``group_<i>`` or by their keys, ``counts``,
``size``, ``n_neuron``).

The population is described by a ``GroupLayout``: contiguous, ordered
neuron ranges, one per group.  Inside ``RecurrentNN`` the fused engine runs
the whole population in one step kernel, each neuron with its own model's
update (``_engine``); a standalone step routes each range of the input
through its group's HIP step and writes the spikes into one output tensor.
"""

from __future__ import annotations

from collections.abc import Mapping, Sequence
from dataclasses import dataclass

import torch
from torch import nn

from .two_compartment import TwoCompartmentGLIF


@dataclass(frozen=True)
class GroupLayout:
    """Names, first neuron and neuron count of each group, in axis order."""

    names: tuple[str, ...]
    starts: tuple[int, ...]
    counts: tuple[int, ...]

    @property
    def total(self) -> int:
        return self.starts[-1] + self.counts[-1]

    def span(self, g: int) -> slice:
        return slice(self.starts[g], self.starts[g] + self.counts[g])

    def group_of(self, neuron: int) -> int:
        for g in range(len(self.names) - 1, -1, -1):
            if neuron >= self.starts[g]:
                return g
        raise IndexError(neuron)


def _normalise(groups) -> list[tuple[str, int, nn.Module]]:
    if isinstance(groups, Mapping):
        entries = [(str(k), spec) for k, spec in groups.items()]
    else:
        entries = [(f"group_{i}", spec) for i, spec in enumerate(groups)]
    if len(entries) == 0:
        raise ValueError("MixedNeuronPopulation needs at least one (count, neuron) group.")
    out = []
    for name, spec in entries:
        count, neuron = spec
        count = int(count)
        if count < 1:
            raise ValueError(f"MixedNeuronPopulation group {name!r}: count {count} is not "
                             "positive.")
        size = int(getattr(neuron, "size", count))
        out.append((name, size, neuron))
    return out


class MixedNeuronPopulation(nn.Module):
    """Several neuron models on one neuron axis (see the module doc)."""

    def __init__(self, groups: Sequence[tuple[int, nn.Module]] | Mapping[str, tuple[int, nn.Module]],
                 step_mode: str = "s"):
        super().__init__()
        entries = _normalise(groups)
        starts, at = [], 0
        for name, size, neuron in entries:
            self.add_module(name, neuron)
            starts.append(at)
            at += size
        self.layout = GroupLayout(tuple(e[0] for e in entries), tuple(starts),
                                  tuple(e[1] for e in entries))
        self.counts = list(self.layout.counts)
        self._group_names = list(self.layout.names)
        self.size = at
        self.n_neuron = (at,)
        self.step_mode = step_mode

    def groups(self) -> list[nn.Module]:
        return [self._modules[name] for name in self.layout.names]

    def population_param(self, name: str) -> torch.Tensor:
        """Parameter ``name`` of every group laid out along the neuron axis
        (a group's scalar repeated over its range)."""
        lay = self.layout
        ref = None
        for g in self.groups():
            t = getattr(g, name, None)
            if t is None:
                raise AttributeError(f"{type(g).__name__} has no parameter {name!r}")
            ref = t if ref is None else ref
        out = torch.empty(lay.total, dtype=ref.dtype, device=ref.device)
        for g, mod in enumerate(self.groups()):
            out[lay.span(g)] = getattr(mod, name).detach().reshape(-1).expand(lay.counts[g])
        return out

    @property
    def v_threshold(self) -> torch.Tensor:
        return self.population_param("v_threshold")

    @property
    def v_reset(self) -> torch.Tensor:
        return self.population_param("v_reset")

    def single_step_forward(self, x: torch.Tensor, i_apical: torch.Tensor | None = None):
        lay = self.layout
        spikes = None
        for g, mod in enumerate(self.groups()):
            sl = lay.span(g)
            soma = x[..., sl]
            if isinstance(mod, TwoCompartmentGLIF):
                out = mod(soma, None if i_apical is None else i_apical[..., sl])
            else:
                out = mod(soma)
            z = out[0] if isinstance(out, tuple) else out
            if spikes is None:
                spikes = z.new_empty(*z.shape[:-1], lay.total)
            spikes[..., sl] = z
        return spikes

    def multi_step_forward(self, x: torch.Tensor, i_apical: torch.Tensor | None = None):
        steps = [self.single_step_forward(x[t], None if i_apical is None else i_apical[t])
                 for t in range(x.shape[0])]
        return torch.stack(steps)

    def forward(self, x: torch.Tensor, i_apical: torch.Tensor | None = None):
        step = self.multi_step_forward if self.step_mode == "m" else self.single_step_forward
        return step(x, i_apical)

    def extra_repr(self) -> str:
        lay = self.layout
        parts = [f"n_neuron={self.n_neuron}"]
        parts += [f"{name}[{s}:{s + c}]" for name, s, c in zip(lay.names, lay.starts, lay.counts)]
        return ", ".join(parts)


__all__ = ["GroupLayout", "MixedNeuronPopulation"]
