"""Configuration of the sparse connectivity backend.

Same registry interface as btorch/_sparse_config.py:55-113
(``backend(name)``, ``selected``, ``resolve(name, override)``).  There is one
compute backend, the HIP event-driven product; the reference's backend names
select its accumulation mode so existing configurations keep working:

  "hip"     HipSparseConfig()               ascending-source, bit-exact
  "native"  HipSparseConfig(ordered=True)   same sums as torch.sparse.mm on CPU
  "triton"  HipSparseConfig(ordered=False)  order-free float atomics
"""

from __future__ import annotations

from copy import deepcopy
from dataclasses import asdict, dataclass, fields, is_dataclass
from typing import Any


@dataclass
class SparseBackendConfig:
    """Base configuration for sparse backends."""


@dataclass
class HipSparseConfig(SparseBackendConfig):
    """Event-driven HIP product settings.

    ordered: sum each destination's terms in ascending source order (bit-exact
        against the CPU reference); False uses float atomics.
    slots: per-destination terms the ordered mode sorts (2..16); a
        destination with more terms in one step is reduced by pulling its
        incoming edges (exact, slower for that destination).
    bin_cap: terms per (lane, 1024-destination segment) bin of the recurrent
        engine's ordered mode, 0 = the maximum (8192); a bin that receives
        more is reduced by pulling every destination of the segment.
    dense_pull: let the recurrent engine deliver a step whose push would
        bin more than bin_cap / 2 terms per bin on average (more than 1024
        when lanes x segments < 256) as a dense pull over every destination's
        incoming edges (same bits); False always pushes into the bins.
    """

    ordered: bool = True
    slots: int = 16
    bin_cap: int = 0
    dense_pull: bool = True

    def validate(self) -> None:
        if self.slots < 2 or self.slots > 16:
            raise ValueError("slots must be in [2, 16].")
        if self.bin_cap < 0 or self.bin_cap > 8192:
            raise ValueError("bin_cap must be in [0, 8192].")


class SparseConfigRegistry:
    """Selected backend name plus one mutable config template per name."""

    def __init__(self, selected: str | None = None):
        self._configs: dict[str, SparseBackendConfig] = {
            "hip": HipSparseConfig(),
            "native": HipSparseConfig(ordered=True),
            "triton": HipSparseConfig(ordered=False),
        }
        self._selected = None
        if selected:
            self.backend(selected)

    @property
    def selected(self) -> str | None:
        return self._selected

    def names(self) -> list[str]:
        return sorted(self._configs)

    def backend(self, name: str | None = None) -> SparseBackendConfig:
        if name is not None:
            key = name.lower()
            if key not in self._configs:
                raise ValueError(
                    f"Unknown sparse backend '{name}'; expected one of {', '.join(self.names())}."
                )
            self._selected = key
        if self._selected is None:
            raise RuntimeError(
                "No global sparse backend is selected; call config.sparse.backend('hip') first."
            )
        return self._configs[self._selected]

    def resolve(self, name: str, override: SparseBackendConfig | dict[str, Any] | None = None):
        key = name.lower()
        if key not in self._configs:
            raise ValueError(
                f"Unknown sparse backend '{name}'; expected one of {', '.join(self.names())}."
            )
        cfg = deepcopy(self._configs[key])
        if override is not None:
            values = asdict(override) if is_dataclass(override) else dict(override)
            valid = {f.name for f in fields(cfg)}
            unknown = set(values) - valid
            if unknown:
                raise ValueError(f"Unknown {name} sparse config fields: {', '.join(sorted(unknown))}.")
            for k, v in values.items():
                setattr(cfg, k, v)
        validate = getattr(cfg, "validate", None)
        if validate is not None:
            validate()
        return cfg


__all__ = ["SparseBackendConfig", "HipSparseConfig", "SparseConfigRegistry"]
