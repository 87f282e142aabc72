"""Network-wide state helpers (btorch/models/functional.py:26-500).

Hidden states are addressed by dotted keys ``<module path>.<memory name>``
(e.g. ``neuron.v``, ``synapse.psc``), the keys ``named_hidden_states``
returns and ``set_hidden_states`` accepts.
"""

from __future__ import annotations

import logging
from collections.abc import Sequence
from contextlib import contextmanager
from typing import Any

import torch
from torch import nn

from . import base


def _move_net(net: nn.Module, kwargs: dict[str, Any]) -> None:
    move = {k: kwargs[k] for k in ("device", "dtype") if kwargs.get(k) is not None}
    if move:
        net.to(**move)


def init_net_state(net: nn.Module, batch_size=None, **kwargs):
    """``init_state(batch_size, **kwargs)`` on every stateful submodule."""
    _move_net(net, kwargs)
    for m in net.modules():
        if hasattr(m, "init_state") and callable(m.init_state):
            if not isinstance(m, base.MemoryModule):
                logging.warning("Trying to call `init_state()` of %s, which is not MemoryModule", m)
            m.init_state(batch_size, **kwargs)


def reset_net(net: nn.Module, batch_size=None, **kwargs):
    """``reset(batch_size, **kwargs)`` on every stateful submodule."""
    _move_net(net, kwargs)
    for m in net.modules():
        if hasattr(m, "reset") and callable(m.reset):
            m.reset(batch_size, **kwargs)


reset_net_state = reset_net


@contextmanager
def prepare_sparse_modules(net: nn.Module, batch_size: int | None = None):
    """Pack sparse weights once for an enclosing multi-step run."""
    from .linear import BaseSparseConn

    prepared = []
    try:
        for m in net.modules():
            if isinstance(m, BaseSparseConn):
                m.prepare_sparse(batch_size=batch_size)
                prepared.append(m)
        yield
    finally:
        for m in reversed(prepared):
            m.finish_sparse()


def _strip_self(names) -> set[str]:
    return {s.removeprefix("self.").removeprefix("self") for s in names}


def named_hidden_states(mod: nn.Module, names: Sequence[str] | None = None,
                        allow_buffer: bool = False, clone: bool = False) -> dict[str, Any]:
    """Dotted dict of every memory (optionally only ``names``)."""
    wanted = _strip_self(names) if names is not None else None
    out = {}
    for name, m in mod.named_modules():
        if not (allow_buffer or isinstance(m, base.MemoryModule)):
            continue
        items = m._memories.items() if isinstance(m, base.MemoryModule) else m._buffers.items()
        prefix = "" if name == "" else f"{name}."
        for k, v in items:
            key = prefix + k
            if wanted is None or name in wanted or key in wanted:
                out[key] = v.clone() if clone and torch.is_tensor(v) else v
    return out


named_memory_values = filter_hidden_states = named_hidden_states


def hidden_state_names(mod: nn.Module, names: Sequence[str] | None = None,
                       allow_buffer: bool = False) -> list[str]:
    """The keys ``named_hidden_states`` would return, without reading any
    memory (so memories held elsewhere between runs are not written back)."""
    wanted = _strip_self(names) if names is not None else None
    out = []
    for name, m in mod.named_modules():
        if isinstance(m, base.MemoryModule):
            keys = list(m._memories_rv)
        elif allow_buffer:
            keys = list(m._buffers)
        else:
            continue
        prefix = "" if name == "" else f"{name}."
        for k in keys:
            key = prefix + k
            if wanted is None or name in wanted or key in wanted:
                out.append(key)
    return out


def set_hidden_states(mod: nn.Module, hidden_states: dict[str, Any], allow_buffer: bool = False,
                      inplace: bool = False):
    """Assign memories from a dotted dict (or ``{"module": {"v": ...}}``)."""
    if hidden_states is None:
        return
    for name, value in hidden_states.items():
        if value is None:
            continue
        path = name.removeprefix("self.").removeprefix("self").split(".")
        path = [p for p in path if p != ""]
        m = mod
        for p in path[:-1]:
            m = getattr(m, p)
        leaf = path[-1] if path else None
        target = getattr(m, leaf) if leaf is not None and hasattr(m, leaf) else None
        if leaf is None or isinstance(target, nn.Module):
            owner = m if leaf is None else target
            for k, v in value.items():
                _assign(owner, k, v, inplace)
        else:
            _assign(m, leaf, value, inplace)


def _assign(m: nn.Module, key: str, value, inplace: bool):
    if isinstance(m, base.MemoryModule):
        assert key in m._memories_rv, f"{key} is not a memory of {type(m).__name__}"
    if inplace:
        getattr(m, key).copy_(value)
    else:
        setattr(m, key, value)


set_memory_values = set_hidden_states


def named_memory_reset_values(mod: nn.Module, names: Sequence[str] | None = None):
    wanted = _strip_self(names) if names is not None else None
    out = {}
    for name, m in mod.named_modules():
        if isinstance(m, base.MemoryModule):
            prefix = "" if name == "" else f"{name}."
            for k, v in m._memories_rv.items():
                if wanted is None or name in wanted or prefix + k in wanted:
                    out[prefix + k] = v
    return out


def detach_net(net: nn.Module):
    for m in net.modules():
        if isinstance(m, base.MemoryModule):
            m.detach()


__all__ = [
    "init_net_state", "reset_net", "reset_net_state", "prepare_sparse_modules",
    "named_hidden_states", "filter_hidden_states", "hidden_state_names", "set_hidden_states",
    "named_memory_reset_values", "detach_net",
]
