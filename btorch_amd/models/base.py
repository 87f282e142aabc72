"""Stateful-module runtime: step-mode dispatch, parameter/buffer definition
and managed memories (per-lane state tensors).

Interface mirrors btorch/models/base.py:
  * ``StepModule`` -- ``step_mode`` "s"/"m" dispatch (base.py:17-52)
  * ``ParamBufferMixin.def_param`` -- per-neuron or scalar parameter/buffer,
    uniform values saved as 0-d tensors (base.py:75-457)
  * ``MemoryModule.register_memory / init_state / reset(inplace)``
    (base.py:643-962): memories are buffers of shape (*batch, *sizes)
  * ``BaseNode`` -- charge -> adaptation -> fire -> reset (base.py:967-1124)
This layer only manages tensors; the arithmetic of every neuron/synapse
step runs in the HIP library.
"""

from __future__ import annotations

from collections.abc import Callable, Sequence
from dataclasses import dataclass
from numbers import Number
from typing import Any

import numpy as np
import torch

from .surrogate import Sigmoid


class StepModule:
    """``forward`` dispatches on ``step_mode``: "s" single step, "m" sequence."""

    def supported_step_mode(self) -> tuple[str, ...]:
        return ("s", "m")

    @property
    def step_mode(self) -> str:
        return self._step_mode

    @step_mode.setter
    def step_mode(self, value: str):
        if value not in self.supported_step_mode():
            raise ValueError(
                f"step_mode can only be {self.supported_step_mode()}, but got \"{value}\"!"
            )
        self._step_mode = value

    def single_step_forward(self, x, *args, **kwargs):
        raise NotImplementedError

    def multi_step_forward(self, x_seq, *args, **kwargs):
        raise NotImplementedError

    def forward(self, *args, **kwargs):
        if self.step_mode == "s":
            return self.single_step_forward(*args, **kwargs)
        if self.step_mode == "m":
            return self.multi_step_forward(*args, **kwargs)
        raise ValueError(self.step_mode)


def _broadcastable(a: Sequence[int], b: Sequence[int]) -> bool:
    try:
        torch.broadcast_shapes(tuple(a), tuple(b))
        return True
    except RuntimeError:
        return False


def _fits(value_shape: Sequence[int], target: Sequence[int]) -> bool:
    """``value_shape`` broadcasts *to* ``target`` without growing it."""
    if not _broadcastable(value_shape, target):
        return False
    return tuple(torch.broadcast_shapes(tuple(value_shape), tuple(target))) == tuple(target)


def _expand_to(v: torch.Tensor, sizes: Sequence[int]) -> torch.Tensor:
    sizes = tuple(sizes)
    if tuple(v.shape) == sizes:
        return v
    while v.ndim < len(sizes):
        v = v.unsqueeze(0)
    return v.expand(sizes).clone()


def _is_uniform(t: torch.Tensor, atol: float = 1e-6) -> bool:
    if t.numel() <= 1:
        return True
    first = t.reshape(-1)[0]
    if t.dtype.is_floating_point:
        return bool(torch.allclose(t, first, atol=atol))
    return bool((t == first).all())


class ParamBufferMixin(torch.nn.Module):
    """Define parameters/buffers in their minimal broadcastable form.

    ``def_param(name, val, sizes=..., trainable_param=...)`` registers a
    ``Parameter`` when trainable, else a persistent buffer.  Uniform values
    are saved to ``state_dict`` as 0-d tensors and loaded back either way.
    """

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self._param_shape_mode: dict[str, str] = {}
        self._param_allow_compact: dict[str, bool] = {}

    def _resolve_is_trainable(self, name, trainable_param=None) -> bool:
        if isinstance(trainable_param, bool):
            return trainable_param
        if isinstance(trainable_param, (set, frozenset, list, tuple)):
            return name in trainable_param
        return name in getattr(self, "trainable_param", ())

    def def_param(
        self,
        name: str,
        val,
        *,
        sizes: Sequence[int] | None = None,
        trainable_param=None,
        trainable_shape: str = "auto",
        normalize_to_sizes: bool = False,
        **kwargs,
    ):
        if trainable_shape not in ("auto", "scalar", "full"):
            raise ValueError(
                f"Invalid trainable_shape={trainable_shape!r}. Use 'auto', 'scalar', or 'full'."
            )
        if sizes is None:
            if not hasattr(self, "n_neuron"):
                raise ValueError("sizes is required when module has no n_neuron.")
            sizes = tuple(self.n_neuron)
        sizes = tuple(int(s) for s in sizes)
        trainable = self._resolve_is_trainable(name, trainable_param)
        value = torch.as_tensor(val, **kwargs)
        if kwargs.get("dtype") is None and value.dtype == torch.float64 and not torch.is_tensor(val):
            value = value.to(torch.get_default_dtype())
        numel = int(np.prod(sizes)) if sizes else 1
        if value.ndim == 1 and value.numel() == numel and len(sizes) > 0:
            value = value.reshape(sizes)
        if len(sizes) > 1 and value.ndim == len(sizes) - 1 and tuple(value.shape) == sizes[:-1]:
            value = value[..., None]
        if not _fits(value.shape, sizes):
            raise ValueError(
                f"{name} shape {tuple(value.shape)} is not broadcastable to {sizes}."
            )
        allow_compact = sizes == tuple(getattr(self, "n_neuron", sizes))
        if trainable and trainable_shape == "full":
            value = _expand_to(value, sizes)
        if trainable and trainable_shape == "scalar":
            if not _is_uniform(value):
                raise ValueError(
                    f"{name} with trainable_shape='scalar' must be uniform, but got "
                    f"shape {tuple(value.shape)} with non-uniform values."
                )
            value = value.reshape(-1)[:1].reshape(())
        if normalize_to_sizes and trainable_shape != "scalar":
            value = _expand_to(value, sizes)
        if hasattr(self, name):
            delattr(self, name)
        self._param_shape_mode[name] = trainable_shape if trainable else "auto"
        self._param_allow_compact[name] = allow_compact
        if trainable:
            self.register_parameter(name, torch.nn.Parameter(value))
        else:
            self.register_buffer(name, value, persistent=True)

    def _replace_tensor(self, name: str, value: torch.Tensor) -> None:
        cur = getattr(self, name)
        value = value.to(device=cur.device, dtype=cur.dtype)
        if name in self._parameters:
            rg = bool(self._parameters[name].requires_grad)
            delattr(self, name)
            self.register_parameter(name, torch.nn.Parameter(value, requires_grad=rg))
        else:
            persistent = name not in self._non_persistent_buffers_set
            delattr(self, name)
            self.register_buffer(name, value, persistent=persistent)

    def _save_to_state_dict(self, destination, prefix, keep_vars):
        super()._save_to_state_dict(destination, prefix, keep_vars)
        for name, mode in self._param_shape_mode.items():
            key = prefix + name
            if mode == "full" or not self._param_allow_compact.get(name, True):
                continue
            value = destination.get(key)
            if torch.is_tensor(value) and value.numel() > 1 and _is_uniform(value):
                destination[key] = value.reshape(-1)[:1].reshape(())

    def _load_from_state_dict(self, state_dict, prefix, local_metadata, strict, missing_keys,
                              unexpected_keys, error_msgs):
        for name, mode in self._param_shape_mode.items():
            key = prefix + name
            if key not in state_dict or not hasattr(self, name):
                continue
            loaded = state_dict[key]
            if not torch.is_tensor(loaded):
                continue
            cur_shape = tuple(getattr(self, name).shape)
            new_shape = tuple(loaded.shape)
            if mode == "full" or not self._param_allow_compact.get(name, True):
                if new_shape != cur_shape and _fits(new_shape, cur_shape):
                    state_dict[key] = _expand_to(loaded, cur_shape)
                elif new_shape != cur_shape:
                    self._replace_tensor(name, loaded.detach().clone())
                continue
            if mode == "scalar" and loaded.numel() != 1 and name in self._parameters:
                error_msgs.append(
                    f"{key}: received non-scalar checkpoint tensor for "
                    "trainable_shape='scalar' trainable parameter."
                )
                continue
            if _is_uniform(loaded):
                scalar = loaded.reshape(-1)[:1].reshape(())
                state_dict[key] = scalar
                if cur_shape != ():
                    self._replace_tensor(name, scalar)
            elif new_shape != cur_shape:
                self._replace_tensor(name, loaded.detach().clone())
        super()._load_from_state_dict(state_dict, prefix, local_metadata, strict, missing_keys,
                                      unexpected_keys, error_msgs)


def normalize_n_neuron(n_neuron: int | Sequence[int]) -> tuple[tuple[int, ...], int]:
    shape = (int(n_neuron),) if isinstance(n_neuron, (int, np.integer)) else tuple(n_neuron)
    if len(shape) == 0:
        raise ValueError("n_neuron must contain at least one dimension.")
    return shape, int(np.prod(shape))


def flatten_neuron(x: torch.Tensor, n_neuron: tuple[int, ...], size: int):
    if len(n_neuron) == 1:
        return x, x.shape[:-1]
    lead = x.shape[: -len(n_neuron)]
    return x.reshape(*lead, size), lead


def unflatten_neuron(x: torch.Tensor, leading_shape, n_neuron: tuple[int, ...]):
    if len(n_neuron) == 1:
        return x
    return x.reshape(*leading_shape, *n_neuron)


@dataclass
class ResetValue:
    """How a memory is (re)initialised: value, per-lane sizes, forced dtype."""

    value: Any
    sizes: tuple[int, ...]
    dtype: torch.dtype | None = None
    persistent: bool | None = None
    has_batch: bool = False


def _memory_tensor(rv: ResetValue, batch_size, **fmt) -> torch.Tensor:
    sizes = tuple(rv.sizes)
    if batch_size is not None:
        bs = (int(batch_size),) if isinstance(batch_size, (int, np.integer)) else tuple(batch_size)
        sizes = bs + sizes
    if rv.dtype is not None:
        fmt["dtype"] = rv.dtype
    if callable(rv.value):
        v = torch.as_tensor(rv.value(sizes, batch_size=batch_size)).to(**fmt)
    elif rv.value is None:
        v = torch.empty(sizes, **{k: val for k, val in fmt.items() if val is not None})
    else:
        v = torch.as_tensor(rv.value).to(**{k: val for k, val in fmt.items() if val is not None})
        v = v.detach().clone()
    if tuple(v.shape) != sizes:
        v = _expand_to(v, sizes)
    return v


class MemoryModule(StepModule, torch.nn.Module):
    """Module whose per-lane state tensors ("memories") are managed buffers.

    ``register_memory(name, value, sizes)`` declares a memory; ``init_state``
    (re)allocates every memory with a leading batch shape; ``reset`` restores
    the registered values, in place when ``inplace=True`` (buffer identity and
    address kept, as graph replay requires).
    """

    @property
    def supported_backends(self) -> tuple[str, ...]:
        return ("torch",)

    @property
    def backend(self) -> str:
        return self._backend

    @backend.setter
    def backend(self, value: str):
        if value not in self.supported_backends:
            raise NotImplementedError(f"{value} is not a supported backend of {self._get_name()}!")
        self._backend = value

    def __init__(self):
        super().__init__()
        self._memories_rv: dict[str, ResetValue] = {}
        self._backend = "torch"
        self._step_mode = "s"

    def single_step_forward(self, x, *args, **kwargs):
        raise NotImplementedError

    def multi_step_forward(self, x_seq, *args, **kwargs):
        ys = [self.single_step_forward(x_seq[t], *args, **kwargs) for t in range(x_seq.shape[0])]
        return torch.stack(ys, 0)

    def register_memory(self, name: str, value, sizes, dtype=None, persistent=None):
        assert not hasattr(self, name), f"{name} has been set as a member variable!"
        sizes = (int(sizes),) if isinstance(sizes, (int, np.integer)) else tuple(sizes)
        if isinstance(value, Number):
            value = np.asarray(value)
        elif isinstance(value, Sequence) and not isinstance(value, str):
            assert len(value) != 0, f"Memory cannot be empty list or sequence: got {value}"
            value = np.asarray(value)
        if value is not None and not callable(value):
            if not _fits(np.shape(value), sizes):
                raise ValueError(f"memory {name}: value shape {np.shape(value)} vs sizes {sizes}")
        self._memories_rv[name] = ResetValue(value=value, sizes=sizes, dtype=dtype,
                                             persistent=persistent)
        self.register_buffer(name, None, persistent=False)

    def set_reset_value(self, name: str, value, *, strict: bool = True, **kwargs):
        if isinstance(value, ResetValue):
            self._memories_rv[name] = value
            return
        if isinstance(value, Number):
            value = np.asarray(value)
        rv = self._memories_rv.get(name)
        if rv is None:
            if "sizes" not in kwargs:
                raise ValueError(f"'sizes' is required when registering new memory_rv '{name}'")
            self._memories_rv[name] = ResetValue(value=value, **kwargs)
            return
        sizes = kwargs.get("sizes", rv.sizes)
        if strict and tuple(sizes) != tuple(rv.sizes):
            raise ValueError(f"Memory_rv '{name}' sizes mismatch: existing={rv.sizes}, new={sizes}")
        self._memories_rv[name] = ResetValue(
            value=value, sizes=tuple(sizes), dtype=kwargs.get("dtype", rv.dtype),
            persistent=kwargs.get("persistent", rv.persistent),
            has_batch=kwargs.get("has_batch", rv.has_batch))

    @torch.no_grad()
    def init_state(self, batch_size=None, dtype=None, device=None, persistent=False,
                   skip_mem_name: tuple[str, ...] = ()):
        skip = set(skip_mem_name)
        for key, rv in self._memories_rv.items():
            if key in skip:
                continue
            v = _memory_tensor(rv, batch_size, dtype=rv.dtype or dtype, device=device)
            self.register_buffer(key, v, persistent=bool(rv.persistent or persistent))

    def _batch_dims(self, key):
        buf = self._buffers.get(key)
        sizes = self._memories_rv[key].sizes
        if buf is None or tuple(buf.shape) == tuple(sizes):
            return None
        return tuple(buf.shape[: buf.ndim - len(sizes)])

    @torch.no_grad()
    def reset(self, batch_size=None, dtype=None, device=None,
              skip_mem_name: tuple[str, ...] = (), inplace: bool = False):
        skip = set(skip_mem_name)
        for key, rv in self._memories_rv.items():
            if key in skip:
                continue
            buf = self._buffers.get(key)
            bs = batch_size if batch_size is not None else self._batch_dims(key)
            fmt = {
                "device": device or (buf.device if buf is not None else None),
                "dtype": rv.dtype or dtype or (buf.dtype if buf is not None else None),
            }
            v = _memory_tensor(rv, bs, **fmt)
            if not inplace or buf is None:
                setattr(self, key, v)
                continue
            if tuple(buf.shape) != tuple(v.shape):
                raise ValueError(
                    f"reset(inplace=True) cannot resize memory '{key}' from {tuple(buf.shape)} "
                    f"to {tuple(v.shape)}; call reset()/init_state() to (re)allocate, or keep "
                    "batch_size fixed."
                )
            buf.copy_(v)

    def _sync_memories(self) -> None:
        """Bring every memory buffer up to date.  A memory may be kept
        elsewhere between runs (the fused engine holds a delay history on the
        device) and written back here only when it is read."""

    def memories(self):
        self._sync_memories()
        for name in self._memories_rv:
            yield self._buffers[name]

    def named_memories(self):
        self._sync_memories()
        for name in self._memories_rv:
            yield name, self._buffers[name]

    def detach(self):
        self._sync_memories()
        for key in self._memories_rv:
            buf = self._buffers.get(key)
            if buf is not None:
                self._buffers[key] = buf.detach()

    @property
    def _memories(self):
        self._sync_memories()
        return {name: self._buffers[name] for name in self._memories_rv}

    @_memories.setter
    def _memories(self, value: dict):
        for k, v in value.items():
            assert k in self._memories_rv
            setattr(self, k, v)

    @property
    def memories_rv(self):
        return self._memories_rv

    def set_memories_rv(self, value: dict, strict: bool = False):
        for k, v in value.items():
            assert k in self._memories_rv
            self.set_reset_value(k, v, strict=strict)

    @staticmethod
    def _format_repr_value(value) -> str:
        if value is None:
            return "None"
        if isinstance(value, torch.Tensor):
            return f"{value.item():.4g}" if value.numel() == 1 else f"shape={tuple(value.shape)}"
        return str(value)


class BaseNode(ParamBufferMixin, MemoryModule):
    """Spiking neuron: charge -> adaptation -> fire -> reset each step.

    Constructor arguments and buffers follow btorch/models/base.py:995-1052:
    memory ``v`` (reset value ``v_reset``), parameters ``v_threshold`` and
    ``v_reset`` (trainable when named in ``trainable_param``).
    """

    def __init__(self, n_neuron, v_threshold=1.0, v_reset=0.0, trainable_param=frozenset(),
                 surrogate_function: Callable | None = None, detach_reset: bool = False,
                 hard_reset: bool = False, pre_spike_v: bool = False, step_mode: str = "s",
                 backend: str = "torch", device=None, dtype=None):
        super().__init__()
        self.n_neuron, self.size = normalize_n_neuron(n_neuron)
        self.register_memory("v", v_reset, self.n_neuron)
        self.pre_spike_v = pre_spike_v
        if pre_spike_v:
            self.register_memory("v_pre_spike", v_reset, self.n_neuron, persistent=False)
        self.trainable_param = set(trainable_param)
        fk = {"device": device, "dtype": dtype}
        self.def_param("v_threshold", v_threshold, sizes=self.n_neuron,
                       trainable_param=self.trainable_param, **fk)
        self.def_param("v_reset", v_reset, sizes=self.n_neuron,
                       trainable_param=self.trainable_param, **fk)
        self.detach_reset = detach_reset
        self.surrogate_function = surrogate_function if surrogate_function is not None else Sigmoid()
        self.hard_reset = hard_reset
        self.step_mode = step_mode
        self.backend = backend

    def multi_step_forward(self, x_seq):
        return torch.stack([self.single_step_forward(x) for x in x_seq])

    def extra_repr(self):
        parts = [
            f"n_neuron={self.n_neuron}",
            f"v_threshold={self._format_repr_value(self.v_threshold)}",
            f"v_reset={self._format_repr_value(self.v_reset)}",
            f"surrogate={self.surrogate_function.__class__.__name__}",
        ]
        if self.hard_reset:
            parts.append("hard_reset=True")
        return ", ".join(parts)
