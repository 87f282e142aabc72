"""Spike history for integer synaptic delays (btorch/models/history.py:28-358).

Memory ``history`` has shape (*batch, max_delay, *n_neuron).  Circular mode
writes the newest spikes at ``_cursor`` and reads delay d from slot
(cursor - 1 - d) mod max_delay; cat mode keeps delay d at index d.
``get_flattened(D)`` interleaves neurons and delays, index n*D + d, matching
the rows of ``expand_conn_for_delays``.

Inside ``RecurrentNN`` the history is replaced by a ring of compacted spike
lists on the device (no (B, N*D) matrix is ever materialised).  Between runs
the ring stays on the device and the next run continues from it; the dense
memory is written back from the ring only when something reads it (the
``history`` or ``_cursor`` attribute, ``memories()``, ``state_dict``,
``.to()``), so a run followed by another run never pays for the (B, D, N)
write.  Assigning the memory, ``set_cursor``, ``reset`` and ``init_state``
end that deferral first, so the module's own value always wins.
"""

from __future__ import annotations

import torch

from .base import MemoryModule, flatten_neuron, normalize_n_neuron


class SpikeHistory(MemoryModule):
    """Rolling buffer of the last ``max_delay_steps`` spike patterns."""

    def __init__(self, n_neuron, max_delay_steps: int = 5, use_circular_buffer: bool = True):
        super().__init__()
        self.n_neuron, self.size = normalize_n_neuron(n_neuron)
        self.max_delay = int(max_delay_steps)
        self.use_circular_buffer = use_circular_buffer
        self.register_memory("history", torch.tensor(0.0, dtype=torch.float32),
                             (self.max_delay, *self.n_neuron))
        self._cursor_host = 0
        # deferred write-back of the memory (see the module doc): a callable
        # that fills the ``history`` buffer and the device cursor, or None
        self._deferred = None
        if use_circular_buffer:
            self.register_buffer("_cursor", torch.tensor(0, dtype=torch.long), persistent=False)

    # -- deferred write-back ------------------------------------------------
    def _sync_memories(self) -> None:
        fn = self.__dict__.get("_deferred")
        if fn is not None:
            self.__dict__["_deferred"] = None
            fn(self)

    def defer_writeback(self, fn, cursor: int) -> None:
        """The memory now lives elsewhere: ``fn(self)`` writes it (and the
        device cursor) back when it is read.  ``cursor`` is the host cursor
        after the run that produced it."""
        self.__dict__["_deferred"] = fn
        self._cursor_host = int(cursor) % self.max_delay

    def deferred_writeback(self):
        """The pending write-back callable, or None."""
        return self.__dict__.get("_deferred")

    def discard_writeback(self) -> None:
        self.__dict__["_deferred"] = None

    def __getattr__(self, name):
        if name in ("history", "_cursor"):
            self._sync_memories()
        return super().__getattr__(name)

    def __setattr__(self, name, value):
        if name in ("history", "_cursor") and "_deferred" in self.__dict__:
            # an assignment replaces what the deferred write-back would bring
            if name == "_cursor":
                self._sync_memories()
            else:
                self.__dict__["_deferred"] = None
        super().__setattr__(name, value)

    def named_buffers(self, *args, **kwargs):
        self._sync_memories()
        return super().named_buffers(*args, **kwargs)

    def _apply(self, fn, recurse=True):
        self._sync_memories()
        return super()._apply(fn, recurse=recurse)

    def _save_to_state_dict(self, destination, prefix, keep_vars):
        self._sync_memories()
        super()._save_to_state_dict(destination, prefix, keep_vars)

    def _load_from_state_dict(self, state_dict, prefix, *args, **kwargs):
        self._sync_memories()
        super()._load_from_state_dict(state_dict, prefix, *args, **kwargs)

    # -- cursor ---------------------------------------------------------------
    @property
    def cursor(self) -> int:
        """Host copy of the circular cursor (0 in cat mode)."""
        return self._cursor_host if self.use_circular_buffer else 0

    def set_cursor(self, value: int) -> None:
        self._sync_memories()
        self._cursor_host = int(value) % self.max_delay
        if self.use_circular_buffer:
            self._cursor.fill_(self._cursor_host)

    def init_state(self, batch_size=None, dtype=None, device=None, persistent=False,
                   skip_mem_name=()):
        if "history" in tuple(skip_mem_name):
            self._sync_memories()
        self.discard_writeback()
        super().init_state(batch_size, dtype, device, persistent, tuple(skip_mem_name))
        if self.use_circular_buffer:
            self._cursor = torch.tensor(0, dtype=torch.long, device=device)
        self._cursor_host = 0

    def reset(self, batch_size=None, dtype=None, device=None, skip_mem_name=(), inplace=False):
        if "history" in tuple(skip_mem_name):
            self._sync_memories()
        self.discard_writeback()
        super().reset(batch_size, dtype, device, tuple(skip_mem_name), inplace=inplace)
        if self.use_circular_buffer:
            if inplace:
                self._cursor.zero_()
            else:
                self._cursor = torch.tensor(0, dtype=torch.long, device=self._cursor.device)
        self._cursor_host = 0

    def _delay_axis(self) -> int:
        return self.history.ndim - 1 - len(self.n_neuron)

    def update(self, spike: torch.Tensor) -> "SpikeHistory":
        if self.history is None:
            raise RuntimeError(
                "SpikeHistory buffer not initialized. Call init_state() before first update()."
            )
        spike = spike.to(self.history.dtype)
        ax = self._delay_axis()
        if self.use_circular_buffer:
            self.history.select(ax, self._cursor_host).copy_(spike)
            self.set_cursor(self._cursor_host + 1)
        else:
            hist = self.history
            self.history = torch.cat([spike.unsqueeze(ax), hist.narrow(ax, 0, self.max_delay - 1)],
                                     dim=ax)
        return self

    def slot_of_delay(self, delay_steps: int) -> int:
        if self.use_circular_buffer:
            return (self._cursor_host - 1 - delay_steps) % self.max_delay
        return delay_steps

    def get_delay(self, delay_steps: int) -> torch.Tensor:
        if delay_steps >= self.max_delay:
            raise IndexError(
                f"delay_steps ({delay_steps}) >= max_delay ({self.max_delay}). "
                "Increase max_delay_steps or use smaller delays."
            )
        return self.history.select(self._delay_axis(), self.slot_of_delay(delay_steps))

    def get_recent(self, n_steps: int) -> list[torch.Tensor]:
        if n_steps > self.max_delay:
            raise IndexError(f"n_steps ({n_steps}) > max_delay ({self.max_delay})")
        return [self.get_delay(d) for d in range(n_steps)]

    def get_flattened(self, n_delays: int) -> torch.Tensor:
        if n_delays > self.max_delay:
            raise IndexError(f"n_delays ({n_delays}) > max_delay ({self.max_delay})")
        stacked = torch.stack([self.get_delay(d) for d in range(n_delays)], dim=-1)
        # (*batch, *n_neuron, D) -> (*batch, size*D), index n*D + d
        lead = stacked.shape[: stacked.ndim - 1 - len(self.n_neuron)]
        return stacked.reshape(*lead, self.size * n_delays)

    def extra_repr(self) -> str:
        mode = "circular" if self.use_circular_buffer else "cat"
        return f"n_neuron={self.n_neuron}, max_delay={self.max_delay}, {mode}"


__all__ = ["SpikeHistory", "flatten_neuron"]
