"""Thread-local computation environment (``dt`` and friends).

Same interface as btorch/models/environ.py:24-140: ``context(**kw)`` pushes
values for the duration of a ``with`` block (or a decorated call), ``get``
reads the innermost value, then the global default set by ``set``.
"""

from __future__ import annotations

import functools
import threading
from typing import Any


class _State(threading.local):
    def __init__(self):
        self.stack: dict[str, list[Any]] = {}
        self.defaults: dict[str, Any] = {}


_STATE = _State()


class context:
    """Push ``kwargs`` onto the thread-local environment while active."""

    def __init__(self, **kwargs):
        self.kwargs = kwargs

    def __enter__(self):
        for key, value in self.kwargs.items():
            _STATE.stack.setdefault(key, []).append(value)
        return all()

    def __exit__(self, exc_type, exc, tb):
        for key in self.kwargs:
            _STATE.stack[key].pop()
        return False

    def __call__(self, func):
        @functools.wraps(func)
        def wrapped(*args, **kwargs):
            with self:
                return func(*args, **kwargs)

        return wrapped


def get(key: str, desc: str | None = None):
    """Innermost context value for ``key``, else its global default."""
    values = _STATE.stack.get(key)
    if values:
        return values[-1]
    if key in _STATE.defaults:
        return _STATE.defaults[key]
    msg = (
        f"'{key}' is not found in the context. You can set it by "
        f"`environ.context({key}=value)` locally or `environ.set({key}=value)` globally."
    )
    if desc is not None:
        msg += f"\nDescription: {desc}"
    raise KeyError(msg)


def all() -> dict:  # noqa: A001 - same name as the reference API
    out = {k: v[-1] for k, v in _STATE.stack.items() if v}
    for k, v in _STATE.defaults.items():
        out.setdefault(k, v)
    return out


def set(**kwargs):  # noqa: A001 - same name as the reference API
    _STATE.defaults.update(kwargs)
