"""Process-wide flags (btorch/config.py:13-37): ``BTORCH_SPARSE_BACKEND``
selects the sparse backend name at import, ``config.sparse`` is the registry."""

from __future__ import annotations

import os

from ._sparse_config import HipSparseConfig, SparseBackendConfig, SparseConfigRegistry

SPARSE_BACKEND = os.environ.get("BTORCH_SPARSE_BACKEND")
SPARSE_BACKEND = SPARSE_BACKEND.lower() if SPARSE_BACKEND else None
sparse = SparseConfigRegistry(SPARSE_BACKEND)

__all__ = ["SPARSE_BACKEND", "sparse", "HipSparseConfig", "SparseBackendConfig"]
