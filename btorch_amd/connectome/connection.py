"""Delay expansion of a connection matrix (btorch/connectome/connection.py:973-1027).

Row ``pre*n_delay_bins + clip(delay, 0, n_delay_bins-1)`` of the expanded
(n_pre*n_delay_bins, n_post) matrix carries synapse (pre, post); the format
must match the reference so that ``SpikeHistory.get_flattened`` and the
fused engine's virtual rows index the same entries.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse


def expand_conn_for_delays(conn, delays: np.ndarray, n_delay_bins: int = 5):
    delays = np.asarray(delays)
    if np.any(delays < 0):
        raise ValueError("delays must be non-negative")
    coo = scipy.sparse.coo_array(conn)
    n_pre, n_post = coo.shape
    bins = np.clip(delays, 0, n_delay_bins - 1).astype(np.int64)
    rows = coo.row.astype(np.int64) * n_delay_bins + bins
    return scipy.sparse.coo_array((coo.data, (rows, coo.col)),
                                  shape=(n_pre * n_delay_bins, n_post))
