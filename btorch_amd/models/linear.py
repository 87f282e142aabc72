"""Sparse connectivity linear: ``y = x @ conn`` for a fixed sparse ``conn``.

Mirrors btorch/models/linear.py:222-591 (``BaseSparseConn``/``SparseConn``):
constructor ``SparseConn(conn, bias=None, enforce_dale=True,
sparse_backend=None, sparse_config=None, device=None, dtype=None)``; buffer
``indices`` (2, E) int64 holding the coalesced (dst, src) pairs and
Parameter ``magnitude`` (E,) in that order; ``prepare_sparse``/
``finish_sparse``; ``get_sparse_matrix``; ``constrain``.

The product runs in the HIP library as an event-driven push over a
source-major CSR (``_csr_*`` non-persistent buffers, rebuilt on load): the
input is compacted to its nonzero sources per lane, and each active row is
streamed to its destinations.  The default ordered mode sums every
destination in ascending source order with fmaf, which equals the CPU
``torch.sparse.mm`` of the reference's native branch bit for bit.
"""

from __future__ import annotations

from typing import Any

import numpy as np
import scipy.sparse
import torch
import torch.nn as nn

from .. import _kernels, _lib
from .. import config as btorch_config
from .._sparse_config import HipSparseConfig, SparseBackendConfig


def _resolve_sparse_backend(name: str | None) -> str:
    if name is None:
        name = btorch_config.sparse.selected or "hip"
    name = name.lower()
    if name == "torch_sparse":
        name = "triton"
    if name not in btorch_config.sparse.names():
        raise ValueError(
            f"sparse_backend must be one of {btorch_config.sparse.names()}, got '{name}'."
        )
    return name


def available_sparse_backends(*, include_experimental: bool = False) -> list[str]:
    """Backend names accepted by ``SparseConn`` (all run the HIP product)."""
    return btorch_config.sparse.names()


class _SparseProduct(torch.autograd.Function):
    """y = x @ W with W packed in CSR order; grads for x and the packed weights
    (autograd carries the latter back to ``magnitude`` through the gather)."""

    @staticmethod
    def forward(ctx, x2d, w_csr, conn: "BaseSparseConn"):
        active, counts = _kernels.compact_dense(x2d)
        ws = conn._workspace_for(x2d.shape[0])
        cfg: HipSparseConfig = conn.sparse_config
        w_dst = conn.effective_weight_f32()
        out = _kernels.spmv_event(conn._layout, w_csr, x2d, active, counts,
                                  ordered=cfg.ordered, slots=cfg.slots, workspace=ws,
                                  w_dst=w_dst)
        ctx.conn = conn
        ctx.save_for_backward(x2d, w_csr, active, counts)
        return out

    @staticmethod
    def backward(ctx, g):
        x2d, w_csr, active, counts = ctx.saved_tensors
        layout = ctx.conn._layout
        gx = gw = None
        if ctx.needs_input_grad[0]:
            gx = _kernels.spmv_transpose(layout, w_csr, g)
        if ctx.needs_input_grad[1]:
            gw = torch.zeros(layout.nnz, dtype=torch.float32, device=g.device)
            _kernels.spmv_weight_grad(layout, x2d, active, counts, g, gw)
        return gx, gw, None


class BaseSparseConn(nn.Module):
    """Fixed sparse connection ``conn`` of shape (n_src, n_dst); see module doc."""

    def __init__(self, conn, bias=None, sparse_backend: str | None = None,
                 sparse_config: SparseBackendConfig | dict[str, Any] | None = None,
                 device=None, dtype=None):
        super().__init__()
        self.sparse_backend = _resolve_sparse_backend(sparse_backend)
        self.sparse_config = btorch_config.sparse.resolve(self.sparse_backend, sparse_config)
        self._prepared_sparse_weight: torch.Tensor | None = None
        self._sparse_prepare_depth = 0
        self._workspaces: dict[int, torch.Tensor] = {}
        # no-grad packed (CSR order) and destination-order weights, reused
        # while the tensors they derive from are unchanged (weight_key)
        self._weight_cache: dict[str, tuple] = {}
        prebuilt = None
        if isinstance(conn, _kernels.BuiltConn):
            # built on the device by btx_conn_build (see ``from_edges``)
            self.in_features, self.out_features = conn.shape
            indices = conn.indices
            value = conn.values.to(dtype or torch.float32)
            prebuilt = conn.layout
        elif isinstance(conn, torch.Tensor) and conn.is_sparse:
            # torch COO (n_src, n_dst): transpose and coalesce on its own device
            # (sorted by (dst, src), duplicates summed), e.g. on the GPU
            self.in_features, self.out_features = conn.shape
            t = conn.t().coalesce()
            indices = t.indices().to(device=device, dtype=torch.long)
            value = t.values().to(device=device, dtype=dtype or torch.float32)
        else:
            if not isinstance(conn, scipy.sparse.coo_array):
                conn = scipy.sparse.coo_array(conn)
            self.in_features, self.out_features = conn.shape
            conn_t = conn.T.tocoo()
            conn_t.sum_duplicates()  # sorted by (dst, src)
            indices = torch.stack([
                torch.as_tensor(conn_t.row, dtype=torch.long),
                torch.as_tensor(conn_t.col, dtype=torch.long),
            ]).to(device)
            value = torch.as_tensor(conn_t.data, dtype=dtype or torch.float32).to(device)
        self.shape = (self.in_features, self.out_features)
        self.register_buffer("indices", indices)
        self._rebuild_layout(prebuilt)
        self.bias = nn.Parameter(bias) if bias is not None else None
        self._init_weights(value)

    # -- layout --------------------------------------------------------------
    @classmethod
    def from_edges(cls, pre: torch.Tensor, post: torch.Tensor, weight: torch.Tensor,
                   delay: torch.Tensor | None = None, n_delay_bins: int = 1,
                   shape: tuple[int, int] | None = None, **kwargs):
        """The connection of the delay-expanded matrix of edges (pre, post,
        delay, weight), built on the device (``btx_conn_build``): equal to
        ``cls(expand_conn_for_delays(coo((weight, (pre, post))), delay,
        n_delay_bins))`` without the host round trip.  ``shape`` is
        (n_pre, n_post); by default the largest indices + 1."""
        n_pre, n_post = shape if shape is not None else (None, None)
        built = _kernels.build_conn(pre, post, weight, delay, n_delay_bins, n_pre, n_post)
        return cls(built, **kwargs)

    def _rebuild_layout(self, lay: _kernels.CsrLayout | None = None) -> None:
        if lay is None:
            lay = _kernels.build_csr(self.indices[0], self.indices[1], self.in_features,
                                     self.out_features)
        for name, t in (("_csr_row_ptr", lay.row_ptr), ("_csr_col", lay.col),
                        ("_csr_perm", lay.perm), ("_csc_ptr", lay.csc_ptr),
                        ("_csc_src", lay.csc_src)):
            if name in self._buffers:
                setattr(self, name, t)
            else:
                self.register_buffer(name, t, persistent=False)
        self._workspaces = {}

    @property
    def _layout(self) -> _kernels.CsrLayout:
        return _kernels.CsrLayout(self._csr_row_ptr, self._csr_col, self._csr_perm,
                                  self.in_features, self.out_features, self._csc_ptr,
                                  self._csc_src)

    @property
    def nnz(self) -> int:
        return int(self.indices.shape[1])

    def _workspace_for(self, lanes: int) -> torch.Tensor:
        cfg: HipSparseConfig = self.sparse_config
        dev = self._csr_col.device
        need = _lib.load().btx_spmv_workspace_bytes(lanes, self.out_features, cfg.slots,
                                                    int(cfg.ordered))
        key = (int(cfg.ordered), cfg.slots, str(dev))
        ws = self._workspaces.get(key)
        if ws is None or ws.numel() < need:
            ws = torch.zeros(need, dtype=torch.uint8, device=dev)
            self._workspaces[key] = ws
        return ws

    def _apply(self, fn, recurse=True):
        self._prepared_sparse_weight = None
        self._sparse_prepare_depth = 0
        self._workspaces = {}
        self._weight_cache = {}
        return super()._apply(fn, recurse=recurse)

    def _load_from_state_dict(self, state_dict, prefix, local_metadata, strict, missing_keys,
                              unexpected_keys, error_msgs):
        super()._load_from_state_dict(state_dict, prefix, local_metadata, strict, missing_keys,
                                      unexpected_keys, error_msgs)
        self._prepared_sparse_weight = None
        self._sparse_prepare_depth = 0
        self._weight_cache = {}
        self._rebuild_layout()

    # -- weights -------------------------------------------------------------
    def _init_weights(self, value: torch.Tensor):
        raise NotImplementedError

    def _get_effective_weight(self) -> torch.Tensor:
        raise NotImplementedError

    _LAYOUT_BUFFERS = ("_csr_row_ptr", "_csr_col", "_csr_perm", "_csc_ptr", "_csc_src")

    def weight_key(self) -> tuple:
        """Identity, storage and in-place version of every parameter and
        buffer the effective weight derives from.  Writes through
        ``tensor.data`` keep the version, so code that edits weights that
        way assigns ``.data`` (a new storage) or calls ``touch_weights()``."""
        parts = [self._weight_epoch]
        for name, t in self._parameters.items():
            parts.append((name, None if t is None else (id(t), t._version, t.data_ptr())))
        for name, t in self._buffers.items():
            if name not in self._LAYOUT_BUFFERS:
                parts.append((name, None if t is None else (id(t), t._version, t.data_ptr())))
        return tuple(parts)

    _weight_epoch = 0

    def touch_weights(self) -> None:
        """Drop the cached packed weights (after an edit ``weight_key`` cannot see)."""
        self._weight_epoch += 1
        self._weight_cache = {}

    def _cached_weight(self, kind: str, make) -> torch.Tensor:
        key = (self.weight_key(), id(self._csr_perm))
        hit = self._weight_cache.get(kind)
        if hit is not None and hit[0] == key:
            return hit[1]
        with torch.no_grad():
            w = make()
        self._weight_cache[kind] = (key, w)
        return w

    def effective_weight_f32(self) -> torch.Tensor:
        """Detached float32 effective weights in the coalesced (dst, src) order."""
        return self._cached_weight(
            "dst", lambda: self._get_effective_weight().detach().to(torch.float32).contiguous())

    def _pack(self) -> torch.Tensor:
        if torch.is_grad_enabled():
            w = self._get_effective_weight()
            return w.index_select(0, self._csr_perm).contiguous()
        return self._cached_weight(
            "csr", lambda: self._get_effective_weight().index_select(0, self._csr_perm)
            .contiguous())

    def packed_weight(self) -> torch.Tensor:
        """Effective weights in CSR order (``magnitude[perm]``)."""
        if self._prepared_sparse_weight is not None:
            return self._prepared_sparse_weight
        return self._pack()

    def prepare_sparse(self, batch_size: int | None = None) -> None:
        """Pack weights once for a multi-step run (reference-counted)."""
        if self._prepared_sparse_weight is None:
            self._prepared_sparse_weight = self._pack()
        if batch_size is not None and self._csr_col.is_cuda:
            self._workspace_for(int(batch_size))
        self._sparse_prepare_depth += 1

    def finish_sparse(self) -> None:
        if self._sparse_prepare_depth <= 0:
            raise RuntimeError("finish_sparse called without prepare_sparse.")
        self._sparse_prepare_depth -= 1
        if self._sparse_prepare_depth == 0:
            self._prepared_sparse_weight = None

    def get_sparse_matrix(self) -> torch.Tensor:
        idx = torch.stack([self.indices[1], self.indices[0]], dim=0)
        return torch.sparse_coo_tensor(idx, self._get_effective_weight(),
                                       size=(self.in_features, self.out_features),
                                       is_coalesced=True)

    # -- forward -------------------------------------------------------------
    def forward(self, x: torch.Tensor) -> torch.Tensor:
        _lib.require_hip(x, self._csr_col)
        no_batch = x.ndim == 1
        if no_batch:
            x = x[None, :]
        lead = x.shape[:-1]
        x2d = x.reshape(-1, x.shape[-1])
        if x2d.dtype != torch.float32:
            raise TypeError("The HIP sparse backend currently supports float32 only.")
        x2d = x2d.contiguous()
        out = _SparseProduct.apply(x2d, self.packed_weight(), self)
        out = out.reshape(*lead, self.out_features)
        if no_batch:
            out = out[0]
        if self.bias is not None:
            out = out + self.bias
        return out


class SparseConn(BaseSparseConn):
    """Sparse linear with one learnable magnitude per connection; optional
    Dale's law via a fixed sign and ReLU projection in ``constrain``."""

    def __init__(self, conn, bias=None, enforce_dale: bool = True,
                 sparse_backend: str | None = None,
                 sparse_config: SparseBackendConfig | dict[str, Any] | None = None,
                 device=None, dtype=None):
        self.enforce_dale = enforce_dale
        super().__init__(conn, bias=bias, sparse_backend=sparse_backend,
                         sparse_config=sparse_config, device=device, dtype=dtype)

    def _init_weights(self, value: torch.Tensor):
        if self.enforce_dale:
            self.register_buffer("initial_sign", torch.sign(value), persistent=False)
        self.magnitude = nn.Parameter(value)

    def _get_effective_weight(self) -> torch.Tensor:
        return self.magnitude

    @torch.no_grad()
    def constrain(self, *args, **kwargs):
        if self.enforce_dale:
            self.magnitude.data = (self.magnitude * self.initial_sign).relu() * self.initial_sign


class SparseConstrainedConn(BaseSparseConn):
    """Sparse linear whose connections share one learnable magnitude per
    constraint group (btorch/models/linear.py:594-867): the effective weight
    of a connection is ``initial_weight * magnitude[group]``.  ``constraint``
    has the shape of ``conn`` and holds 1-based group ids.  The product and
    its backward are the same HIP kernels as ``SparseConn``'s; the per-edge
    weight gradient reaches ``magnitude`` through the gather, which sums it
    by group."""

    def __init__(self, conn, constraint, enforce_dale: bool = True, bias=None,
                 sparse_backend: str | None = None,
                 sparse_config: SparseBackendConfig | dict[str, Any] | None = None,
                 device=None, dtype=None, persist_initial_weight: bool = False):
        self.enforce_dale = enforce_dale
        self.persist_initial_weight = persist_initial_weight
        self.constraint_info = None
        constraint = scipy.sparse.coo_array(constraint).T.tocoo()
        constraint.eliminate_zeros()
        self._constraint_matrix = constraint
        super().__init__(conn, bias=bias, sparse_backend=sparse_backend,
                         sparse_config=sparse_config, device=device, dtype=dtype)

    def _init_weights(self, value: torch.Tensor):
        self.register_buffer("initial_weight", value, persistent=self.persist_initial_weight)
        num_groups = int(self._constraint_matrix.data.max())
        self.magnitude = nn.Parameter(torch.empty(num_groups, device=value.device))
        idx = self._precompute_scatter_indices(self._constraint_matrix)
        self.register_buffer("_constraint_scatter_indices",
                             torch.as_tensor(idx, dtype=torch.long, device=value.device),
                             persistent=False)
        self.reset_parameters()

    def reset_parameters(self) -> None:
        nn.init.ones_(self.magnitude)

    def _precompute_scatter_indices(self, constraint) -> np.ndarray:
        """0-based group of every connection, in the coalesced (dst, src)
        order (the reference matches the (row, col) pairs by a left join)."""
        idx = self.indices.cpu().numpy()
        n_src = self.in_features
        key = idx[0].astype(np.int64) * n_src + idx[1].astype(np.int64)
        ckey = constraint.row.astype(np.int64) * n_src + constraint.col.astype(np.int64)
        order = np.argsort(ckey, kind="stable")
        ckey, group = ckey[order], constraint.data[order].astype(np.int64)
        pos = np.searchsorted(ckey, key)
        hit = pos < ckey.size
        hit[hit] = ckey[pos[hit]] == key[hit]
        if not hit.all():
            raise AssertionError("Constraint missing for some connections.")
        return group[pos] - 1

    def _get_effective_weight(self) -> torch.Tensor:
        return self.initial_weight * self.magnitude[self._constraint_scatter_indices]

    @torch.no_grad()
    def constrain(self, *args, **kwargs):
        if self.enforce_dale:
            self.magnitude.data = self.magnitude.relu()

    @classmethod
    def from_hetersynapse(cls, conn, constraint, receptor_type_index, enforce_dale: bool = True,
                          bias=None, sparse_backend=None, sparse_config=None, device=None,
                          dtype=None, persist_initial_weight: bool = False):
        """From the output of make_hetersynapse_constrained_conn, keeping the
        receptor table for ``set_group_magnitude(receptor_pair=...)``."""
        inst = cls(conn, constraint, enforce_dale=enforce_dale, bias=bias,
                   sparse_backend=sparse_backend, sparse_config=sparse_config, device=device,
                   dtype=dtype, persist_initial_weight=persist_initial_weight)
        inst.constraint_info = {"receptor_type_index": receptor_type_index}
        return inst

    def get_group_info(self, include_weights: bool = False):
        """One row per group: group_id, num_connections, current_magnitude
        (and the mean/std of the initial weights)."""
        import pandas as pd

        n_groups = len(self.magnitude)
        gid = self._constraint_scatter_indices
        rows = []
        for g in range(n_groups):
            mask = gid == g
            info = {"group_id": g, "num_connections": int(mask.sum()),
                    "current_magnitude": float(self.magnitude[g].detach())}
            if include_weights:
                w = self.initial_weight[mask]
                info["mean_initial_weight"] = float(w.mean())
                info["std_initial_weight"] = float(w.std(unbiased=False))
            rows.append(info)
        df = pd.DataFrame(rows)
        if self.constraint_info is not None:
            rti = self.constraint_info["receptor_type_index"]
            if len(rti) == n_groups:
                df = pd.concat([df, rti.reset_index(drop=True)], axis=1)
        return df

    def set_group_magnitude(self, group_id: int | None = None,
                            receptor_pair: tuple[str, str] | None = None,
                            value: float | torch.Tensor = 1.0):
        if (group_id is None) == (receptor_pair is None):
            raise ValueError("Exactly one of group_id or receptor_pair must be specified")
        if receptor_pair is not None:
            if self.constraint_info is None:
                raise ValueError("receptor_pair can only be used if created via "
                                 "from_hetersynapse")
            rti = self.constraint_info["receptor_type_index"]
            if "pre_receptor_type" not in rti.columns:
                raise ValueError("receptor_pair requires neuron mode "
                                 "(pre_receptor_type, post_receptor_type)")
            table = rti.set_index(["pre_receptor_type", "post_receptor_type"])
            group_id = int(table.loc[tuple(receptor_pair), "receptor_index"])
        if isinstance(value, (int, float)):
            value = torch.tensor(value, dtype=self.magnitude.dtype, device=self.magnitude.device)
        self.magnitude.data[group_id] = value

    def get_weights_by_group(self) -> dict[int, torch.Tensor]:
        gid = self._constraint_scatter_indices
        return {g: self.initial_weight[gid == g] * self.magnitude[g]
                for g in range(len(self.magnitude))}


__all__ = ["BaseSparseConn", "SparseConn", "SparseConstrainedConn", "available_sparse_backends"]
