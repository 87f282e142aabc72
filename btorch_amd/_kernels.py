"""Tensor-level wrappers over the C ABI of libbtx.so.

Each wrapper validates shapes/devices on the host (so no kernel is launched
on operands it does not expect), allocates outputs with torch and enqueues
the kernel on the current stream.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

from . import _lib


def _i32(t):
    return t.dtype == torch.int32 and t.is_contiguous()


@dataclass
class CsrLayout:
    """Source-major CSR of a connection matrix (n_src rows, n_dst columns).

    ``perm[k]`` is the position, in the reference's coalesced (dst, src)
    order (btorch/models/linear.py:276-291), of CSR edge k; packed weights
    are ``magnitude[perm]``.
    """

    row_ptr: torch.Tensor  # (n_src + 1,) int64
    col: torch.Tensor      # (nnz,) int32, destinations, ascending per row
    perm: torch.Tensor     # (nnz,) int64
    n_src: int
    n_dst: int
    csc_ptr: torch.Tensor | None = None  # (n_dst + 1,) int64, dst-major row pointer
    csc_src: torch.Tensor | None = None  # (nnz,) int32, sources in coalesced order

    @property
    def nnz(self) -> int:
        return int(self.col.numel())

    def to(self, device) -> "CsrLayout":
        mv = (lambda t: None if t is None else t.to(device))
        return CsrLayout(self.row_ptr.to(device), self.col.to(device), self.perm.to(device),
                         self.n_src, self.n_dst, mv(self.csc_ptr), mv(self.csc_src))


def build_csr(dst: torch.Tensor, src: torch.Tensor, n_src: int, n_dst: int) -> CsrLayout:
    """Source-major CSR from coalesced (dst, src) index vectors.

    A stable sort by source keeps destinations ascending inside each row
    (the input is sorted by (dst, src)).  Runs on the indices' device.
    """
    if n_src >= 2**31 or n_dst >= 2**31:
        raise ValueError("connection dimensions must fit int32 indices")
    dst = dst.to(torch.int64)
    src = src.to(torch.int64)
    perm = torch.argsort(src, stable=True)
    col = dst[perm].to(torch.int32).contiguous()
    counts = torch.bincount(src, minlength=n_src)
    row_ptr = torch.zeros(n_src + 1, dtype=torch.int64, device=dst.device)
    if counts.numel():
        row_ptr[1:] = torch.cumsum(counts, 0)
    dcounts = torch.bincount(dst, minlength=n_dst)
    csc_ptr = torch.zeros(n_dst + 1, dtype=torch.int64, device=dst.device)
    if dcounts.numel():
        csc_ptr[1:] = torch.cumsum(dcounts, 0)
    return CsrLayout(row_ptr.contiguous(), col, perm.contiguous(), n_src, n_dst,
                     csc_ptr.contiguous(), src.to(torch.int32).contiguous())


@dataclass
class BuiltConn:
    """A connection built on the device by ``build_conn``: coalesced
    (dst, src) ``indices``, their ``values``, the CSR ``layout`` and the
    matrix ``shape`` (n_src, n_dst)."""

    indices: torch.Tensor
    values: torch.Tensor
    layout: CsrLayout
    shape: tuple[int, int]


def build_conn(pre: torch.Tensor, post: torch.Tensor, w: torch.Tensor,
               delay: torch.Tensor | None = None, n_delay_bins: int = 1,
               n_pre: int | None = None, n_post: int | None = None) -> BuiltConn:
    """Delay-expanded, coalesced connectivity and its CSR layout built on the
    device by ``btx_conn_build`` (the reference's expand_conn_for_delays,
    btorch/connectome/connection.py:973-1027, and BaseSparseConn's coalescing,
    btorch/models/linear.py:276-291).  Virtual source row of an edge:
    pre*n_delay_bins + clip(delay, 0, n_delay_bins - 1)."""
    _lib.require_hip(pre, post, w)
    if delay is not None:
        _lib.require_hip(delay)
    nnz = int(pre.numel())
    if post.numel() != nnz or w.numel() != nnz or (delay is not None and delay.numel() != nnz):
        raise ValueError("pre, post, w (and delay) must have one entry per edge")
    if n_delay_bins < 1:
        raise ValueError("n_delay_bins must be at least 1")
    dev = pre.device
    pre = pre.reshape(-1).to(torch.int64).contiguous()
    post = post.reshape(-1).to(torch.int64).contiguous()
    w = w.reshape(-1).to(torch.float32).contiguous()
    if delay is not None:
        delay = delay.reshape(-1).to(torch.int64).contiguous()
    if nnz:
        lo = torch.stack([pre.min(), post.min()] + ([delay.min()] if delay is not None else []))
        hi = torch.stack([pre.max(), post.max()])
        lo, hi = lo.tolist(), hi.tolist()
        if min(lo[:2]) < 0:
            raise ValueError("pre and post indices must be non-negative")
        if delay is not None and lo[2] < 0:
            raise ValueError("delays must be non-negative")
        n_pre = hi[0] + 1 if n_pre is None else n_pre
        n_post = hi[1] + 1 if n_post is None else n_post
        if hi[0] >= n_pre or hi[1] >= n_post:
            raise ValueError("an edge lies outside the (n_pre, n_post) matrix")
    n_pre = int(n_pre or 0)
    n_post = int(n_post or 0)
    n_src = n_pre * n_delay_bins
    m = max(nnz, 1)
    i64 = dict(dtype=torch.int64, device=dev)
    i32 = dict(dtype=torch.int32, device=dev)
    out_dst, out_src, perm = (torch.empty(m, **i64) for _ in range(3))
    out_w = torch.empty(m, dtype=torch.float32, device=dev)
    col, csc_src = torch.empty(m, **i32), torch.empty(m, **i32)
    row_ptr = torch.empty(n_src + 1, **i64)
    csc_ptr = torch.empty(n_post + 1, **i64)
    nnz_out = torch.zeros(1, **i64)
    lib = _lib.load()
    ws = torch.empty(lib.btx_conn_build_workspace_bytes(nnz), dtype=torch.uint8, device=dev)
    ptr = (lambda t: None if t is None else t.data_ptr())
    _lib.check(lib.btx_conn_build(ptr(pre), ptr(post), ptr(delay), ptr(w), nnz, n_pre, n_post,
                                  int(n_delay_bins), ptr(out_dst), ptr(out_src), ptr(out_w),
                                  ptr(nnz_out), ptr(row_ptr), ptr(col), ptr(perm), ptr(csc_ptr),
                                  ptr(csc_src), ptr(ws), ws.numel(), _lib.stream_ptr()),
               "btx_conn_build")
    u = int(nnz_out.item())
    layout = CsrLayout(row_ptr, col[:u].contiguous(), perm[:u].contiguous(), n_src, n_post,
                       csc_ptr, csc_src[:u].contiguous())
    indices = torch.stack([out_dst[:u], out_src[:u]])
    return BuiltConn(indices, out_w[:u].contiguous(), layout, (n_src, n_post))


def compact_dense(x2d: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """Ascending nonzero columns of every row of ``x2d`` (lanes, n)."""
    _lib.require_hip(x2d)
    if x2d.dtype != torch.float32:
        raise TypeError("compaction expects float32 input")
    x2d = x2d.contiguous()
    lanes, n = x2d.shape
    lib = _lib.load()
    idx = torch.empty((lanes, max(n, 1)), dtype=torch.int32, device=x2d.device)
    cnt = torch.empty((lanes,), dtype=torch.int32, device=x2d.device)
    tiles = torch.empty((lanes, lib.btx_compact_tiles(n)), dtype=torch.int32, device=x2d.device)
    _lib.check(lib.btx_compact_dense(x2d.data_ptr(), lanes, n, n, idx.data_ptr(), idx.shape[1],
                                     cnt.data_ptr(), tiles.data_ptr(), _lib.stream_ptr()),
               "btx_compact_dense")
    return idx, cnt


def compact_bits(words: torch.Tensor, n: int) -> tuple[torch.Tensor, torch.Tensor]:
    """Ascending set-bit positions of every row of ``words`` (lanes, ceil(n/64)) int64."""
    _lib.require_hip(words)
    lanes = words.shape[0]
    if words.shape[1] != (n + 63) // 64 or not words.is_contiguous():
        raise ValueError("words must be contiguous (lanes, ceil(n/64))")
    idx = torch.empty((lanes, max(n, 1)), dtype=torch.int32, device=words.device)
    cnt = torch.empty((lanes,), dtype=torch.int32, device=words.device)
    _lib.check(_lib.load().btx_compact_bits(words.data_ptr(), lanes, n, words.shape[1],
                                            idx.data_ptr(), idx.shape[1], cnt.data_ptr(),
                                            _lib.stream_ptr()), "btx_compact_bits")
    return idx, cnt


def pack_rows(x: torch.Tensor) -> torch.Tensor:
    """(rows, n) float32 -> (rows, ceil(n/64)) int64 spike words (x != 0)."""
    _lib.require_hip(x)
    rows, n = x.shape
    if x.stride(1) != 1 or x.dtype != torch.float32:
        raise ValueError("rows must be float32 with unit inner stride")
    words = torch.empty((rows, (n + 63) // 64), dtype=torch.int64, device=x.device)
    _lib.check(_lib.load().btx_pack_bits(x.data_ptr(), rows, n, x.stride(0), words.data_ptr(),
                                         _lib.stream_ptr()), "btx_pack_bits")
    return words


def unpack_rows(words: torch.Tensor, n: int, out: torch.Tensor | None = None) -> torch.Tensor:
    """(rows, ceil(n/64)) int64 words -> (rows, n) float32 0/1 (into ``out`` rows if given)."""
    _lib.require_hip(words)
    rows = words.shape[0]
    if words.shape[1] != (n + 63) // 64 or not words.is_contiguous():
        raise ValueError("words must be contiguous (rows, ceil(n/64))")
    if out is None:
        out = torch.empty((rows, n), dtype=torch.float32, device=words.device)
    if out.shape != (rows, n) or out.stride(1) != 1 or out.dtype != torch.float32:
        raise ValueError("out must be float32 (rows, n) with unit inner stride")
    _lib.check(_lib.load().btx_unpack_bits(words.data_ptr(), rows, n, out.data_ptr(),
                                           out.stride(0), _lib.stream_ptr()), "btx_unpack_bits")
    return out


def spmv_workspace(lanes: int, n_dst: int, slots: int, ordered: bool, device) -> torch.Tensor:
    nbytes = _lib.load().btx_spmv_workspace_bytes(lanes, n_dst, slots, int(ordered))
    return torch.zeros(nbytes, dtype=torch.uint8, device=device)


def spmv_event(layout: CsrLayout, w_csr: torch.Tensor, x2d: torch.Tensor, active: torch.Tensor,
               counts: torch.Tensor, *, ordered: bool, slots: int,
               workspace: torch.Tensor, w_dst: torch.Tensor) -> torch.Tensor:
    """out (lanes, n_dst) = x2d @ W from the compacted active lists; ``w_dst``
    holds the same weights in the coalesced (dst, src) order."""
    _lib.require_hip(x2d, w_csr, active, counts, workspace, w_dst)
    if layout.csc_ptr is None or w_dst.numel() != layout.nnz or not w_dst.is_contiguous():
        raise ValueError("destination-major arrays missing or mismatched")
    lanes = x2d.shape[0]
    if x2d.shape[1] != layout.n_src or w_csr.numel() != layout.nnz:
        raise ValueError("operand shapes do not match the connection layout")
    if not (_i32(active) and _i32(counts)) or active.shape[0] != lanes:
        raise ValueError("active/counts must be int32 (lanes, cap)/(lanes,)")
    if w_csr.dtype != torch.float32 or x2d.dtype != torch.float32:
        raise TypeError("the event-driven product supports float32 only")
    out = torch.empty((lanes, layout.n_dst), dtype=torch.float32, device=x2d.device)
    rc = _lib.load().btx_spmv_event(
        layout.row_ptr.data_ptr(), layout.col.data_ptr(), w_csr.data_ptr(), layout.n_src,
        layout.n_dst, layout.nnz, layout.csc_ptr.data_ptr(), layout.csc_src.data_ptr(),
        w_dst.data_ptr(), active.data_ptr(), counts.data_ptr(), lanes, active.shape[1],
        x2d.data_ptr(), x2d.shape[1], out.data_ptr(), int(ordered), slots,
        workspace.data_ptr(), workspace.numel(), _lib.stream_ptr())
    _lib.check(rc, "btx_spmv_event")
    return out


def spmv_transpose(layout: CsrLayout, w_csr: torch.Tensor, g: torch.Tensor) -> torch.Tensor:
    """gx (lanes, n_src) = g @ W^T."""
    _lib.require_hip(g, w_csr)
    g = g.contiguous()
    if g.shape[1] != layout.n_dst:
        raise ValueError("gradient width does not match n_dst")
    lanes = g.shape[0]
    gx = torch.empty((lanes, layout.n_src), dtype=torch.float32, device=g.device)
    _lib.check(_lib.load().btx_spmv_transpose(layout.row_ptr.data_ptr(), layout.col.data_ptr(),
                                              w_csr.data_ptr(), layout.n_src, layout.n_dst,
                                              g.data_ptr(), lanes, gx.data_ptr(),
                                              _lib.stream_ptr()), "btx_spmv_transpose")
    return gx


def spmv_weight_grad(layout: CsrLayout, x2d: torch.Tensor | None, active: torch.Tensor,
                     counts: torch.Tensor, g: torch.Tensor, gw_csr: torch.Tensor) -> None:
    """gw_csr += event-driven sum over active (lane, source) of x * g[:, col]."""
    _lib.require_hip(g, gw_csr, active, counts)
    g = g.contiguous()
    lanes = g.shape[0]
    _lib.check(_lib.load().btx_spmv_weight_grad(
        layout.row_ptr.data_ptr(), layout.col.data_ptr(), layout.n_src, layout.n_dst,
        active.data_ptr(), counts.data_ptr(), lanes, active.shape[1],
        None if x2d is None else x2d.data_ptr(), 0 if x2d is None else x2d.shape[1],
        g.data_ptr(), gw_csr.data_ptr(), _lib.stream_ptr()), "btx_spmv_weight_grad")


def poisson_fill(seed: torch.Tensor, prob: torch.Tensor, t0: torch.Tensor, kick: float, T: int,
                 n: int, out: torch.Tensor | None = None) -> torch.Tensor:
    """(T, lanes, n) float32 Poisson input: ``kick`` where the fused step's
    in-kernel drive (per-lane seed, rate and clock ``t0``) would fire, else
    0 (``btx_poisson_fill``).  A run given this tensor as x produces the same
    bits as a run given the drive."""
    _lib.require_hip(seed, prob, t0)
    lanes = seed.shape[0]
    for t_, dt_ in ((seed, torch.int64), (prob, torch.float32), (t0, torch.int64)):
        if t_.shape != (lanes,) or t_.dtype != dt_ or not t_.is_contiguous():
            raise ValueError("seed/prob/t0 must be contiguous (lanes,) int64/float32/int64")
    if out is None:
        out = torch.empty((T, lanes, n), dtype=torch.float32, device=seed.device)
    if out.shape != (T, lanes, n) or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError("out must be contiguous float32 (T, lanes, n)")
    # one launch covers at most 65535 steps
    for s0 in range(0, T, 65535):
        s1 = min(T, s0 + 65535)
        t0s = t0 + s0 if s0 else t0
        _lib.check(_lib.load().btx_poisson_fill(seed.data_ptr(), prob.data_ptr(), t0s.data_ptr(),
                                                float(kick), s1 - s0, lanes, n,
                                                out[s0:s1].data_ptr(), _lib.stream_ptr()),
                   "btx_poisson_fill")
    return out
