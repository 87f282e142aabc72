"""Recorded spikes in the reference's coordinate encoding
(btorch/io/serialization.py:8-20, 80-196).

A (T, B, N) spike array is stored as the index arrays of its nonzeros in
row-major (np.nonzero) order -- ``{name}_idx_time``, ``{name}_idx_batch``,
``{name}_idx_neuron`` (int32) -- their values ``{name}_data``, and a marker
``{name}`` holding the count and the original shape, dims and dtype.  The
engine records spikes as 64-bit words, so the coordinates are produced on
the device from the words without a dense array (``btx_word_popcount`` and
``btx_raster_coo``).  The xarray/Zarr container around these arrays is not
rebuilt; the dictionary carries the same variables and attributes.
"""

from __future__ import annotations

import torch

from . import _lib

DIMS = ("time", "batch", "neuron")


def raster_to_sparse(words: torch.Tensor, n: int, var_name: str = "spikes") -> dict:
    """Coordinates of the spikes of a (T, B, ceil(n/64)) word raster (the
    engine's raster), as ``to_sparse_repr`` lays them out."""
    if words.dim() != 3 or words.shape[-1] != (n + 63) // 64:
        raise ValueError("words must be (T, B, ceil(n/64))")
    if words.dtype not in (torch.int64, torch.uint64):
        raise TypeError("words must be 64-bit integers")
    _lib.require_hip(words)
    words = words.contiguous()
    T, B, _ = words.shape
    rows = T * B
    dev = words.device
    lib = _lib.load()
    stream = _lib.stream_ptr()
    counts = torch.empty(words.numel(), dtype=torch.int64, device=dev)
    _lib.check(lib.btx_word_popcount(words.data_ptr(), rows, n, counts.data_ptr(), stream),
               "btx_word_popcount")
    ends = torch.cumsum(counts, 0)
    nnz = int(ends[-1]) if ends.numel() else 0
    offsets = (ends - counts).contiguous()
    idx = [torch.empty(nnz, dtype=torch.int32, device=dev) for _ in range(3)]
    if nnz:
        _lib.check(lib.btx_raster_coo(words.data_ptr(), rows, n, B, offsets.data_ptr(),
                                      idx[0].data_ptr(), idx[1].data_ptr(), idx[2].data_ptr(),
                                      stream), "btx_raster_coo")
    out = {f"{var_name}_idx_{d}": t for d, t in zip(DIMS, idx)}
    out[f"{var_name}_data"] = torch.ones(nnz, dtype=torch.float32, device=dev)
    out[var_name] = {"_btorch_sparse": True, "original_shape": [T, B, n],
                     "original_dims": list(DIMS), "original_dtype": "float32", "nnz": nnz}
    return out


def spikes_to_sparse(z: torch.Tensor, var_name: str = "spikes") -> dict:
    """The same for a dense (T, B, N) 0/1 spike tensor (a model's output)."""
    if z.dim() != 3:
        raise ValueError("spikes must be (T, B, N)")
    _lib.require_hip(z)
    T, B, n = z.shape
    z = z.to(torch.float32).contiguous()
    words = torch.empty((T, B, (n + 63) // 64), dtype=torch.int64, device=z.device)
    _lib.check(_lib.load().btx_pack_bits(z.data_ptr(), T * B, n, n, words.data_ptr(),
                                         _lib.stream_ptr()), "btx_pack_bits")
    return raster_to_sparse(words, n, var_name)


def sparse_to_spikes(d: dict, var_name: str = "spikes") -> torch.Tensor:
    """Dense (T, B, N) spikes back from the coordinate encoding
    (serialization.py:151-196)."""
    meta = d[var_name]
    shape = tuple(meta["original_shape"])
    ref = d[f"{var_name}_data"]
    out = torch.zeros(shape, dtype=ref.dtype, device=ref.device)
    idx = tuple(d[f"{var_name}_idx_{dim}"].long() for dim in meta["original_dims"])
    out[idx] = ref
    return out


__all__ = ["raster_to_sparse", "spikes_to_sparse", "sparse_to_spikes"]
