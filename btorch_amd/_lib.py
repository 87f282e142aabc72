"""ctypes binding of libbtx.so (declared in include/btorch_amd.h).

The library is the only compute backend of this package: nothing here falls
back to another implementation.  Loading fails loudly when the shared object
is missing, and every call that needs a GPU raises when tensors are not on a
HIP device.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import torch

from ._build import LIB_PATH

MAX_BINS = 64
RING_CHUNK = 1024

NEURON_LIF = 0
NEURON_GLIF3 = 1
NEURON_TWO_COMPARTMENT = 2
NEURON_MIXED = 3
MIXED_HARD_RESET = 4

SYN_EXP = 0
SYN_ALPHA = 1
SYN_ALPHA_BILLEH = 2
SYN_DUAL_EXP = 3

# indices into RsnnPlan.tc (BTX_TC_*)
TC_DA, TC_DA1, TC_DS, TC_DS1, TC_DTH1, TC_R_S, TC_E_L, TC_DT, TC_DTC = range(9)
TC_W_CA, TC_THETA_CA, TC_W_SA, TC_W_AS, TC_DELTA_TH = range(9, 14)
TC_NPARAM = 14

SURR_SIGMOID = 0
SURR_SUPERSPIKE = 1
SURR_ERF = 2

_c_i64 = ctypes.c_int64
_c_int = ctypes.c_int
_c_f = ctypes.c_float
_vp = ctypes.c_void_p


class RsnnPlan(ctypes.Structure):
    """Mirror of ``btx_rsnn_plan``; field order and types match the header."""

    _fields_ = [
        ("n", _c_i64),
        ("lanes", _c_i64),
        ("neuron_model", _c_int),
        ("hard_reset", _c_int),
        ("use_refractory", _c_int),
        ("ordered", _c_int),
        ("slots", _c_int),
        ("dt", _c_f),
        ("tau", _vp), ("tau_stride", _c_i64),
        ("v_th", _vp), ("v_th_stride", _c_i64),
        ("v_reset", _vp), ("v_reset_stride", _c_i64),
        ("c_m", _vp), ("c_m_stride", _c_i64),
        ("tau_ref", _vp), ("tau_ref_stride", _c_i64),
        ("tau_syn", _vp), ("tau_syn_stride", _c_i64),
        ("k_syn", _vp), ("k_syn_stride", _c_i64),
        ("v", _vp),
        ("refractory", _vp),
        ("psc", _vp),
        ("x", _vp), ("x_step_stride", _c_i64),
        ("in_seed", _vp),
        ("in_prob", _vp),
        ("in_t0", _vp),
        ("in_kick", _c_f),
        ("row_ptr", _vp),
        ("col", _vp),
        ("w", _vp),
        ("nnz", _c_i64),
        ("csr_d", _c_int),
        ("ring_d", _c_int),
        ("ring_idx", _vp),
        ("ring_cnt", _vp),
        ("ring_cursor", _c_int),
        ("nbins", _c_int),
        ("bin_delay", _c_int * MAX_BINS),
        ("bin_sub", _c_int * MAX_BINS),
        ("csc_ptr", _vp),
        ("csc_src", _vp),
        ("csc_w", _vp),
        ("ring_words", _vp),
        ("teacher", _vp),
        ("raster", _vp),
        ("rec_z", _vp),
        ("rec_v", _vp),
        ("rec_refractory", _vp),
        ("rec_psc", _vp),
        ("rec_v_pre", _vp),
        ("rec_v_prev", _vp),
        ("counts", _vp),
        ("rec_sample", _vp),
        ("n_sample", _c_i64),
        ("lane_t_end", _vp),
        ("n_asc", _c_int),
        ("v_rest", _vp), ("v_rest_stride", _c_i64),
        ("k_v", _vp), ("k_v_stride", _c_i64),
        ("asc_k", _vp),
        ("asc_kk", _vp),
        ("asc_amp", _vp),
        ("iasc", _vp),
        ("ring_rebuild", _c_int),
        ("bin_cap", _c_int),
        ("x_apical", _vp), ("x_apical_step_stride", _c_i64),
        ("i_a", _vp),
        ("i_bap", _vp),
        ("theta_th", _vp),
        ("tc", _vp * TC_NPARAM),
        ("tc_stride", _c_i64 * TC_NPARAM),
        ("rec_i_a", _vp),
        ("rec_i_bap", _vp),
        ("rec_theta_th", _vp),
        ("syn_kind", _c_int),
        ("syn_h1", _vp),
        ("syn_h2", _vp),
        ("syn_gain", _vp), ("syn_gain_stride", _c_i64),
        ("tau_syn2", _vp), ("tau_syn2_stride", _c_i64),
        ("k_syn2", _vp), ("k_syn2_stride", _c_i64),
        ("syn_a", _vp), ("syn_a_stride", _c_i64),
        ("rec_syn_h1", _vp),
        ("rec_syn_h2", _vp),
        ("workspace", _vp),
        ("workspace_bytes", ctypes.c_size_t),
        ("tables_ready", _c_int),
        ("neuron_kinds", _vp),
        ("mixed_models", _c_int),
        ("ring_lane_bits", _vp),
    ]


class AdjLifArgs(ctypes.Structure):
    """Mirror of ``btx_adj_lif_args``."""

    _fields_ = [
        ("n", _c_i64),
        ("lanes", _c_i64),
        ("dt", _c_f),
        ("hard_reset", _c_int),
        ("detach_reset", _c_int),
        ("surrogate", _c_int),
        ("alpha", _c_f),
        ("damping", _c_f),
        ("tau", _vp), ("tau_stride", _c_i64),
        ("v_th", _vp), ("v_th_stride", _c_i64),
        ("v_reset", _vp), ("v_reset_stride", _c_i64),
        ("c_m", _vp), ("c_m_stride", _c_i64),
        ("tau_syn", _vp), ("tau_syn_stride", _c_i64),
        ("k_syn", _vp), ("k_syn_stride", _c_i64),
        ("v_prev", _vp),
        ("v_pre", _vp),
        ("z", _vp),
        ("ref_prev", _vp),
        ("gz", _vp),
        ("gsyn", _vp),
        ("gv", _vp),
        ("gpsc", _vp),
        ("gx", _vp),
        ("gtau", _vp),
    ]


_SIGNATURES = {
    "btx_version": (_c_int, []),
    "btx_last_error": (ctypes.c_char_p, []),
    "btx_compact_tiles": (_c_i64, [_c_i64]),
    "btx_compact_dense": (_c_int, [_vp, _c_i64, _c_i64, _c_i64, _vp, _c_i64, _vp, _vp, _vp]),
    "btx_compact_bits": (_c_int, [_vp, _c_i64, _c_i64, _c_i64, _vp, _c_i64, _vp, _vp]),
    "btx_spmv_workspace_bytes": (ctypes.c_size_t, [_c_i64, _c_i64, _c_int, _c_int]),
    "btx_word_popcount": (_c_int, [_vp, _c_i64, _c_i64, _vp, _vp]),
    "btx_raster_coo": (_c_int, [_vp, _c_i64, _c_i64, _c_i64, _vp, _vp, _vp, _vp, _vp]),
    "btx_conn_build_workspace_bytes": (ctypes.c_size_t, [_c_i64]),
    "btx_conn_build": (_c_int, [_vp, _vp, _vp, _vp, _c_i64, _c_i64, _c_i64, _c_int, _vp, _vp,
                                _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp]),
    "btx_pack_bits": (_c_int, [_vp, _c_i64, _c_i64, _c_i64, _vp, _vp]),
    "btx_unpack_bits": (_c_int, [_vp, _c_i64, _c_i64, _vp, _c_i64, _vp]),
    "btx_spmv_event": (
        _c_int,
        [_vp, _vp, _vp, _c_i64, _c_i64, _c_i64, _vp, _vp, _vp, _vp, _vp, _c_i64, _c_i64,
         _vp, _c_i64, _vp, _c_int, _c_int, _vp, ctypes.c_size_t, _vp],
    ),
    "btx_spmv_event_noalloc_max_bytes": (ctypes.c_size_t, []),
    "btx_spmv_event_noalloc": (
        _c_int,
        [_vp, _vp, _vp, _c_i64, _c_i64, _c_i64, _vp, _vp, _vp, _vp, _vp, _c_i64, _c_i64,
         _vp, _c_i64, _vp, _c_int, _c_int, _vp],
    ),
    "btx_spmv_transpose": (_c_int, [_vp, _vp, _vp, _c_i64, _c_i64, _vp, _c_i64, _vp, _vp]),
    "btx_spmv_weight_grad": (
        _c_int,
        [_vp, _vp, _c_i64, _c_i64, _vp, _vp, _c_i64, _c_i64, _vp, _c_i64, _vp, _vp, _vp],
    ),
    "btx_rsnn_workspace_bytes": (ctypes.c_size_t, [_c_i64, _c_i64, _c_int, _c_int]),
    "btx_poisson_fill": (_c_int, [_vp, _vp, _vp, _c_f, _c_i64, _c_i64, _c_i64, _vp, _vp]),
    "btx_rsnn_run": (_c_int, [ctypes.POINTER(RsnnPlan), _c_i64, _vp]),
    "btx_rsnn_run_timed": (_c_int, [ctypes.POINTER(RsnnPlan), _c_i64, _vp,
                                    ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_int64),
                                    ctypes.POINTER(ctypes.c_double), _c_i64]),
    "btx_lif_step": (_c_int, [ctypes.POINTER(RsnnPlan), _vp, _vp, _vp, _vp]),
    "btx_exp_decay": (_c_int, [_vp, _c_i64, _c_i64, _vp, _c_i64, _vp, _c_i64, _vp]),
    "btx_syn_decay": (_c_int, [ctypes.POINTER(RsnnPlan), _vp]),
    "btx_adj_lif_step": (_c_int, [ctypes.POINTER(AdjLifArgs), _vp]),
    "btx_adj_lif_chunk": (_c_int, [ctypes.POINTER(AdjLifArgs), _c_i64, _vp, _vp, _vp, _vp, _vp,
                                   _vp]),
    "btx_adj_transpose": (_c_int, [_vp, _vp, _vp, _c_i64, _c_i64, _vp, _vp, _vp]),
    "btx_adj_wgrad": (_c_int, [_vp, _vp, _c_i64, _c_i64, _vp, _vp, _vp, _vp]),
    "btx_adj_transpose_delayed": (_c_int, [_vp, _vp, _vp, _c_i64, _c_int, _c_i64, _vp, _c_i64,
                                           _c_i64, _vp, _vp, _c_int, _vp]),
    "btx_adj_wgrad_delayed": (_c_int, [_vp, _vp, _c_i64, _c_int, _c_i64, _vp, _vp, _c_i64,
                                       _c_i64, _vp, _vp, _vp]),
}

_lock = threading.Lock()
_lib: ctypes.CDLL | None = None


def library_path() -> Path:
    return LIB_PATH


def load() -> ctypes.CDLL:
    """Load libbtx.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"btorch_amd: {LIB_PATH} is missing. Build it with "
                    "`python -m btorch_amd._build` (hipcc, gfx950)."
                )
            lib = ctypes.CDLL(str(LIB_PATH))
            for name, (res, args) in _SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def declared_symbols() -> list[str]:
    return list(_SIGNATURES)


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().btx_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed (status {rc}): {msg}")


def require_hip(*tensors: torch.Tensor) -> None:
    """Raise unless every tensor lives on a HIP device."""
    seen = False
    for t in tensors:
        if t is not None:
            if t.device.type != "cuda":
                raise RuntimeError(
                    "btorch_amd runs its kernels on a HIP device only; got a tensor on "
                    f"{t.device}. Move the module and inputs to 'cuda'."
                )
            seen = True
    # a tensor on a HIP device already shows that one is visible
    if not seen and not torch.cuda.is_available():
        raise RuntimeError("btorch_amd: no HIP device is visible")


def stream_ptr(device: torch.device | None = None) -> int:
    """The current HIP stream of `device` (default: the current device)."""
    if device is None:
        # torch's raw-stream query, without building a Stream object
        return torch._C._cuda_getCurrentRawStream(torch._C._cuda_getDevice())
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()
