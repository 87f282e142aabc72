"""btx_spmv_event_noalloc, the product entry point that allocates nothing,
against btx_spmv_event (bit-exact in ordered mode), called through the C ABI
on two calls in a row, the second with a smaller input."""

import ctypes

import numpy as np
import pytest
import scipy.sparse
import torch

from btorch_amd import _kernels, _lib
from btorch_amd.models import SparseConn

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _case(n_src, n_dst, lanes, seed, density=0.05, p=0.2):
    rng = np.random.default_rng(seed)
    dense = (rng.random((n_src, n_dst)) < p) * rng.lognormal(-1, 0.5, (n_src, n_dst))
    dense *= rng.choice([-1.0, 1.0], size=dense.shape)
    conn = SparseConn(scipy.sparse.coo_array(dense.astype(np.float32)), enforce_dale=False).to(DEV)
    x = torch.from_numpy((rng.random((lanes, n_src)) < density).astype(np.float32)).to(DEV)
    return conn, x


def _both(conn, x, ordered, slots=8):
    lay = conn._layout
    w_csr = conn.packed_weight().detach().contiguous()
    w_dst = conn.effective_weight_f32()
    act, cnt = _kernels.compact_dense(x)
    lanes = x.shape[0]
    ref = _kernels.spmv_event(lay, w_csr, x, act, cnt, ordered=ordered, slots=slots,
                              workspace=_kernels.spmv_workspace(lanes, lay.n_dst, slots, ordered,
                                                                DEV),
                              w_dst=w_dst)
    out = torch.full((lanes, lay.n_dst), float("nan"), device=DEV)
    rc = _lib.load().btx_spmv_event_noalloc(
        lay.row_ptr.data_ptr(), lay.col.data_ptr(), w_csr.data_ptr(), lay.n_src, lay.n_dst,
        lay.nnz, lay.csc_ptr.data_ptr(), lay.csc_src.data_ptr(), w_dst.data_ptr(),
        act.data_ptr(), cnt.data_ptr(), lanes, act.shape[1], x.data_ptr(), x.shape[1],
        out.data_ptr(), int(ordered), slots, _lib.stream_ptr())
    _lib.check(rc, "btx_spmv_event_noalloc")
    torch.cuda.synchronize()
    return out, ref


@pytest.mark.parametrize("ordered", [True, False])
def test_noalloc_product_equals_workspace_product_two_calls(ordered):
    # a larger call, then a smaller one: the second must not see the first's
    # leftovers in the shared scratch array
    for n_src, n_dst, lanes, seed in ((3000, 2500, 6, 1), (700, 300, 2, 2)):
        conn, x = _case(n_src, n_dst, lanes, seed)
        out, ref = _both(conn, x, ordered)
        assert not torch.isnan(out).any()
        if ordered:
            assert torch.equal(out.view(torch.int32), ref.view(torch.int32))
        else:
            assert torch.allclose(out, ref, rtol=1e-5, atol=1e-5)


def test_noalloc_rejects_sizes_beyond_its_scratch():
    lib = _lib.load()
    cap = lib.btx_spmv_event_noalloc_max_bytes()
    assert cap == 32 << 20
    lanes, n_dst = 64, 1 << 16
    assert lib.btx_spmv_workspace_bytes(lanes, n_dst, 8, 1) > cap
    # rejected on the host before any launch: every pointer may be NULL
    rc = lib.btx_spmv_event_noalloc(None, None, None, 10, n_dst, 0, ctypes.c_void_p(1),
                                    ctypes.c_void_p(1), ctypes.c_void_p(1), None, None, lanes,
                                    0, ctypes.c_void_p(1), 10, None, 1, 8, None)
    assert rc != 0
    assert b"scratch" in lib.btx_last_error()
