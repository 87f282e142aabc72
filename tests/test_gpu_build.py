"""GPU parity of the device-built connectivity (btx_conn_build) with the
reference's host path: expand_conn_for_delays (connection.py:973-1027), then
BaseSparseConn's transpose + scipy sum_duplicates (linear.py:276-291), then
the CSR layout; every array equal, weights bit for bit."""

import numpy as np
import pytest
import scipy.sparse
import torch

from btorch_amd import workloads
from btorch_amd.connectome import expand_conn_for_delays
from btorch_amd.models import (
    LIF, ExponentialPSC, HeterSynapsePSC, RecurrentNN, SparseConn, environ, functional,
)

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _host(pre, post, w, delay, D, n_pre, n_post):
    m = scipy.sparse.coo_array((w, (pre, post)), shape=(n_pre, n_post))
    if delay is not None:
        m = expand_conn_for_delays(m, delay, D)
    return SparseConn(m)


def _device(pre, post, w, delay, D, n_pre, n_post):
    t = (lambda a: None if a is None else torch.from_numpy(a).to(DEV))
    return SparseConn.from_edges(t(pre), t(post), t(w), t(delay), D, shape=(n_pre, n_post))


def _assert_same(a: SparseConn, b: SparseConn):
    assert (a.in_features, a.out_features) == (b.in_features, b.out_features)
    assert torch.equal(a.indices.cpu(), b.indices.cpu())
    assert torch.equal(a.magnitude.detach().cpu(), b.magnitude.detach().cpu())
    for name in ("_csr_row_ptr", "_csr_col", "_csr_perm", "_csc_ptr", "_csc_src"):
        x, y = getattr(a, name).cpu(), getattr(b, name).cpu()
        assert x.dtype == y.dtype and torch.equal(x, y), name


def _edges(n_pre, n_post, nnz, D, seed, dup=False):
    rng = np.random.default_rng(seed)
    pre = rng.integers(0, n_pre, nnz)
    post = rng.integers(0, n_post, nnz)
    delay = rng.integers(0, D + 2, nnz) if D > 1 else None   # some past the last bin
    if not dup:
        key = (pre * n_post + post) * (D + 2) + (delay if delay is not None else 0)
        _, first = np.unique(key, return_index=True)
        keep = np.sort(first)
        pre, post = pre[keep], post[keep]
        delay = delay[keep] if delay is not None else None
    w = rng.normal(size=pre.size).astype(np.float32)
    return pre.astype(np.int64), post.astype(np.int64), w, delay


@pytest.mark.parametrize("n_pre,n_post,nnz,D", [(1000, 1000, 50_000, 7), (300, 1, 2000, 3),
                                                (1, 500, 400, 1), (5000, 3000, 200_000, 21)])
@pytest.mark.parametrize("dup", [False, True])
def test_device_build_equals_host_path(n_pre, n_post, nnz, D, dup):
    pre, post, w, delay = _edges(n_pre, n_post, nnz, D, seed=nnz + D, dup=dup)
    print(f"\nbuild n_pre={n_pre} n_post={n_post} nnz={pre.size} D={D} dup={dup}", flush=True)
    host = _host(pre, post, w, delay, D, n_pre, n_post)
    dev = _device(pre, post, w, delay, D, n_pre, n_post)
    if dup:
        assert host.nnz < pre.size, "no duplicates were drawn"
    _assert_same(host, dev)


def test_device_build_edge_cases():
    # no edges; one edge; every edge in one row
    for pre, post, w, shape in [
            (np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0, np.float32), (4, 6)),
            (np.array([3]), np.array([2]), np.array([0.5], np.float32), (4, 6)),
            (np.full(100, 2), np.arange(100) % 7, np.linspace(-1, 1, 100).astype(np.float32),
             (4, 7))]:
        host = SparseConn(scipy.sparse.coo_array((w, (pre, post)), shape=shape))
        dev = SparseConn.from_edges(torch.from_numpy(pre).to(DEV), torch.from_numpy(post).to(DEV),
                                    torch.from_numpy(w).to(DEV), shape=shape)
        _assert_same(host, dev)
    with pytest.raises(ValueError):
        SparseConn.from_edges(torch.tensor([0, 5], device=DEV), torch.tensor([0, 1], device=DEV),
                              torch.ones(2, device=DEV), shape=(4, 4))
    with pytest.raises(ValueError):
        SparseConn.from_edges(torch.tensor([0, 1], device=DEV), torch.tensor([0, 1], device=DEV),
                              torch.ones(2, device=DEV), torch.tensor([1, -1], device=DEV), 3)


def test_engine_on_device_built_connectome_equals_host_built():
    n, B, T, D = 800, 2, 120, 9
    c = workloads.powerlaw_net_fast(n=n, k_min=5, k_max=80, delay_lo=1, delay_hi=D - 1, seed=3)
    x = torch.from_numpy(workloads.poisson_input(T, B, n, 400.0, 0.1, seed=3) * 25.0).to(DEV)
    rasters = []
    for conn in (SparseConn(c.expanded(D)),
                 SparseConn.from_edges(torch.from_numpy(c.pre).to(DEV),
                                       torch.from_numpy(c.post).to(DEV),
                                       torch.from_numpy(c.w).to(DEV),
                                       torch.from_numpy(c.delay).to(DEV), D, shape=(n, n))):
        syn = HeterSynapsePSC(n, 1, None, conn, base_psc=ExponentialPSC, max_delay_steps=D,
                              use_circular_buffer=True, tau_syn=5.0)
        model = RecurrentNN(LIF(n, tau_ref=2.0), syn, step_mode="m",
                            update_state_names=[]).to(DEV)
        functional.init_net_state(model, batch_size=B, device=DEV, dtype=torch.float32)
        with environ.context(dt=0.1), torch.no_grad():
            z, _ = model(x)
        rasters.append(z.cpu())
    assert rasters[0].sum() > 50
    assert torch.equal(rasters[0], rasters[1])


@pytest.mark.parametrize("n_pre,n_post,nnz", [(3, 2, 20_000), (10, 10, 2_000), (50, 40, 30_000)])
def test_device_build_heavy_duplicates_sum_like_scipy(n_pre, n_post, nnz):
    # segments of ~3300, ~20 and ~15 duplicates: the recursive, the
    # eight-accumulator and the sequential branches of the pairwise sum
    rng = np.random.default_rng(nnz)
    pre = rng.integers(0, n_pre, nnz).astype(np.int64)
    post = rng.integers(0, n_post, nnz).astype(np.int64)
    w = rng.normal(size=nnz).astype(np.float32)
    host = _host(pre, post, w, None, 1, n_pre, n_post)
    dev = _device(pre, post, w, None, 1, n_pre, n_post)
    assert host.nnz <= n_pre * n_post
    _assert_same(host, dev)
