"""Seeded synthetic networks and inputs for the benchmark configurations.

Host-side numpy generators (``numpy.random.default_rng(seed)``); shared by
the benchmark and the tests so that every implementation sees the same
connectome, delays and Poisson input.

  plain_lif_net      N neurons, Erdos-Renyi p, lognormal(-1, 0.5) weights,
                     excitatory first, inhibitory x4 negative, no autapses
  powerlaw_net       out-degree power law (exponent, k_min, k_max), targets
                     uniform without replacement, integer delays U{d_lo..d_hi}
  poisson_input      Bernoulli(rate*dt) spikes
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse
import torch


@dataclass
class Connectome:
    n: int
    pre: np.ndarray        # int64 (E,)
    post: np.ndarray       # int64 (E,)
    w: np.ndarray          # float32 (E,)
    delay: np.ndarray | None = None  # int64 (E,) steps
    n_exc: int = 0

    @property
    def nnz(self) -> int:
        return int(self.w.size)

    def matrix(self) -> scipy.sparse.coo_array:
        return scipy.sparse.coo_array((self.w, (self.pre, self.post)), shape=(self.n, self.n))

    def expanded(self, n_delay_bins: int) -> scipy.sparse.coo_array:
        """Delay-expanded (n*D, n) matrix, row pre*D + clip(delay, 0, D-1)."""
        d = np.clip(self.delay, 0, n_delay_bins - 1)
        return scipy.sparse.coo_array((self.w, (self.pre * n_delay_bins + d, self.post)),
                                      shape=(self.n * n_delay_bins, self.n))


def _signed_lognormal(rng, pre, n_exc, inh_scale=4.0, mu=-1.0, sigma=0.5):
    w = rng.lognormal(mu, sigma, size=pre.size).astype(np.float32)
    inh = pre >= n_exc
    w[inh] *= np.float32(-inh_scale)
    return w


def plain_lif_net(n=4096, p=0.02, exc_frac=0.8, seed=0) -> Connectome:
    """Erdos-Renyi recurrent net, no autapses, sorted by (pre, post)."""
    rng = np.random.default_rng(seed)
    n_exc = int(round(exc_frac * n))
    pres, posts = [], []
    chunk = max(1, 2**24 // max(n, 1))
    for lo in range(0, n, chunk):
        hi = min(n, lo + chunk)
        mask = rng.random((hi - lo, n)) < p
        r, c = np.nonzero(mask)
        r = r + lo
        keep = r != c
        pres.append(r[keep])
        posts.append(c[keep])
    pre = np.concatenate(pres).astype(np.int64)
    post = np.concatenate(posts).astype(np.int64)
    w = _signed_lognormal(rng, pre, n_exc)
    return Connectome(n, pre, post, w, None, n_exc)


def powerlaw_degrees(rng, n, exponent, k_min, k_max) -> np.ndarray:
    """Discrete power-law out-degrees P(k) ~ k^-exponent on [k_min, k_max]."""
    ks = np.arange(k_min, k_max + 1, dtype=np.float64)
    pk = ks ** (-exponent)
    pk /= pk.sum()
    return rng.choice(ks.astype(np.int64), size=n, p=pk)


def powerlaw_net(n=140_000, exponent=2.1, k_min=20, k_max=5000, exc_frac=0.7,
                 delay_lo=1, delay_hi=20, seed=1) -> Connectome:
    """Power-law out-degree net with uniform targets and integer delays."""
    rng = np.random.default_rng(seed)
    n_exc = int(round(exc_frac * n))
    k = powerlaw_degrees(rng, n, exponent, k_min, min(k_max, n - 1))
    pre = np.repeat(np.arange(n, dtype=np.int64), k)
    post = np.empty(pre.size, dtype=np.int64)
    off = 0
    for i in range(n):
        ki = int(k[i])
        # uniform targets without replacement, excluding i
        t = rng.choice(n - 1, size=ki, replace=False)
        t[t >= i] += 1
        t.sort()
        post[off:off + ki] = t
        off += ki
    w = _signed_lognormal(rng, pre, n_exc)
    delay = rng.integers(delay_lo, delay_hi + 1, size=pre.size, dtype=np.int64)
    return Connectome(n, pre, post, w, delay, n_exc)


def powerlaw_net_fast(n=140_000, exponent=2.1, k_min=20, k_max=5000, exc_frac=0.7,
                      delay_lo=1, delay_hi=20, seed=1, w_mu=-1.0, w_sigma=0.5,
                      inh_scale=4.0) -> Connectome:
    """Same distribution as ``powerlaw_net`` with vectorised target sampling
    (targets drawn with replacement then de-duplicated per row; the realised
    degree can fall short of the draw by the collision count, <1% here)."""
    rng = np.random.default_rng(seed)
    n_exc = int(round(exc_frac * n))
    k = powerlaw_degrees(rng, n, exponent, k_min, min(k_max, n - 1))
    pre = np.repeat(np.arange(n, dtype=np.int64), k)
    post = rng.integers(0, n - 1, size=pre.size, dtype=np.int64)
    post[post >= pre] += 1
    key = pre * n + post
    key = np.unique(key)
    pre, post = key // n, key % n
    w = _signed_lognormal(rng, pre, n_exc, inh_scale, w_mu, w_sigma)
    delay = rng.integers(delay_lo, delay_hi + 1, size=pre.size, dtype=np.int64)
    return Connectome(n, pre, post, w, delay, n_exc)


def input_weights(n_in, n, density=0.05, seed=0) -> scipy.sparse.coo_array:
    """Fixed random positive input projection (n_in, n), lognormal(-1, 0.5)."""
    rng = np.random.default_rng(seed + 1000)
    mask = rng.random((n_in, n)) < density
    r, c = np.nonzero(mask)
    w = rng.lognormal(-1.0, 0.5, size=r.size).astype(np.float32)
    return scipy.sparse.coo_array((w, (r, c)), shape=(n_in, n))


def poisson_input(T, lanes, n, rate_hz, dt_ms, seed=0) -> np.ndarray:
    """Bernoulli(rate*dt) spikes (T, lanes, n) float32."""
    rng = np.random.default_rng(seed + 2000)
    p = rate_hz * dt_ms * 1e-3
    return (rng.random((T, lanes, n)) < p).astype(np.float32)


def er_net_fast(n=4096, p=0.02, exc_frac=0.8, seed=0, w_mu=-1.0, w_sigma=0.5,
                inh_scale=4.0) -> Connectome:
    """Erdos-Renyi net with Binomial(n-1, p) out-degrees and uniform targets
    (drawn with replacement, then de-duplicated per row; no autapses)."""
    rng = np.random.default_rng(seed)
    n_exc = int(round(exc_frac * n))
    k = rng.binomial(n - 1, p, size=n)
    pre = np.repeat(np.arange(n, dtype=np.int64), k)
    post = rng.integers(0, n - 1, size=pre.size, dtype=np.int64)
    post[post >= pre] += 1
    key = np.unique(pre * n + post)
    pre, post = key // n, key % n
    w = _signed_lognormal(rng, pre, n_exc, inh_scale, w_mu, w_sigma)
    return Connectome(n, pre, post, w, None, n_exc)


def sheet_net_device(side: int = 1000, mean_degree: float = 1000.0, sigma: float = 0.1,
                     delay_lo: int = 1, delay_hi: int = 40, exc_frac: float = 0.8,
                     w_mu: float = 0.0, w_sigma: float = 0.5, inh_scale: float = 4.0,
                     seed: int = 4, device="cuda", chunk: int = 1 << 27):
    """Edges of a distance-dependent random net on a periodic unit square,
    generated on the device: side*side neurons on a grid (neuron i at
    ((i % side + 0.5)/side, (i // side + 0.5)/side)); out-degree Poisson with
    the given mean; each target at a Gaussian displacement (sigma in units of
    the side) from the source, wrapped, snapped to its grid cell; delays
    U{delay_lo..delay_hi}; lognormal weights, negative and x inh_scale for
    the inhibitory sources (every fifth neuron with exc_frac = 0.8).
    Returns device tensors (pre, post, w, delay) in source order."""
    n = side * side
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    deg = torch.poisson(torch.full((n,), float(mean_degree), device=device), generator=g)
    deg = deg.to(torch.int64)
    pre = torch.repeat_interleave(torch.arange(n, device=device), deg)
    E = pre.numel()
    post = torch.empty(E, dtype=torch.int64, device=device)
    w = torch.empty(E, dtype=torch.float32, device=device)
    period = max(1, round(1.0 / max(1e-9, 1.0 - exc_frac)))
    for s in range(0, E, chunk):
        p = pre[s:s + chunk]
        m = p.numel()
        x = (p % side).to(torch.float32) + 0.5
        y = torch.div(p, side, rounding_mode="floor").to(torch.float32) + 0.5
        tx = torch.floor(x + torch.randn(m, generator=g, device=device) * (sigma * side))
        ty = torch.floor(y + torch.randn(m, generator=g, device=device) * (sigma * side))
        tx = torch.remainder(tx.to(torch.int64), side)
        ty = torch.remainder(ty.to(torch.int64), side)
        post[s:s + chunk] = ty * side + tx
        mag = torch.exp(w_mu + w_sigma * torch.randn(m, generator=g, device=device))
        inh = (p % period) == 0
        w[s:s + chunk] = torch.where(inh, -inh_scale * mag, mag)
    delay = torch.randint(delay_lo, delay_hi + 1, (E,), generator=g, device=device)
    return pre, post, w, delay


def connectome_140k(n: int = 140_000, k_min: int = 100, k_max: int = 5000, delay_max: int = 20,
                    seed: int = 1, w_mu: float = -3.7, device="cuda", sparse_backend="native"):
    """The 140k-neuron connectome workload: power-law out-degree (exponent
    2.1) on [k_min, k_max] (E ~ 50M at k_min 100), 70/30 E/I,
    lognormal(w_mu, 0.5) weights with inhibitory x4, delays U{1..delay_max}
    (D = delay_max + 1 bins), built on the device as a delay-expanded
    SparseConn; per-neuron tau_m ~ U(10, 30) ms.  Returns (conn, tau, nnz)."""
    import torch

    from .models import SparseConn

    D = delay_max + 1
    c = powerlaw_net_fast(n=n, k_min=k_min, k_max=k_max, exc_frac=0.7, delay_lo=1,
                          delay_hi=delay_max, seed=seed, w_mu=w_mu)
    edges = [torch.from_numpy(a).to(device) for a in (c.pre, c.post, c.w, c.delay)]
    conn = SparseConn.from_edges(*edges[:3], edges[3], D, shape=(n, n),
                                 sparse_backend=sparse_backend)
    rng = np.random.default_rng(seed + 7)
    tau = torch.from_numpy(rng.uniform(10.0, 30.0, n).astype(np.float32))
    return conn, tau, c.nnz
