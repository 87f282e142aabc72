"""Writes ``sparse_mm_order.npz``, the fixture that pins the summation order
of the reference's native product (btorch/models/linear.py:474-483, which
calls torch.sparse.mm on the CPU): a seeded coalesced (dst, src) matrix whose
weights span six decades (so the order of the fp32 additions shows in the
result), 0/1 spike rows, and torch-CPU's output.  Run from the repository
root: ``python tests/golden/make_sparse_mm_order.py``."""

from pathlib import Path

import numpy as np
import torch

OUT = Path(__file__).with_name("sparse_mm_order.npz")


def main() -> None:
    rng = np.random.default_rng(2024)
    n_src, n_dst, lanes = 300, 40, 6
    mask = rng.random((n_dst, n_src)) < 0.3
    dst, src = np.nonzero(mask)                      # row-major: coalesced (dst, src) order
    w = (rng.normal(size=dst.size) * 10.0 ** rng.uniform(-3, 3, dst.size)).astype(np.float32)
    x = (rng.random((lanes, n_src)) < 0.5).astype(np.float32)
    a = torch.sparse_coo_tensor(torch.from_numpy(np.stack([dst, src])), torch.from_numpy(w),
                                (n_dst, n_src)).coalesce()
    out = torch.sparse.mm(a, torch.from_numpy(x).T).T.contiguous().numpy()
    np.savez(OUT, dst=dst.astype(np.int64), src=src.astype(np.int64), w=w, x=x, out=out,
             n_src=np.int64(n_src), n_dst=np.int64(n_dst),
             torch_version=np.array(torch.__version__))
    print(f"wrote {OUT} ({dst.size} edges, torch {torch.__version__})")


if __name__ == "__main__":
    main()
