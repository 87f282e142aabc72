"""CPU oracle for the recurrent SNN time step -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package, and only as the checker.  The product
(``btorch_amd``) never imports, links or runs it.

Contents
  rsnn.py        op-for-op torch-CPU restatement of btorch's native path
                 (float32 parity oracle, float64 error oracle)
  csrc/seq_spmm.c  independent C restatement of the per-destination
                 sequential fmaf sum that torch.sparse.mm performs on CPU
  Makefile       builds liboracle.so from csrc/

Pinning (see tests/test_oracle.py): torch.sparse.mm vs the C sequential sum
(bit-exact), dense float64 matmul, the reference's delay-expansion known
answers, closed-form exponential-PSC decay, the LIF Euler closed form and
refractory period, the reference's delayed-PSC exact values, surrogate peak
and half-width invariants, and finite differences for the gradients.
"""
