/* Sequential sparse product: the summation order torch.sparse.mm uses on CPU
 * for a coalesced (dst, src) COO matrix A (btorch/models/linear.py:474-483).
 * For every lane b and every destination, a fresh 0.0f accumulator takes
 * acc = fmaf(w, x[b, src], acc) over the destination's entries in ascending
 * source order.  Test infrastructure only (CPU oracle). */
#include <math.h>
#include <stdint.h>
#include <string.h>

void seq_spmm(const int64_t* dst, const int64_t* src, const float* w, int64_t nnz,
              const float* x, int64_t lanes, int64_t n_src, int64_t n_dst, float* out) {
    memset(out, 0, sizeof(float) * (size_t)(lanes * n_dst));
    for (int64_t b = 0; b < lanes; ++b) {
        const float* xb = x + b * n_src;
        float* ob = out + b * n_dst;
        for (int64_t e = 0; e < nnz; ++e) ob[dst[e]] = fmaf(w[e], xb[src[e]], ob[dst[e]]);
    }
}
