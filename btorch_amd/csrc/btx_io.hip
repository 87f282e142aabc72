// Spike rasters to the coordinate (COO) encoding of recorded spikes.
//
// The reference stores a recorded (T, B, N) spike array as the index arrays
// of its nonzeros in row-major order plus their values
// (btorch/io/serialization.py:80-147, to_sparse_repr over np.nonzero).  The
// engine records spikes as 64-bit words (bit j of word w of a row is neuron
// 64w + j), so the same index arrays come straight from the words: a
// popcount per word, an exclusive prefix sum of those counts (the caller's),
// then one thread per word writes the coordinates of its set bits in
// ascending bit order.  Words are row-major over (row, word), so the output
// order is np.nonzero's.

#include <hip/hip_runtime.h>

#include "btx_common.h"

namespace btx {

// Bits of word w of a row that lie below n.
__device__ __forceinline__ uint64_t row_bits(uint64_t word, int64_t w, int64_t n) {
    const int64_t room = n - w * 64;
    return room >= 64 ? word : (room <= 0 ? 0ull : word & ((1ull << room) - 1ull));
}

__global__ __launch_bounds__(256) void k_word_popcount(const uint64_t* __restrict__ words,
                                                       int64_t count, int64_t n, int64_t nwords,
                                                       int64_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    out[i] = __popcll(row_bits(words[i], i % nwords, n));
}

__global__ __launch_bounds__(256) void k_raster_coo(const uint64_t* __restrict__ words,
                                                    int64_t count, int64_t n, int64_t nwords,
                                                    const int64_t* __restrict__ offsets,
                                                    int64_t row_div, int32_t* __restrict__ idx_a,
                                                    int32_t* __restrict__ idx_b,
                                                    int32_t* __restrict__ idx_n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const int64_t row = i / nwords, w = i - row * nwords;
    uint64_t bits = row_bits(words[i], w, n);
    int64_t pos = offsets[i];
    const int32_t a = (int32_t)(row / row_div), b = (int32_t)(row - (row / row_div) * row_div);
    while (bits) {
        const int j = __ffsll((unsigned long long)bits) - 1;
        idx_a[pos] = a;
        idx_b[pos] = b;
        idx_n[pos] = (int32_t)(w * 64 + j);
        ++pos;
        bits &= bits - 1;
    }
}

}  // namespace btx

using namespace btx;

BTX_API int btx_word_popcount(const uint64_t* words, int64_t rows, int64_t n, int64_t* counts,
                              void* stream) {
    if (rows <= 0 || n <= 0) return 0;
    const int64_t nwords = (n + 63) / 64, count = rows * nwords;
    hipLaunchKernelGGL(k_word_popcount, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, words, count, n, nwords, counts);
    return check_launch("btx_word_popcount");
}

BTX_API int btx_raster_coo(const uint64_t* words, int64_t rows, int64_t n, int64_t row_div,
                           const int64_t* offsets, int32_t* idx_a, int32_t* idx_b,
                           int32_t* idx_n, void* stream) {
    if (rows <= 0 || n <= 0) return 0;
    if (row_div < 1) return fail("btx_raster_coo: row_div must be >= 1");
    if (n >= ((int64_t)1 << 31) || rows / row_div >= ((int64_t)1 << 31))
        return fail("btx_raster_coo: indices exceed int32");
    const int64_t nwords = (n + 63) / 64, count = rows * nwords;
    hipLaunchKernelGGL(k_raster_coo, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, words, count, n, nwords, offsets, row_div, idx_a,
                       idx_b, idx_n);
    return check_launch("btx_raster_coo");
}
