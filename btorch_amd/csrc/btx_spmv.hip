// Spike compaction and the event-driven sparse product on gfx950.
//
// Layout: connectivity is CSR over (virtual) source rows -- row_ptr int64,
// col int32 (destination), w fp32 -- with each row's destinations ascending.
// A row of the delay-expanded matrix is pre*csr_d + delay_bin
// (btorch/connectome/connection.py:1020-1027).
//
// The product is push-style: each active source row is streamed once,
// coalesced, by a group of G lanes, and every edge delivers its term to its
// destination.  Two accumulation modes:
//  * ordered: the term (source id, weight) goes into one of K per-destination
//    slots claimed with an int atomic (overflow list past K); a later pass
//    sums each destination's terms in ascending source order, bit-identical
//    to the CPU reference.
//  * fast: float atomics straight into the accumulator (order-free, like the
//    reference's Triton kernels).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "btx_common.h"
#include "btx_spmv.h"

namespace btx {

thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(const std::string& msg) {
    set_error(msg);
    return 1;
}

int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return (int)e;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Compaction: one 1024-thread workgroup per lane.  Thread t owns a contiguous
// range of the lane's words (or elements), counts its active entries, and a
// block scan gives it the position of its first index, so indices come out
// ascending.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(1024) void k_compact_bits(const uint64_t* __restrict__ words,
                                                       int64_t n, int64_t nwords, int64_t wstride,
                                                       int32_t* __restrict__ idx,
                                                       int64_t istride,
                                                       int32_t* __restrict__ counts) {
    const int lane = blockIdx.x;
    const uint64_t* wl = words + (int64_t)lane * wstride;
    const int64_t per = (nwords + blockDim.x - 1) / blockDim.x;
    const int64_t w0 = (int64_t)threadIdx.x * per;
    const int64_t w1 = w0 + per < nwords ? w0 + per : nwords;
    // bits past n in the last word never count
    const uint64_t tail = (n & 63) ? ((1ull << (n & 63)) - 1) : ~0ull;
    int c = 0;
    for (int64_t w = w0; w < w1; ++w) c += __popcll(w == nwords - 1 ? (wl[w] & tail) : wl[w]);
    int total;
    const int excl = block_excl_scan<16>(c, total);
    int32_t* out = idx + (int64_t)lane * istride + excl;
    for (int64_t w = w0; w < w1; ++w) {
        uint64_t m = w == nwords - 1 ? (wl[w] & tail) : wl[w];
        while (m) {
            const int b = __ffsll((unsigned long long)m) - 1;
            *out++ = (int32_t)(w * 64 + b);
            m &= m - 1;
        }
    }
    if (threadIdx.x == 0) counts[lane] = total;
}

// Compaction of a dense fp32 matrix, in tiles of kCTile consecutive columns
// of one lane: each wavefront owns kCTile / 4 columns and walks them 64 at a
// time (one coalesced 256-byte load per step), turning the nonzero test into
// a 64-bit ballot.  Pass 1 counts each tile; pass 2 adds the counts of the
// lane's earlier tiles, the earlier wavefronts of its own tile and, within a
// wavefront, the set lanes below (mbcnt), so the indices come out ascending.
constexpr int kCTile = 4096;
constexpr int kCWave = kCTile / 4;

__device__ __forceinline__ int wave_nonzeros(const float* __restrict__ xl, int64_t base,
                                             int64_t n) {
    const int ln = threadIdx.x & 63;
    int c = 0;
#pragma unroll 4
    for (int j = 0; j < kCWave / 64; ++j) {
        const int64_t e = base + j * 64 + ln;
        c += __popcll(__ballot(e < n && xl[e] != 0.f));
    }
    return c;
}

__global__ __launch_bounds__(256) void k_compact_count(const float* __restrict__ x, int64_t n,
                                                       int64_t xstride, int64_t ntiles,
                                                       int32_t* __restrict__ tile_counts) {
    __shared__ int s_c[4];
    const int64_t lane = blockIdx.y, tile = blockIdx.x;
    const int w = threadIdx.x >> 6;
    const int c = wave_nonzeros(x + lane * xstride, tile * kCTile + w * kCWave, n);
    if ((threadIdx.x & 63) == 0) s_c[w] = c;
    __syncthreads();
    if (threadIdx.x == 0) tile_counts[lane * ntiles + tile] = s_c[0] + s_c[1] + s_c[2] + s_c[3];
}

__global__ __launch_bounds__(256) void k_compact_write(const float* __restrict__ x, int64_t n,
                                                       int64_t xstride, int64_t ntiles,
                                                       const int32_t* __restrict__ tile_counts,
                                                       int32_t* __restrict__ idx,
                                                       int64_t istride,
                                                       int32_t* __restrict__ counts) {
    __shared__ int s_c[4];
    __shared__ int s_pre[4];
    const int64_t lane = blockIdx.y, tile = blockIdx.x;
    const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
    const float* xl = x + lane * xstride;
    // nonzeros of the lane's earlier tiles
    int pre = 0;
    for (int64_t t = threadIdx.x; t < tile; t += 256) pre += tile_counts[lane * ntiles + t];
    for (int o = 32; o > 0; o >>= 1) pre += __shfl_down(pre, o, 64);
    const int64_t base = tile * kCTile + w * kCWave;
    const int c = wave_nonzeros(xl, base, n);
    if (ln == 0) {
        s_pre[w] = pre;
        s_c[w] = c;
    }
    __syncthreads();
    int off = s_pre[0] + s_pre[1] + s_pre[2] + s_pre[3];
    for (int k = 0; k < w; ++k) off += s_c[k];
    int32_t* out = idx + lane * istride;
    const uint64_t below = (ln == 0) ? 0ull : (~0ull >> (64 - ln));
    for (int j = 0; j < kCWave / 64; ++j) {
        const int64_t e = base + j * 64 + ln;
        const bool nz = e < n && xl[e] != 0.f;
        const uint64_t b = __ballot(nz);
        if (nz) out[off + __popcll(b & below)] = (int32_t)e;
        off += __popcll(b);
    }
    if (tile == ntiles - 1 && threadIdx.x == 255) counts[lane] = off;
}

// ---------------------------------------------------------------------------
// Push kernel.  Work items are (lane, bin, k): the k-th active neuron of the
// ring slot that bin reads for that lane.  Every workgroup scans the
// (lane, bin) segment sizes into LDS and maps each item to its segment by a
// binary search; each group of G lanes streams one active row at a time.
// ---------------------------------------------------------------------------

constexpr int kPushThreads = 256;
constexpr int kMaxSeg = 12288;

template <int G, int ORDERED, bool HASX>
__global__ __launch_bounds__(kPushThreads) void k_push(PushArgs a) {
    extern __shared__ int s_off[];  // nseg + 1 segment offsets
    __shared__ int s_slot[BTX_MAX_BINS];
    __shared__ int s_sub[BTX_MAX_BINS];
    const int nseg = a.lanes * a.nbins;
    if (threadIdx.x < a.nbins) {
        s_slot[threadIdx.x] = a.bin_slot[threadIdx.x];
        s_sub[threadIdx.x] = a.bin_sub[threadIdx.x];
    }
    __syncthreads();
    // Segment sizes, chunked per thread, then one block scan.
    const int per = (nseg + kPushThreads - 1) / kPushThreads;
    const int s0 = threadIdx.x * per;
    const int s1 = s0 + per < nseg ? s0 + per : nseg;
    int c = 0;
    for (int s = s0; s < s1; ++s) {
        const int b = s / a.nbins, j = s - b * a.nbins;
        c += a.ring_cnt[(int64_t)s_slot[j] * a.ring_lanes + a.lane_base + b];
    }
    int total;
    int run = block_excl_scan<kPushThreads / 64>(c, total);
    for (int s = s0; s < s1; ++s) {
        const int b = s / a.nbins, j = s - b * a.nbins;
        s_off[s] = run;
        run += a.ring_cnt[(int64_t)s_slot[j] * a.ring_lanes + a.lane_base + b];
    }
    __syncthreads();
    if (threadIdx.x == 0) s_off[nseg] = total;
    __syncthreads();

    constexpr int gpb = kPushThreads / G;
    const int gl = threadIdx.x % G;
    const int ngroups = gridDim.x * gpb;
    for (int item = blockIdx.x * gpb + threadIdx.x / G; item < total; item += ngroups) {
        int lo = 0, hi = nseg;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (s_off[mid] <= item) lo = mid;
            else hi = mid;
        }
        const int64_t b = a.lane_base + lo / a.nbins;
        const int j = lo % a.nbins;
        const int k = item - s_off[lo];
        const int src = a.ring_idx[((int64_t)s_slot[j] * a.ring_lanes + b) * a.ring_cap + k];
        const int64_t r = (int64_t)src * a.csr_d + s_sub[j];
        const int64_t beg = a.row_ptr[r], end = a.row_ptr[r + 1];
        const float xv = HASX ? a.x[b * a.x_stride + src] : 1.f;
        const int64_t base = b * a.n_dst;
        for (int64_t e = beg + gl; e < end; e += G) {
            const int dst = a.col[e];
            const float wv = a.w[e];
            const int64_t key = base + dst;
            if (ORDERED) {
                const int pos = atomicAdd(&a.cnt[key], 1);
                if (pos < a.K)
                    a.slots[(int64_t)pos * a.ring_lanes * a.n_dst + key] =
                        make_int2((int)r, __float_as_int(wv));
            } else {
                atomicAdd(&a.acc[key], HASX ? wv * xv : wv);
            }
        }
    }
}

template <int G>
static void launch_push_g(const PushArgs& a, int blocks, hipStream_t s) {
    const bool hasx = a.x != nullptr;
    const size_t lds = sizeof(int) * ((size_t)a.lanes * a.nbins + 1);
    if (a.ordered) {
        if (hasx) hipLaunchKernelGGL((k_push<G, 1, true>), dim3(blocks), dim3(kPushThreads), lds, s, a);
        else hipLaunchKernelGGL((k_push<G, 1, false>), dim3(blocks), dim3(kPushThreads), lds, s, a);
    } else {
        if (hasx) hipLaunchKernelGGL((k_push<G, 0, true>), dim3(blocks), dim3(kPushThreads), lds, s, a);
        else hipLaunchKernelGGL((k_push<G, 0, false>), dim3(blocks), dim3(kPushThreads), lds, s, a);
    }
}

// Launches the push over all ring lanes, in chunks small enough for the
// per-workgroup segment table in LDS.
int launch_push(const PushArgs& all, int group, int blocks, hipStream_t s) {
    if (all.nbins < 1 || all.nbins > BTX_MAX_BINS) return fail("event product: bad bin count");
    const int64_t chunk = kMaxSeg / all.nbins;
    for (int64_t l0 = 0; l0 < all.ring_lanes; l0 += chunk) {
        PushArgs a = all;
        a.lane_base = l0;
        a.lanes = (int)(all.ring_lanes - l0 < chunk ? all.ring_lanes - l0 : chunk);
        switch (group) {
            case 8: launch_push_g<8>(a, blocks, s); break;
            case 16: launch_push_g<16>(a, blocks, s); break;
            case 32: launch_push_g<32>(a, blocks, s); break;
            default: launch_push_g<64>(a, blocks, s); break;
        }
        if (int rc = check_launch("k_push")) return rc;
    }
    return 0;
}

int pick_group(double avg_row) {
    if (avg_row >= 48.0) return 64;
    if (avg_row >= 24.0) return 32;
    if (avg_row >= 12.0) return 16;
    return 8;
}

int push_blocks() {
    static int cached = 0;
    if (cached == 0) {
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess)
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                hipSuccess)
                cus = 256;
        cached = 4 * (cus > 0 ? cus : 256);
    }
    return cached;
}

// ---------------------------------------------------------------------------
// Reduction of the standalone product: out = ordered sum (or the atomic
// accumulator), leaving the workspace zeroed.
// ---------------------------------------------------------------------------

// Ordered reduction of the standalone product.  Destinations with at most K
// terms sort their slots; a destination with more terms pulls every incoming
// edge in the coalesced (dst, src) order instead -- x = 0 terms leave the
// accumulator unchanged, so the pull equals the event sum bit for bit.
__global__ __launch_bounds__(256) void k_reduce_ordered(int64_t m, int64_t n_dst,
                                                        int* __restrict__ cnt,
                                                        const int2* __restrict__ slots, int K,
                                                        const int64_t* __restrict__ csc_ptr,
                                                        const int32_t* __restrict__ csc_src,
                                                        const float* __restrict__ csc_w,
                                                        const float* __restrict__ x,
                                                        int64_t x_stride,
                                                        float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int c = cnt[i];
    float acc = 0.f;
    if (c) {
        const int64_t b = i / n_dst;
        const float* xl = x + b * x_stride;
        if (c <= K) {
            acc = ordered_sum<true>(i, c, slots, m, xl);
        } else {
            const int64_t d = i - b * n_dst;
            for (int64_t e = csc_ptr[d]; e < csc_ptr[d + 1]; ++e)
                acc = fmaf(csc_w[e], xl[csc_src[e]], acc);
        }
        cnt[i] = 0;
    }
    out[i] = acc;
}

__global__ __launch_bounds__(256) void k_zero(float* __restrict__ p, int64_t m) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) p[i] = 0.f;
}

__global__ __launch_bounds__(256) void k_copy_zero(float* __restrict__ acc, float* __restrict__ out,
                                                   int64_t m) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) {
        out[i] = acc[i];
        acc[i] = 0.f;
    }
}

// ---------------------------------------------------------------------------
// Backward helpers.
// ---------------------------------------------------------------------------

// gx[b, s] = sum_e w[e] * g[b, col[e]] over row s.  One wavefront per row;
// lanes split the row's edges, a wave reduction finishes each lane's sum.
__global__ __launch_bounds__(256) void k_spmv_transpose(const int64_t* __restrict__ rp,
                                                        const int32_t* __restrict__ col,
                                                        const float* __restrict__ w,
                                                        int64_t n_src, int64_t n_dst,
                                                        const float* __restrict__ g,
                                                        int64_t lanes, float* __restrict__ gx) {
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int ln = threadIdx.x & 63;
    if (row >= n_src) return;
    const int64_t beg = rp[row], end = rp[row + 1];
    for (int64_t b = 0; b < lanes; ++b) {
        const float* gb = g + b * n_dst;
        float acc = 0.f;
        for (int64_t e = beg + ln; e < end; e += 64) acc += w[e] * gb[col[e]];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (ln == 0) gx[b * n_src + row] = acc;
    }
}

// gw[e] += x[b, s] * g[b, col[e]] for every active (b, s).
__global__ __launch_bounds__(256) void k_weight_grad(const int64_t* __restrict__ rp,
                                                     const int32_t* __restrict__ col,
                                                     int64_t n_dst,
                                                     const int32_t* __restrict__ active,
                                                     const int32_t* __restrict__ counts,
                                                     int64_t lanes, int64_t astride,
                                                     const float* __restrict__ x,
                                                     int64_t x_stride,
                                                     const float* __restrict__ g,
                                                     float* __restrict__ gw) {
    const int64_t b = blockIdx.y;
    const int cnt = counts[b];
    const int wid = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int ln = threadIdx.x & 63;
    const int nw = gridDim.x * 4;
    for (int k = wid; k < cnt; k += nw) {
        const int s = active[b * astride + k];
        const float xv = x ? x[b * x_stride + s] : 1.f;
        const int64_t beg = rp[s], end = rp[s + 1];
        for (int64_t e = beg + ln; e < end; e += 64)
            atomicAdd(&gw[e], xv * g[b * n_dst + col[e]]);
    }
}

// Spike words of rows of a dense fp32 matrix: bit j of word w of a row is
// (x[row, 64w + j] != 0).  One thread per element, one ballot per word.
__global__ __launch_bounds__(256) void k_pack_bits(const float* __restrict__ x, int64_t n,
                                                   int64_t x_stride, uint64_t* __restrict__ words,
                                                   int64_t nwords) {
    const int64_t row = blockIdx.y;
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int on = j < n ? (x[row * x_stride + j] != 0.f) : 0;
    const uint64_t m = __ballot(on);
    if ((threadIdx.x & 63) == 0) {
        const int64_t w = j >> 6;
        if (w < nwords) words[row * nwords + w] = m;
    }
}

// out[row, j] = bit j of the row's words, as 0.0f / 1.0f.
__global__ __launch_bounds__(256) void k_unpack_bits(const uint64_t* __restrict__ words,
                                                     int64_t n, int64_t nwords,
                                                     float* __restrict__ out,
                                                     int64_t out_stride) {
    const int64_t row = blockIdx.y;
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const uint64_t w = words[row * nwords + (j >> 6)];
    out[row * out_stride + j] = ((w >> (j & 63)) & 1ull) ? 1.f : 0.f;
}

// The same, 16 floats per thread as four 16-byte stores (out rows 16-byte
// aligned): grid (ceil(n / 4096), rows), each workgroup a 16 KB span of a
// row.  Four consecutive columns share one word, so a store takes a nibble.
__global__ __launch_bounds__(256) void k_unpack_bits16(const uint64_t* __restrict__ words,
                                                       int64_t n, int64_t nwords,
                                                       float* __restrict__ out,
                                                       int64_t out_stride) {
    const int64_t row = blockIdx.y;
    const uint64_t* wr = words + row * nwords;
    float* orow = out + row * out_stride;
    const int64_t base = (int64_t)blockIdx.x * 4096;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t j = base + (int64_t)(k * 256 + threadIdx.x) * 4;
        if (j >= n) break;
        const unsigned b = (unsigned)(wr[j >> 6] >> (j & 63)) & 0xFu;
        if (j + 4 <= n) {
            float4 v;
            v.x = (b & 1u) ? 1.f : 0.f;
            v.y = (b & 2u) ? 1.f : 0.f;
            v.z = (b & 4u) ? 1.f : 0.f;
            v.w = (b & 8u) ? 1.f : 0.f;
            *reinterpret_cast<float4*>(orow + j) = v;
        } else {
            for (int64_t q = j; q < n; ++q) orow[q] = ((b >> (q - j)) & 1u) ? 1.f : 0.f;
        }
    }
}

}  // namespace btx

using namespace btx;

BTX_API int btx_version(void) { return 100; }

// Rows are launched in chunks of at most 65535 (the grid's y limit).
constexpr int64_t kRowChunk = 65535;

BTX_API int btx_pack_bits(const float* x, int64_t rows, int64_t n, int64_t x_row_stride,
                          uint64_t* words, void* stream) {
    if (rows <= 0 || n <= 0) return 0;
    const int64_t nwords = (n + 63) / 64;
    for (int64_t r0 = 0; r0 < rows; r0 += kRowChunk) {
        const int64_t r = std::min(kRowChunk, rows - r0);
        hipLaunchKernelGGL(k_pack_bits, dim3((unsigned)((n + 255) / 256), (unsigned)r), dim3(256),
                           0, (hipStream_t)stream, x + r0 * x_row_stride, n, x_row_stride,
                           words + r0 * nwords, nwords);
    }
    return check_launch("btx_pack_bits");
}

BTX_API int btx_unpack_bits(const uint64_t* words, int64_t rows, int64_t n, float* out,
                            int64_t out_row_stride, void* stream) {
    if (rows <= 0 || n <= 0) return 0;
    const int64_t nwords = (n + 63) / 64;
    const bool vec = (reinterpret_cast<uintptr_t>(out) & 15) == 0 && out_row_stride % 4 == 0;
    for (int64_t r0 = 0; r0 < rows; r0 += kRowChunk) {
        const int64_t r = std::min(kRowChunk, rows - r0);
        const uint64_t* w = words + r0 * nwords;
        float* o = out + r0 * out_row_stride;
        if (vec)
            hipLaunchKernelGGL(k_unpack_bits16, dim3((unsigned)((n + 4095) / 4096), (unsigned)r),
                               dim3(256), 0, (hipStream_t)stream, w, n, nwords, o,
                               out_row_stride);
        else
            hipLaunchKernelGGL(k_unpack_bits, dim3((unsigned)((n + 255) / 256), (unsigned)r),
                               dim3(256), 0, (hipStream_t)stream, w, n, nwords, o,
                               out_row_stride);
    }
    return check_launch("btx_unpack_bits");
}

BTX_API const char* btx_last_error(void) { return g_last_error.c_str(); }

BTX_API int64_t btx_compact_tiles(int64_t n) {
    return n > 0 ? (n + kCTile - 1) / kCTile : 1;
}

BTX_API int btx_compact_dense(const float* x, int64_t lanes, int64_t n, int64_t x_lane_stride,
                              int32_t* idx, int64_t idx_lane_stride, int32_t* counts,
                              int32_t* tile_counts, void* stream) {
    if (lanes <= 0) return 0;
    if (n >= (int64_t)1 << 31) return fail("btx_compact_dense: n exceeds int32 indices");
    if (lanes > 65535) return fail("btx_compact_dense: lanes exceeds 65535");
    if (tile_counts == nullptr) return fail("btx_compact_dense: tile_counts is required");
    const int64_t ntiles = btx_compact_tiles(n);
    const dim3 grid((unsigned)ntiles, (unsigned)lanes);
    const hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_compact_count, grid, dim3(256), 0, s, x, n, x_lane_stride, ntiles,
                       tile_counts);
    hipLaunchKernelGGL(k_compact_write, grid, dim3(256), 0, s, x, n, x_lane_stride, ntiles,
                       tile_counts, idx, idx_lane_stride, counts);
    return check_launch("btx_compact_dense");
}

BTX_API int btx_compact_bits(const uint64_t* words, int64_t lanes, int64_t n,
                             int64_t words_lane_stride, int32_t* idx, int64_t idx_lane_stride,
                             int32_t* counts, void* stream) {
    if (lanes <= 0) return 0;
    if (n >= (int64_t)1 << 31) return fail("btx_compact_bits: n exceeds int32 indices");
    const int64_t nwords = (n + 63) / 64;
    hipLaunchKernelGGL(k_compact_bits, dim3((unsigned)lanes), dim3(1024), 0,
                       (hipStream_t)stream, words, n, nwords, words_lane_stride, idx,
                       idx_lane_stride, counts);
    return check_launch("btx_compact_bits");
}

BTX_API size_t btx_spmv_workspace_bytes(int64_t lanes, int64_t n_dst, int slots, int ordered) {
    return ws_layout(lanes, n_dst, slots, ordered).total;
}

// The event-driven product on a workspace whose term counts (ordered) or
// accumulators (fast) are zero on entry; they are zero again on return.
static int spmv_event_on(const int64_t* row_ptr, const int32_t* col, const float* w,
                         int64_t n_src, int64_t n_dst, int64_t nnz, const int64_t* csc_ptr,
                         const int32_t* csc_src, const float* csc_w, const int32_t* active,
                         const int32_t* counts, int64_t lanes, int64_t active_lane_stride,
                         const float* x, int64_t x_lane_stride, float* out, int ordered,
                         int slots, char* ws, const WsLayout& L, hipStream_t s) {
    const int64_t m = lanes * n_dst;
    PushArgs a{};
    a.row_ptr = row_ptr;
    a.col = col;
    a.w = w;
    a.n_dst = n_dst;
    a.ring_idx = active;
    a.ring_cnt = counts;
    a.ring_cap = active_lane_stride;
    a.ring_lanes = lanes;
    a.lane_base = 0;
    a.lanes = (int)lanes;
    a.nbins = 1;
    a.csr_d = 1;
    a.bin_slot[0] = 0;
    a.bin_sub[0] = 0;
    a.x = x;
    a.x_stride = x_lane_stride;
    a.ordered = ordered;
    a.ctl = (Ctl*)(ws + L.ctl);
    a.K = slots;
    const unsigned rblocks = (unsigned)((m + 255) / 256);
    if (ordered) {
        a.cnt = (int*)(ws + L.cnt);
        a.slots = (int2*)(ws + L.slots);
    } else {
        a.acc = (float*)(ws + L.acc);
    }
    const double avg = n_src > 0 ? (double)nnz / (double)n_src : 64.0;
    int rc = launch_push(a, pick_group(avg), push_blocks(), s);
    if (rc) return rc;
    if (ordered) {
        hipLaunchKernelGGL(k_reduce_ordered, dim3(rblocks), dim3(256), 0, s, m, n_dst, a.cnt,
                           a.slots, slots, csc_ptr, csc_src, csc_w, x, x_lane_stride, out);
    } else {
        hipLaunchKernelGGL(k_copy_zero, dim3(rblocks), dim3(256), 0, s, a.acc, out, m);
    }
    return check_launch("btx_spmv_event reduce");
}

static int spmv_event_check(int64_t n_src, const float* x, const int64_t* csc_ptr,
                            const int32_t* csc_src, const float* csc_w, int ordered, int slots,
                            const char* who) {
    if (ordered && (slots < 2 || slots > kMaxSlots))
        return fail(std::string(who) + ": ordered mode needs 2 <= slots <= 16");
    if (ordered && (x == nullptr || csc_ptr == nullptr || csc_src == nullptr || csc_w == nullptr))
        return fail(std::string(who) + ": ordered mode needs x and the dst-major arrays");
    if (n_src >= (int64_t)1 << 31) return fail(std::string(who) + ": n_src exceeds int32");
    return 0;
}

BTX_API int btx_spmv_event(const int64_t* row_ptr, const int32_t* col, const float* w,
                           int64_t n_src, int64_t n_dst, int64_t nnz, const int64_t* csc_ptr,
                           const int32_t* csc_src, const float* csc_w, const int32_t* active,
                           const int32_t* counts, int64_t lanes, int64_t active_lane_stride,
                           const float* x, int64_t x_lane_stride, float* out, int ordered,
                           int slots, void* workspace, size_t workspace_bytes, void* stream) {
    if (lanes <= 0 || n_dst <= 0) return 0;
    if (int rc = spmv_event_check(n_src, x, csc_ptr, csc_src, csc_w, ordered, slots,
                                  "btx_spmv_event"))
        return rc;
    const WsLayout L = ws_layout(lanes, n_dst, slots, ordered);
    if (workspace == nullptr || workspace_bytes < L.total)
        return fail("btx_spmv_event: workspace too small");
    return spmv_event_on(row_ptr, col, w, n_src, n_dst, nnz, csc_ptr, csc_src, csc_w, active,
                         counts, lanes, active_lane_stride, x, x_lane_stride, out, ordered,
                         slots, (char*)workspace, L, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// The product without a caller workspace: a file-scope device array serves
// as its scratch.  Each GPU holds its own copy of the array, at its own
// address, until the process ends; host code reaches it only through the
// address hipGetSymbolAddress returns on that device.
// ---------------------------------------------------------------------------

constexpr size_t kScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// The scratch array of the current device (looked up once per device).
static int scratch_of_current_device(char** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(std::string("btx_spmv_event_noalloc: hipGetDevice: ") +
                                     hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail("btx_spmv_event_noalloc: device id out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return fail(std::string("btx_spmv_event_noalloc: hipGetSymbolAddress: ") +
                        hipGetErrorString(e));
        g_scratch_ptr[dev] = (char*)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

__global__ __launch_bounds__(256) void k_clear_words(uint32_t* __restrict__ p, int64_t words) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < words) p[i] = 0u;
}

BTX_API size_t btx_spmv_event_noalloc_max_bytes(void) { return kScratchBytes; }

BTX_API int btx_spmv_event_noalloc(const int64_t* row_ptr, const int32_t* col, const float* w,
                                   int64_t n_src, int64_t n_dst, int64_t nnz,
                                   const int64_t* csc_ptr, const int32_t* csc_src,
                                   const float* csc_w, const int32_t* active,
                                   const int32_t* counts, int64_t lanes,
                                   int64_t active_lane_stride, const float* x,
                                   int64_t x_lane_stride, float* out, int ordered, int slots,
                                   void* stream) {
    if (lanes <= 0 || n_dst <= 0) return 0;
    if (int rc = spmv_event_check(n_src, x, csc_ptr, csc_src, csc_w, ordered, slots,
                                  "btx_spmv_event_noalloc"))
        return rc;
    const WsLayout L = ws_layout(lanes, n_dst, slots, ordered);
    if (L.total > kScratchBytes)
        return fail("btx_spmv_event_noalloc: lanes x destinations need " +
                    std::to_string(L.total) + " bytes of scratch, more than the " +
                    std::to_string(kScratchBytes) + " this entry point holds per device");
    char* ws = nullptr;
    if (int rc = scratch_of_current_device(&ws)) return rc;
    // the term counts / accumulators start at zero: clear them on this
    // stream (the array keeps whatever the last call on this device left)
    const size_t beg = ordered ? L.cnt : L.acc;
    const size_t end = ordered ? L.slots : L.total;
    const int64_t words = (int64_t)((end - beg) / 4);
    const hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_clear_words, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s,
                       (uint32_t*)(ws + beg), words);
    if (int rc = check_launch("btx_spmv_event_noalloc clear")) return rc;
    return spmv_event_on(row_ptr, col, w, n_src, n_dst, nnz, csc_ptr, csc_src, csc_w, active,
                         counts, lanes, active_lane_stride, x, x_lane_stride, out, ordered,
                         slots, ws, L, s);
}

BTX_API int btx_spmv_transpose(const int64_t* row_ptr, const int32_t* col, const float* w,
                               int64_t n_src, int64_t n_dst, const float* g, int64_t lanes,
                               float* gx, void* stream) {
    if (n_src <= 0 || lanes <= 0) return 0;
    hipLaunchKernelGGL(k_spmv_transpose, dim3((unsigned)((n_src + 3) / 4)), dim3(256), 0,
                       (hipStream_t)stream, row_ptr, col, w, n_src, n_dst, g, lanes, gx);
    return check_launch("btx_spmv_transpose");
}

BTX_API int btx_spmv_weight_grad(const int64_t* row_ptr, const int32_t* col, int64_t n_src,
                                 int64_t n_dst, const int32_t* active, const int32_t* counts,
                                 int64_t lanes, int64_t active_lane_stride, const float* x,
                                 int64_t x_lane_stride, const float* g, float* gw,
                                 void* stream) {
    (void)n_src;
    if (lanes <= 0) return 0;
    if (lanes > 65535) return fail("btx_spmv_weight_grad: too many lanes");
    hipLaunchKernelGGL(k_weight_grad, dim3(256, (unsigned)lanes), dim3(256), 0,
                       (hipStream_t)stream, row_ptr, col, n_dst, active, counts, lanes,
                       active_lane_stride, x, x_lane_stride, g, gw);
    return check_launch("btx_spmv_weight_grad");
}
