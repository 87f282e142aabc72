// Delay-expanded connectivity built on the device.
//
// The reference builds a delayed connection on the host: expand_conn_for_delays
// turns (pre, post, delay, w) into a (n_pre*D, n_post) matrix whose row is
// pre*D + clip(delay, 0, D-1) (btorch/connectome/connection.py:973-1027), and
// BaseSparseConn transposes and coalesces it with scipy's sum_duplicates
// (btorch/models/linear.py:276-291): edges in (dst, virtual src) order,
// duplicates summed in their input order.  Here the same result comes from
// two stable radix sorts on 64-bit keys:
//
//   1. key dst<<32 | vsrc, value = edge id  -> coalesced order; a flag per
//      segment start and an inclusive scan number the unique edges; the
//      thread at a segment start sums the segment's weights in input order
//      as scipy does it (np.add.reduceat: the first weight plus NumPy's
//      pairwise sum of the others -- sequential below 8 terms, eight
//      accumulators up to 128, halves at multiples of 8 above);
//   2. key vsrc<<32 | dst, value = coalesced position -> the source-major
//      CSR (destinations ascending per row) and its permutation into the
//      coalesced order; row pointers by a lower-bound search per row.
//
// Entries past the unique count carry the key n_src<<32 so that both the
// sort and the pointer searches treat them as past the last row.  Every
// access is bounded by nnz or by the row counts, whatever the index values,
// so out-of-range input gives wrong output but no stray access (the caller
// validates the ranges).

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/functional.hpp>

#include "btx_common.h"

namespace btx {

namespace {

int bit_length(uint64_t v) {
    int b = 0;
    while (v) {
        ++b;
        v >>= 1;
    }
    return b;
}

struct BuildWs {
    size_t keys_a, keys_b, vals_a, vals_b, flags, pos, temp, temp_bytes, total;
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

BuildWs build_ws(int64_t nnz) {
    BuildWs L{};
    const size_t m = (size_t)(nnz > 0 ? nnz : 1);
    size_t sort_bytes = 0, scan_bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, sort_bytes, (const uint64_t*)nullptr,
                                    (uint64_t*)nullptr, (const uint32_t*)nullptr,
                                    (uint32_t*)nullptr, m, 0u, 64u);
    (void)rocprim::inclusive_scan(nullptr, scan_bytes, (const uint32_t*)nullptr,
                                  (uint32_t*)nullptr, m, rocprim::plus<uint32_t>());
    L.temp_bytes = sort_bytes > scan_bytes ? sort_bytes : scan_bytes;
    size_t off = 0;
    L.keys_a = off; off = align256(off + m * 8);
    L.keys_b = off; off = align256(off + m * 8);
    L.vals_a = off; off = align256(off + m * 4);
    L.vals_b = off; off = align256(off + m * 4);
    L.flags = off;  off = align256(off + m * 4);
    L.pos = off;    off = align256(off + m * 4);
    L.temp = off;   off = align256(off + L.temp_bytes);
    L.total = off;
    return L;
}

}  // namespace

// NumPy's pairwise float sum (numpy/_core/src/umath/loops_utils.h.src,
// pairwise_sum) of w[ids[a .. a+n)], n <= 128.
__device__ float pw_leaf(const float* __restrict__ w, const uint32_t* __restrict__ ids, int64_t a,
                         int64_t n) {
    if (n < 8) {
        float res = -0.0f;
        for (int64_t i = 0; i < n; ++i) res = res + w[ids[a + i]];
        return res;
    }
    float r[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = w[ids[a + j]];
    int64_t i = 8;
    for (; i < n - (n % 8); i += 8)
#pragma unroll
        for (int j = 0; j < 8; ++j) r[j] = r[j] + w[ids[a + i + j]];
    float res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res = res + w[ids[a + i]];
    return res;
}

// The same for any n: above 128 terms the range splits at n2 = n/2 rounded
// down to a multiple of 8 and the halves' sums are added (the recursion
// walked with an explicit stack; depth <= 24 for n < 2^31).
__device__ float pw_sum(const float* __restrict__ w, const uint32_t* __restrict__ ids, int64_t a0,
                        int64_t n0) {
    if (n0 <= 128) return pw_leaf(w, ids, a0, n0);
    int64_t fa[32], fn[32];
    float fl[32];
    int fs[32];  // 0: left half pending, 1: right half pending
    int sp = 0;
    fa[0] = a0;
    fn[0] = n0;
    fs[0] = 0;
    for (;;) {
        const int64_t n = fn[sp];
        if (n > 128) {
            int64_t n2 = n / 2;
            n2 -= n2 % 8;
            ++sp;
            fa[sp] = fa[sp - 1];
            fn[sp] = n2;
            fs[sp] = 0;
            continue;
        }
        float val = pw_leaf(w, ids, fa[sp], n);
        for (;;) {
            --sp;
            if (sp < 0) return val;
            if (fs[sp] == 0) {
                fl[sp] = val;
                fs[sp] = 1;
                int64_t n2 = fn[sp] / 2;
                n2 -= n2 % 8;
                ++sp;
                fa[sp] = fa[sp - 1] + n2;
                fn[sp] = fn[sp - 1] - n2;
                fs[sp] = 0;
                break;
            }
            val = fl[sp] + val;
        }
    }
}

__global__ __launch_bounds__(256) void k_edge_keys(const int64_t* __restrict__ pre,
                                                   const int64_t* __restrict__ post,
                                                   const int64_t* __restrict__ delay,
                                                   int64_t nnz, int64_t D,
                                                   uint64_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= nnz) return;
    int64_t d = delay ? delay[e] : 0;
    d = d < 0 ? 0 : (d > D - 1 ? D - 1 : d);
    const uint64_t vsrc = (uint64_t)(pre[e] * D + d);
    keys[e] = ((uint64_t)post[e] << 32) | (vsrc & 0xffffffffull);
    vals[e] = (uint32_t)e;
}

__global__ __launch_bounds__(256) void k_segment_starts(const uint64_t* __restrict__ keys,
                                                        int64_t nnz,
                                                        uint32_t* __restrict__ flags) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nnz) return;
    flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// One thread per segment start: the coalesced edge, its summed weight, and
// the key of the source-major sort.  Slots past the unique count get the
// padding key n_src<<32.
__global__ __launch_bounds__(256) void k_coalesce(const uint64_t* __restrict__ keys,
                                                  const uint32_t* __restrict__ ids,
                                                  const uint32_t* __restrict__ flags,
                                                  const uint32_t* __restrict__ pos,
                                                  const float* __restrict__ w, int64_t nnz,
                                                  int64_t n_src, int64_t* __restrict__ out_dst,
                                                  int64_t* __restrict__ out_src,
                                                  float* __restrict__ out_w,
                                                  uint64_t* __restrict__ skeys,
                                                  uint32_t* __restrict__ svals,
                                                  int64_t* __restrict__ nnz_out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nnz) return;
    const int64_t uniq = (int64_t)pos[nnz - 1];
    if (i == 0) *nnz_out = uniq;
    if (i >= uniq) {
        skeys[i] = (uint64_t)n_src << 32;
        svals[i] = (uint32_t)i;
    }
    if (!flags[i]) return;
    const uint64_t key = keys[i];
    int64_t end = i + 1;
    while (end < nnz && keys[end] == key) ++end;
    float sum = w[ids[i]];
    if (end - i > 1) sum = sum + pw_sum(w, ids, i + 1, end - i - 1);
    const int64_t u = (int64_t)pos[i] - 1;
    const int64_t dst = (int64_t)(key >> 32);
    const int64_t src = (int64_t)(key & 0xffffffffull);
    out_dst[u] = dst;
    out_src[u] = src;
    out_w[u] = sum;
    skeys[u] = ((uint64_t)src << 32) | (uint64_t)dst;
    svals[u] = (uint32_t)u;
}

__global__ __launch_bounds__(256) void k_csr_cols(const uint64_t* __restrict__ skeys,
                                                  const uint32_t* __restrict__ svals,
                                                  const int64_t* __restrict__ nnz_out,
                                                  int64_t nnz, int32_t* __restrict__ col,
                                                  int64_t* __restrict__ perm,
                                                  const int64_t* __restrict__ out_src,
                                                  int32_t* __restrict__ csc_src) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nnz || i >= *nnz_out) return;
    col[i] = (int32_t)(skeys[i] & 0xffffffffull);
    perm[i] = (int64_t)svals[i];
    csc_src[i] = (int32_t)out_src[i];
}

// ptr[r] = first position whose row (key >> 32) is >= r, for r in [0, rows].
__global__ __launch_bounds__(256) void k_row_ptr_keys(const uint64_t* __restrict__ skeys,
                                                      int64_t nnz, int64_t rows,
                                                      int64_t* __restrict__ ptr) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r > rows) return;
    const uint64_t target = (uint64_t)r << 32;
    int64_t lo = 0, hi = nnz;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (skeys[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    ptr[r] = lo;
}

// ptr[c] = first coalesced edge whose destination is >= c, for c in [0, cols].
__global__ __launch_bounds__(256) void k_row_ptr_dst(const int64_t* __restrict__ dst,
                                                     const int64_t* __restrict__ nnz_out,
                                                     int64_t cols, int64_t* __restrict__ ptr) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c > cols) return;
    int64_t lo = 0, hi = *nnz_out;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (dst[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    ptr[c] = lo;
}

__global__ __launch_bounds__(256) void k_fill_i64(int64_t* __restrict__ p, int64_t count,
                                                  int64_t value) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < count) p[i] = value;
}

}  // namespace btx

using namespace btx;

static unsigned blocks_of(int64_t count) { return (unsigned)((count + 255) / 256); }

BTX_API size_t btx_conn_build_workspace_bytes(int64_t nnz) { return build_ws(nnz).total; }

BTX_API int btx_conn_build(const int64_t* pre, const int64_t* post, const int64_t* delay,
                           const float* w, int64_t nnz, int64_t n_pre, int64_t n_post,
                           int n_delay_bins, int64_t* out_dst, int64_t* out_src, float* out_w,
                           int64_t* nnz_out, int64_t* row_ptr, int32_t* col, int64_t* perm,
                           int64_t* csc_ptr, int32_t* csc_src, void* workspace,
                           size_t workspace_bytes, void* stream) {
    const hipStream_t s = (hipStream_t)stream;
    if (nnz < 0 || n_pre < 0 || n_post < 0 || n_delay_bins < 1)
        return fail("btx_conn_build: negative size or n_delay_bins < 1");
    const int64_t n_src = n_pre * (int64_t)n_delay_bins;
    if (n_src >= ((int64_t)1 << 31) || n_post >= ((int64_t)1 << 31))
        return fail("btx_conn_build: rows and columns must fit int32 indices");
    if (nnz >= ((int64_t)1 << 31)) return fail("btx_conn_build: at most 2^31 - 1 edges");
    if (nnz_out == nullptr || row_ptr == nullptr || csc_ptr == nullptr)
        return fail("btx_conn_build: nnz_out and both row pointers are required");
    if (nnz == 0) {
        hipLaunchKernelGGL(k_fill_i64, dim3(blocks_of(n_src + 1)), dim3(256), 0, s, row_ptr,
                           n_src + 1, (int64_t)0);
        hipLaunchKernelGGL(k_fill_i64, dim3(blocks_of(n_post + 1)), dim3(256), 0, s, csc_ptr,
                           n_post + 1, (int64_t)0);
        hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(256), 0, s, nnz_out, (int64_t)1, (int64_t)0);
        return check_launch("btx_conn_build");
    }
    if (pre == nullptr || post == nullptr || w == nullptr || out_dst == nullptr ||
        out_src == nullptr || out_w == nullptr || col == nullptr || perm == nullptr ||
        csc_src == nullptr)
        return fail("btx_conn_build: null edge or output array");
    const BuildWs L = build_ws(nnz);
    if (workspace == nullptr || workspace_bytes < L.total)
        return fail("btx_conn_build: workspace too small");
    char* ws = (char*)workspace;
    uint64_t* ka = (uint64_t*)(ws + L.keys_a);
    uint64_t* kb = (uint64_t*)(ws + L.keys_b);
    uint32_t* va = (uint32_t*)(ws + L.vals_a);
    uint32_t* vb = (uint32_t*)(ws + L.vals_b);
    uint32_t* flags = (uint32_t*)(ws + L.flags);
    uint32_t* pos = (uint32_t*)(ws + L.pos);
    void* temp = ws + L.temp;
    size_t temp_bytes = L.temp_bytes;
    const int n = (int)nnz;
    const unsigned g = blocks_of(nnz);

    hipLaunchKernelGGL(k_edge_keys, dim3(g), dim3(256), 0, s, pre, post, delay, nnz,
                       (int64_t)n_delay_bins, ka, va);
    // coalesced order: (dst, virtual src), stable
    const int end1 = 32 + bit_length((uint64_t)(n_post > 0 ? n_post - 1 : 0));
    if (rocprim::radix_sort_pairs(temp, temp_bytes, ka, kb, va, vb, (size_t)n, 0u,
                                  (unsigned)end1, s) != hipSuccess)
        return fail("btx_conn_build: radix sort failed");
    hipLaunchKernelGGL(k_segment_starts, dim3(g), dim3(256), 0, s, kb, nnz, flags);
    temp_bytes = L.temp_bytes;
    if (rocprim::inclusive_scan(temp, temp_bytes, flags, pos, (size_t)n,
                                rocprim::plus<uint32_t>(), s) != hipSuccess)
        return fail("btx_conn_build: scan failed");
    hipLaunchKernelGGL(k_coalesce, dim3(g), dim3(256), 0, s, kb, vb, flags, pos, w, nnz, n_src,
                       out_dst, out_src, out_w, ka, va, nnz_out);
    // source-major order: (virtual src, dst); padding sorts last
    const int end2 = 32 + bit_length((uint64_t)n_src);
    temp_bytes = L.temp_bytes;
    if (rocprim::radix_sort_pairs(temp, temp_bytes, ka, kb, va, vb, (size_t)n, 0u,
                                  (unsigned)end2, s) != hipSuccess)
        return fail("btx_conn_build: radix sort failed");
    hipLaunchKernelGGL(k_csr_cols, dim3(g), dim3(256), 0, s, kb, vb, nnz_out, nnz, col, perm,
                       out_src, csc_src);
    hipLaunchKernelGGL(k_row_ptr_keys, dim3(blocks_of(n_src + 1)), dim3(256), 0, s, kb, nnz,
                       n_src, row_ptr);
    hipLaunchKernelGGL(k_row_ptr_dst, dim3(blocks_of(n_post + 1)), dim3(256), 0, s, out_dst,
                       nnz_out, n_post, csc_ptr);
    return check_launch("btx_conn_build");
}
