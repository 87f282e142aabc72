// Shared device helpers for the btorch_amd gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#define BTX_API extern "C" __attribute__((visibility("default")))

namespace btx {

// Thread-local message returned by btx_last_error().
void set_error(const std::string& msg);
int fail(const std::string& msg);
int check_launch(const char* what);

constexpr int kWave = 64;

// Inclusive scan across the 64 lanes of a wavefront.
__device__ __forceinline__ int wave_incl_scan(int v) {
    const int ln = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(v, o, 64);
        if (ln >= o) v += t;
    }
    return v;
}

// Exclusive scan across a workgroup of NW wavefronts; `total` receives the sum.
template <int NW>
__device__ __forceinline__ int block_excl_scan(int v, int& total) {
    __shared__ int s_w[NW];
    __shared__ int s_x[NW + 1];
    const int incl = wave_incl_scan(v);
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) s_w[wid] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int i = 0; i < NW; ++i) {
            s_x[i] = run;
            run += s_w[i];
        }
        s_x[NW] = run;
    }
    __syncthreads();
    total = s_x[NW];
    const int r = s_x[wid] + incl - v;
    __syncthreads();
    return r;
}

// Control words at the head of every workspace.
struct Ctl {
    int ovf_count;   // entries appended to the overflow list this step
    int ovf_error;   // sticky: overflow list ran out of capacity
    int pad[62];
};

// Workspace layout shared by the standalone product and the step engine.
struct WsLayout {
    size_t ctl = 0, cnt = 0, slots = 0, ovf = 0, acc = 0, total = 0;
    int64_t ovf_cap = 0;
};

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

inline WsLayout ws_layout(int64_t lanes, int64_t n_dst, int slots, int ordered) {
    WsLayout L;
    const int64_t m = lanes * n_dst;
    size_t off = 0;
    L.ctl = off;
    off = align256(off + sizeof(Ctl));
    if (ordered) {
        L.cnt = off;
        off = align256(off + sizeof(int) * m);
        L.slots = off;
        off = align256(off + sizeof(int2) * m * (int64_t)slots);
        L.ovf_cap = m < 4096 ? 4096 : (m > (1 << 20) ? (1 << 20) : m);
        L.ovf = off;
        off = align256(off + sizeof(int4) * L.ovf_cap);
    } else {
        L.acc = off;
        off = align256(off + sizeof(float) * m);
    }
    L.total = off;
    return L;
}

// Sum the terms a destination received this step in ascending source order,
// starting from a fresh +0.0f accumulator with one fmaf per term.  This is
// the per-destination order of torch.sparse.mm on a coalesced (dst, src) COO
// on CPU (btorch/models/linear.py:474-483).  Terms live in `slots` (the first
// min(c, K) of them) and, past K, in the overflow list keyed by `key`.
// Sources are distinct per destination (coalesced matrix), so a selection by
// strictly increasing source id visits every term exactly once.
template <bool HASX>
__device__ __forceinline__ float ordered_sum(int64_t key, int c, const int2* __restrict__ slots,
                                             int K, const int4* __restrict__ ovf, int novf,
                                             const float* __restrict__ xl) {
    const int2* sl = slots + key * K;
    if (!HASX) {
        if (c == 1) return 0.f + __int_as_float(sl[0].y);
        if (c == 2) return (0.f + __int_as_float(sl[0].y)) + __int_as_float(sl[1].y);
    }
    const int nk = c < K ? c : K;
    const int klo = (int)(key & 0xffffffff);
    const int khi = (int)(key >> 32);
    float acc = 0.f;
    int last = -1;
    for (int k = 0; k < c; ++k) {
        int best = 0x7fffffff;
        float bw = 0.f;
        for (int j = 0; j < nk; ++j) {
            const int2 e = sl[j];
            if (e.x > last && e.x < best) {
                best = e.x;
                bw = __int_as_float(e.y);
            }
        }
        if (c > K) {
            for (int j = 0; j < novf; ++j) {
                const int4 o = ovf[j];
                if (o.x == klo && o.y == khi && o.z > last && o.z < best) {
                    best = o.z;
                    bw = __int_as_float(o.w);
                }
            }
        }
        // every term was found unless an overflow list ran out of room (the
        // run is flagged invalid then); never index with the sentinel
        if (best == 0x7fffffff) break;
        const float xv = HASX ? xl[best] : 1.f;
        acc = fmaf(bw, xv, acc);
        last = best;
    }
    return acc;
}

}  // namespace btx
