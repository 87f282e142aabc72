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

// Control block at the head of every workspace: per-run tables that kernels
// index with a run-time value.  (Indexing an array held in the kernel
// argument block with a run-time index makes the compiler copy the whole
// argument block to per-lane scratch memory; device-memory tables avoid it.)
struct Ctl {
    int age_of_sub[64];  // ring age read by virtual-row offset, -1 = not read
    int bin_delay[64];   // ring age of delay bin j
    int bin_sub[64];     // virtual-row offset of delay bin j
    // step engine, ordered mode: spikes pushed into each ring slot (all
    // lanes, summed one launch after the slot is written), and how the push
    // of the launch at clock c is delivered, at push_mode[c & 3] (decided
    // one launch ahead)
    int slot_total[64];
    int push_mode[4];    // 1 = a dense pull into the accumulator, 0 = term bins
    // the launch at clock c marks the lane bits of the slot it writes when
    // lane_bits_mode[c & 3] (decided one launch ahead, on while activity
    // nears the pull); lane_bits_run counts the consecutive launches, up to
    // the latest decided, that marked
    int lane_bits_mode[4];
    int lane_bits_run;
};

// Slots per destination of the ordered mode are held in registers when the
// terms are sorted, so their count is bounded.
constexpr int kMaxSlots = 16;

// Workspace layout shared by the standalone product and the step engine.
// Ordered mode: cnt[m] term counts and slots[K][m] (structure of arrays, so
// slot j of consecutive destinations is contiguous).  Fast mode: acc[m].
struct WsLayout {
    size_t ctl = 0, cnt = 0, slots = 0, acc = 0, total = 0;
};

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// `planes` accumulators of lanes*n_dst destinations each (one per future step
// of the engine's delivery ring; 1 for the standalone product).
inline WsLayout ws_layout(int64_t lanes, int64_t n_dst, int slots, int ordered,
                          int64_t planes = 1) {
    WsLayout L;
    const int64_t m = lanes * n_dst * planes;
    size_t off = 0;
    L.ctl = off;
    off = align256(off + sizeof(Ctl));
    if (ordered) {
        L.cnt = off;
        off = align256(off + sizeof(int) * m);
        L.slots = off;
        off = align256(off + sizeof(int2) * m * (int64_t)slots);
    } else {
        L.acc = off;
        off = align256(off + sizeof(float) * m);
    }
    L.total = off;
    return L;
}

// Sum the c <= K terms a destination received, in ascending source order,
// from a fresh +0.0f accumulator with one fmaf per term: the per-destination
// order of torch.sparse.mm on a coalesced (dst, src) COO on CPU
// (btorch/models/linear.py:474-483).  Term j of destination `key` is
// slots[j*m + key] = (source row, weight bits).  Sources are distinct per
// destination (coalesced matrix), so selecting strictly increasing source
// ids visits every term once; the terms stay in registers.
template <bool HASX>
__device__ __forceinline__ float ordered_sum(int64_t key, int c, const int2* __restrict__ slots,
                                             int64_t m, const float* __restrict__ xl) {
    if (!HASX) {
        const float w0 = __int_as_float(slots[key].y);
        if (c == 1) return 0.f + w0;
        if (c == 2) return (0.f + w0) + __int_as_float(slots[m + key].y);
    }
    int2 e[kMaxSlots];
#pragma unroll
    for (int j = 0; j < kMaxSlots; ++j)
        e[j] = j < c ? slots[(int64_t)j * m + key] : make_int2(0x7fffffff, 0);
    float acc = 0.f;
    int last = -1;
    for (int k = 0; k < c; ++k) {
        int best = 0x7fffffff;
        float bw = 0.f;
#pragma unroll
        for (int j = 0; j < kMaxSlots; ++j) {
            if (e[j].x > last && e[j].x < best) {
                best = e[j].x;
                bw = __int_as_float(e[j].y);
            }
        }
        const float xv = HASX ? xl[best] : 1.f;
        acc = fmaf(bw, xv, acc);
        last = best;
    }
    return acc;
}

}  // namespace btx
