// Internal launch interface shared by the product and the step engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/btorch_amd.h"
#include "btx_common.h"

namespace btx {

struct PushArgs {
    const int64_t* row_ptr;
    const int32_t* col;
    const float* w;
    int64_t n_dst;
    const int32_t* ring_idx;  // [ring slots][ring_lanes][ring_cap]
    const int32_t* ring_cnt;  // [ring slots][ring_lanes]
    int64_t ring_cap;
    int64_t ring_lanes;       // lane stride of the ring arrays
    int64_t lane_base;        // first lane of this launch
    int lanes;                // lanes of this launch
    int nbins;
    int csr_d;
    int bin_slot[BTX_MAX_BINS];
    int bin_sub[BTX_MAX_BINS];
    const float* x;  // per-source values, NULL = spikes (1.0)
    int64_t x_stride;
    int ordered;
    int K;                    // terms past K are only counted; the reduction pulls them
    int* cnt;
    int2* slots;              // [K][m], m = ring_lanes * n_dst
    Ctl* ctl;
    float* acc;
};

int launch_push(const PushArgs& a, int group, int blocks, hipStream_t s);
int pick_group(double avg_row);
int push_blocks();

__global__ void k_compact_bits(const uint64_t* __restrict__ words, int64_t n, int64_t nwords,
                               int64_t wstride, int32_t* __restrict__ idx, int64_t istride,
                               int32_t* __restrict__ counts);

}  // namespace btx
