// Fused recurrent time step on gfx950.
//
// Per step t, in stream order:
//   F(t)  one elementwise pass over (lane, neuron): the current psc_{t-1}
//         (already holding the previous push), LIF charge/fire/refractory/
//         reset, the exponential-synapse decay psc_dec, spike words -> raster,
//         and the pushed spikes (own or teacher's) as ascending lists per
//         1024-neuron segment into ring slot (cursor + t) % ring_d.
//   P(t)  event-driven push: for every delay bin j, the active neurons of the
//         ring slot bin j points at, each streaming its virtual row
//         n*csr_d + bin_sub[j] of the CSR.
//         ordered: terms are binned by (lane, 4096-destination tile) with one
//                  global reservation per workgroup and tile;
//         fast:    terms are added to psc with float atomics.
//   R(t)  (ordered) one workgroup per bin sorts the bin's terms in LDS by
//         destination, then each destination's terms by virtual row, and adds
//         the ascending sum to psc; a destination with many terms, or every
//         destination of an overflowing bin, pulls its input over its
//         incoming edges in the same order instead.
// After the last step FIN normalises psc (psc + 0.0f).
//
// Op order follows the reference's eager expressions exactly (compiled with
// -ffp-contract=off):
//   LIF charge  v + dt*((-(v - v_reset))/tau + I/c_m)   lif.py:144-154, ode.py:53-55
//   fire        ((v - v_th)/(v_th - v_reset)) >= 0       lif.py:160-168, surrogate/base.py:6-8
//   refractory  s * (ref == 0) ; relu((ref + s*tau_ref) - dt)   lif.py:165-167, 184-187
//   reset soft  v - (v_th - v_reset)*s ; hard v - (v - v_reset)*s   lif.py:176-180
//   exp-Euler   psc + k*((-psc)/tau_s), k = expm1(dt*l)/l, l = -1/tau_s  ode.py:50, synapse.py:171-178
//   adaptation  psc + W^T z, W^T z summed per destination from +0.0f in
//               ascending source row order (torch.sparse.mm on the coalesced
//               (dst, src) matrix, linear.py:474-483)    synapse.py:82-87
// The reference adds the product's +0.0f to destinations without terms,
// which turns a -0.0f psc into +0.0f; the next step's psc + 0.0f does the same.

#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "btx_common.h"
#include "btx_spmv.h"

namespace btx {

constexpr int kChunk = BTX_RING_CHUNK;  // neurons per fused-step workgroup / ring segment
constexpr int kTileShift = 12;
constexpr int kTile = 1 << kTileShift;  // destinations per ordered-mode bin
constexpr int kMaxTiles = 4096;         // n <= 16.7M
constexpr int kBinCap = 6144;           // terms per bin held in LDS by the reduction
constexpr int kSortMax = 32;            // most terms a destination sorts in LDS
constexpr int kRedThreads = 512;
constexpr int kPushThreads = 256;

static_assert(kChunk == 1024, "fused-step workgroups cover 1024 neurons");

struct StepArgs {
    int64_t n;
    int64_t nwords;
    int64_t nchunks;
    float dt;
    int hard;
    int use_ref;
    int first;                   // first step of the run: psc is the stored state
    const float* tau;      int64_t tau_s;
    const float* vth;      int64_t vth_s;
    const float* vr;       int64_t vr_s;
    const float* cm;       int64_t cm_s;
    const float* tref;     int64_t tref_s;
    const float* tsyn;     int64_t tsyn_s;
    const float* ksyn;     int64_t ksyn_s;
    float* v;
    float* ref;
    float* psc;
    const float* x;        // step t, (lanes, n) or NULL
    uint64_t* raster;      // step t, (lanes, nwords)
    const uint64_t* teacher;  // step t, (lanes, nwords) or NULL
    int32_t* ring_idx;     // this step's ring slot, [lanes][n], or NULL
    int32_t* ring_cnt;     // this step's ring slot, [lanes][nchunks]
    uint64_t* ring_words;  // this step's ring slot, [lanes][nwords]
    float* rec_z;          // step t or NULL
    float* rec_v;
    float* rec_ref;
    float* rec_psc_prev;   // step t-1 or NULL
    float* rec_v_pre;
    float* rec_v_prev;
    // in-kernel Poisson drive
    const uint64_t* in_seed;
    const float* in_prob;
    const int64_t* in_t0;
    float in_kick;
    int64_t t;             // step of this launch within the run
    int* counts;
    float* rec_sample;     // step t, [lanes][n_sample]
    int64_t n_sample;
    const int64_t* lane_t_end;
};

__device__ __forceinline__ float prm(const float* p, int64_t stride, int64_t n) {
    return p[stride * n];
}

// SplitMix64 finaliser (Steele, Lea, Flood 2014): a bijective 64-bit mix.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// In-kernel Poisson drive of neuron nidx of `lane` at this step.
__device__ __forceinline__ float poisson_drive(const StepArgs& a, int64_t lane, int64_t nidx) {
    if (a.in_prob == nullptr) return 0.f;
    const uint64_t s = (uint64_t)(a.in_t0[lane] + a.t);
    const uint64_t h = mix64(a.in_seed[lane] ^ mix64(s * 0x100000001b3ull + (uint64_t)nidx));
    const float u = (float)(h >> 40) * (1.0f / 16777216.0f);
    return u < a.in_prob[lane] ? a.in_kick : 0.f;
}

// One neuron of the fused step.  `psc` enters as psc_{t-1} (decayed current
// of the previous step plus its push), leaves as this step's decayed current.
// Returns the spike (0/1).
__device__ __forceinline__ float lif_exp_neuron(const StepArgs& a, int64_t lane, int64_t nidx,
                                                int64_t i, float& psc, float xin, float& v,
                                                float& ref) {
    if (!a.first) {
        psc = psc + 0.f;
        if (a.rec_psc_prev) a.rec_psc_prev[i] = psc;
    }
    xin = xin + poisson_drive(a, lane, nidx);
    const float I = a.psc ? psc + xin : xin;
    const float tau = prm(a.tau, a.tau_s, nidx);
    const float vth = prm(a.vth, a.vth_s, nidx);
    const float vr = prm(a.vr, a.vr_s, nidx);
    const float cm = prm(a.cm, a.cm_s, nidx);
    if (a.rec_v_prev) a.rec_v_prev[i] = v;
    const float dv = (-(v - vr)) / tau + I / cm;
    const float v1 = v + a.dt * dv;
    if (a.rec_v_pre) a.rec_v_pre[i] = v1;
    const float u = (v1 - vth) / (vth - vr);
    float s = (u >= 0.f) ? 1.f : 0.f;
    if (a.use_ref) {
        s = s * ((ref == 0.f) ? 1.f : 0.f);
        float r2 = (ref + s * prm(a.tref, a.tref_s, nidx)) - a.dt;
        ref = (r2 < 0.f) ? 0.f : r2;
        if (a.rec_ref) a.rec_ref[i] = ref;
    }
    v = a.hard ? (v1 - (v1 - vr) * s) : (v1 - (vth - vr) * s);
    if (a.rec_v) a.rec_v[i] = v;
    const bool live = a.lane_t_end == nullptr || a.in_t0[lane] + a.t < a.lane_t_end[lane];
    if (a.counts && s != 0.f && live) a.counts[i] += 1;
    if (a.rec_sample && nidx < a.n_sample && live) a.rec_sample[lane * a.n_sample + nidx] = v;
    if (a.rec_z) a.rec_z[i] = s;
    if (a.psc) {
        const float tsyn = prm(a.tsyn, a.tsyn_s, nidx);
        const float k = prm(a.ksyn, a.ksyn_s, nidx);
        psc = psc + k * ((-psc) / tsyn);
    }
    return s;
}

// Ascending list of one 1024-neuron segment: each thread of the NW-wave
// workgroup holds the pushed bits `bits` of neurons seg_base + first + b
// (bit b, b < 32).  Threads own increasing neuron ranges in thread order.
// Every thread of the workgroup must call it.
template <int NW>
__device__ __forceinline__ void write_segment(uint32_t bits, int first, int64_t seg_base,
                                              int32_t* __restrict__ idx,
                                              int32_t* __restrict__ cnt) {
    int total;
    int pos = block_excl_scan<NW>(__popc(bits), total);
    while (bits) {
        const int b = __ffs(bits) - 1;
        idx[seg_base + pos++] = (int32_t)(seg_base + first + b);
        bits &= bits - 1;
    }
    if (threadIdx.x == 0) *cnt = total;
}

// Scalar form: one neuron per thread, 1024 neurons per workgroup, spike word
// by wavefront ballot.
__global__ __launch_bounds__(1024) void k_lif_exp_step(StepArgs a) {
    const int64_t lane = blockIdx.y;
    const int64_t seg_base = (int64_t)blockIdx.x * kChunk;
    const int64_t nidx = seg_base + threadIdx.x;
    int spk = 0;
    if (nidx < a.n) {
        const int64_t i = lane * a.n + nidx;
        float psc = a.psc ? a.psc[i] : 0.f;
        const float xin = a.x ? a.x[i] : 0.f;
        float v = a.v[i];
        float ref = a.use_ref ? a.ref[i] : 0.f;
        const float s = lif_exp_neuron(a, lane, nidx, i, psc, xin, v, ref);
        a.v[i] = v;
        if (a.use_ref) a.ref[i] = ref;
        if (a.psc) a.psc[i] = psc;
        spk = s != 0.f;
    }
    const uint64_t m = __ballot(spk);
    const int64_t word = nidx >> 6;
    if (a.raster != nullptr && (threadIdx.x & 63) == 0 && word < a.nwords)
        a.raster[lane * a.nwords + word] = m;
    if (a.ring_idx != nullptr) {
        uint32_t pushed = (uint32_t)spk;
        if (a.teacher != nullptr)
            pushed = nidx < a.n
                         ? (uint32_t)((a.teacher[lane * a.nwords + word] >> (nidx & 63)) & 1ull)
                         : 0u;
        const uint64_t pw = a.teacher != nullptr ? __ballot(pushed != 0u) : m;
        if ((threadIdx.x & 63) == 0 && word < a.nwords) a.ring_words[lane * a.nwords + word] = pw;
        write_segment<16>(pushed, threadIdx.x, seg_base, a.ring_idx + lane * a.n,
                          a.ring_cnt + lane * a.nchunks + blockIdx.x);
    }
}

// Vector form (n % 4 == 0, 16-byte aligned state): four consecutive neurons
// per thread with 16-byte loads/stores, 1024 neurons per workgroup; the
// 64-bit spike word of 16 threads is assembled by an OR across them.
__global__ __launch_bounds__(256) void k_lif_exp_step4(StepArgs a) {
    const int64_t lane = blockIdx.y;
    const int64_t seg_base = (int64_t)blockIdx.x * kChunk;
    const int64_t n0 = seg_base + (int64_t)threadIdx.x * 4;
    uint32_t nib = 0;
    if (n0 < a.n) {
        const int64_t i0 = lane * a.n + n0;
        float4 v4 = *(const float4*)(a.v + i0);
        float4 p4 = a.psc ? *(const float4*)(a.psc + i0) : make_float4(0.f, 0.f, 0.f, 0.f);
        float4 r4 = a.use_ref ? *(const float4*)(a.ref + i0) : make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 x4 = a.x ? *(const float4*)(a.x + i0) : make_float4(0.f, 0.f, 0.f, 0.f);
        const float s0 = lif_exp_neuron(a, lane, n0, i0, p4.x, x4.x, v4.x, r4.x);
        const float s1 = lif_exp_neuron(a, lane, n0 + 1, i0 + 1, p4.y, x4.y, v4.y, r4.y);
        const float s2 = lif_exp_neuron(a, lane, n0 + 2, i0 + 2, p4.z, x4.z, v4.z, r4.z);
        const float s3 = lif_exp_neuron(a, lane, n0 + 3, i0 + 3, p4.w, x4.w, v4.w, r4.w);
        *(float4*)(a.v + i0) = v4;
        if (a.use_ref) *(float4*)(a.ref + i0) = r4;
        if (a.psc) *(float4*)(a.psc + i0) = p4;
        nib = (uint32_t)(s0 != 0.f) | ((uint32_t)(s1 != 0.f) << 1) |
              ((uint32_t)(s2 != 0.f) << 2) | ((uint32_t)(s3 != 0.f) << 3);
    }
    const int64_t wi = n0 >> 6;
    uint64_t word = (uint64_t)nib << (4 * (threadIdx.x & 15));
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) word |= __shfl_xor(word, o, 64);
    if (a.raster != nullptr && (threadIdx.x & 15) == 0 && wi < a.nwords)
        a.raster[lane * a.nwords + wi] = word;
    if (a.ring_idx != nullptr) {
        uint32_t pushed = nib;
        uint64_t pw = word;
        if (a.teacher != nullptr) {
            pw = wi < a.nwords ? a.teacher[lane * a.nwords + wi] : 0ull;
            pushed = n0 < a.n ? (uint32_t)((pw >> (4 * (threadIdx.x & 15))) & 15ull) : 0u;
        }
        if ((threadIdx.x & 15) == 0 && wi < a.nwords) a.ring_words[lane * a.nwords + wi] = pw;
        write_segment<4>(pushed, threadIdx.x * 4, seg_base, a.ring_idx + lane * a.n,
                         a.ring_cnt + lane * a.nchunks + blockIdx.x);
    }
}

// Ring lists of all ring slots from spike words [slots][lanes][nwords]
// (grid: segments x lanes x slots; thread = 4 neurons of the segment).
__global__ __launch_bounds__(256) void k_ring_from_words(const uint64_t* __restrict__ words,
                                                         int64_t n, int64_t nwords,
                                                         int64_t lanes, int64_t nchunks,
                                                         int32_t* __restrict__ ring_idx,
                                                         int32_t* __restrict__ ring_cnt) {
    const int64_t seg = blockIdx.x, lane = blockIdx.y, slot = blockIdx.z;
    const int64_t seg_base = seg * kChunk;
    const int64_t n0 = seg_base + (int64_t)threadIdx.x * 4;
    const int64_t wi = n0 >> 6;
    uint32_t nib = 0;
    if (n0 < n) {
        const uint64_t w = words[(slot * lanes + lane) * nwords + wi];
        nib = (uint32_t)((w >> (4 * (threadIdx.x & 15))) & 15ull);
        const int64_t room = n - n0;  // bits past n never count
        if (room < 4) nib &= (1u << room) - 1u;
    }
    write_segment<4>(nib, threadIdx.x * 4, seg_base, ring_idx + (slot * lanes + lane) * n,
                     ring_cnt + (slot * lanes + lane) * nchunks + seg);
}

// ---------------------------------------------------------------------------
// Push
// ---------------------------------------------------------------------------

struct PushStep {
    const int64_t* row_ptr;
    const int32_t* col;
    const float* w;
    int64_t n;
    int64_t lanes;
    int64_t nchunks;
    int cpb;                     // segments per workgroup
    int nbins;
    int csr_d;
    int ring_d;
    int64_t clock;               // cursor + t: ring slot of this step's spikes
    const Ctl* ctl;              // bin_delay / bin_sub tables
    const int32_t* ring_idx;     // [ring_d][lanes][n]
    const int32_t* ring_cnt;     // [ring_d][lanes][nchunks]
    int ntiles;
    int cap;
    int* bin_cnt;                // [lanes][ntiles]
    int32_t* ev_r;               // [lanes][ntiles][cap]
    float* ev_w;
    uint16_t* ev_d;
    float* psc;                  // fast mode target
};

__device__ __forceinline__ int ring_slot(int64_t clock, int delay, int ring_d) {
    const int64_t s = (clock - delay) % ring_d;
    return (int)(s < 0 ? s + ring_d : s);
}

// Workgroup (x, lane) pushes segments [x*cpb, x*cpb + cpb) of every delay bin
// of its lane; groups of G threads stream one active row each.
template <int G, bool ORDERED>
__global__ __launch_bounds__(kPushThreads) void k_push_step(PushStep a) {
    __shared__ int s_off[kPushThreads + 1];
    __shared__ int s_slot[BTX_MAX_BINS];
    __shared__ int s_sub[BTX_MAX_BINS];
    __shared__ int s_tile[ORDERED ? kMaxTiles : 1];
    __shared__ int s_base[ORDERED ? kMaxTiles : 1];
    const int64_t lane = blockIdx.y;
    const int64_t c0 = (int64_t)blockIdx.x * a.cpb;
    const int ncb = (int)(a.nchunks - c0 < a.cpb ? a.nchunks - c0 : a.cpb);
    const int tid = threadIdx.x;
    if (tid < a.nbins) {
        s_slot[tid] = ring_slot(a.clock, a.ctl->bin_delay[tid], a.ring_d);
        s_sub[tid] = a.ctl->bin_sub[tid];
    }
    if (ORDERED)
        for (int t = tid; t < a.ntiles; t += kPushThreads) s_tile[t] = 0;
    __syncthreads();
    const int nseg = a.nbins * ncb;  // <= kPushThreads (host check)
    int c = 0;
    if (tid < nseg) {
        const int j = tid / ncb, cc = tid - j * ncb;
        c = a.ring_cnt[((int64_t)s_slot[j] * a.lanes + lane) * a.nchunks + c0 + cc];
        // a segment never lists more neurons than it has
        const int64_t room = a.n - (c0 + cc) * kChunk;
        const int cap = room < kChunk ? (int)room : kChunk;
        c = c < 0 ? 0 : (c > cap ? cap : c);
    }
    int total;
    const int off = block_excl_scan<kPushThreads / 64>(c, total);
    if (tid < nseg) s_off[tid] = off;
    if (tid == 0) s_off[nseg] = total;
    __syncthreads();
    if (total == 0) return;

    constexpr int gpb = kPushThreads / G;
    const int gl = tid % G;
    const int64_t key0 = lane * a.n;
    for (int pass = 0; pass < (ORDERED ? 2 : 1); ++pass) {
        for (int item = tid / G; item < total; item += gpb) {
            int lo = 0, hi = nseg;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_off[mid] <= item) lo = mid;
                else hi = mid;
            }
            const int j = lo / ncb, cc = lo - j * ncb;
            const int k = item - s_off[lo];
            const int src = a.ring_idx[((int64_t)s_slot[j] * a.lanes + lane) * a.n +
                                       (c0 + cc) * kChunk + k];
            if ((uint32_t)src >= (uint32_t)a.n) continue;  // never indexes past the CSR
            const int r = src * a.csr_d + s_sub[j];
            const int64_t beg = a.row_ptr[r], end = a.row_ptr[r + 1];
            for (int64_t e = beg + gl; e < end; e += G) {
                const int dst = a.col[e];
                if (!ORDERED) {
                    atomicAdd(&a.psc[key0 + dst], a.w[e]);
                } else if (pass == 0) {
                    atomicAdd(&s_tile[dst >> kTileShift], 1);
                } else {
                    const int tl = dst >> kTileShift;
                    const int pos = s_base[tl] + atomicAdd(&s_tile[tl], 1);
                    if (pos < a.cap) {
                        const int64_t q = (lane * a.ntiles + tl) * (int64_t)a.cap + pos;
                        a.ev_r[q] = r;
                        a.ev_w[q] = a.w[e];
                        a.ev_d[q] = (uint16_t)(dst & (kTile - 1));
                    }
                }
            }
        }
        if (ORDERED && pass == 0) {
            __syncthreads();
            for (int t = tid; t < a.ntiles; t += kPushThreads) {
                const int ct = s_tile[t];
                if (ct) {
                    s_base[t] = atomicAdd(&a.bin_cnt[lane * a.ntiles + t], ct);
                    s_tile[t] = 0;
                }
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// Ordered reduction
// ---------------------------------------------------------------------------

struct PullArgs {
    const int64_t* csc_ptr;
    const int32_t* csc_src;
    const float* csc_w;
    const uint64_t* ring_words;  // [ring_d][lanes][nwords], ring-slot order
    int64_t lanes;
    int64_t nwords;
    int ring_d;
    int64_t clock;               // cursor + step of the push being reduced
    int csr_d;
    const int* age_of_sub;       // [csr_d] ring age by virtual-row offset, -1 = not read
};

// Virtual row r = src*csr_d + sub is active in the push being reduced when
// neuron src spiked age_of_sub[sub] steps before it (ring slot clock - age).
__device__ __forceinline__ bool pulled_active(const PullArgs& a, int64_t lane, int r) {
    const int src = r / a.csr_d;
    const int age = a.age_of_sub[r - src * a.csr_d];
    if (age < 0) return false;
    int64_t slot = (a.clock - age) % a.ring_d;
    if (slot < 0) slot += a.ring_d;
    const uint64_t* row = a.ring_words + (slot * a.lanes + lane) * a.nwords;
    return (row[src >> 6] >> (src & 63)) & 1ull;
}

// Input of destination `dst` of `lane`: its active incoming edges summed in
// the coalesced (dst, src) order from +0.0f.  Returns false when none is active.
__device__ bool pull_sum(const PullArgs& a, int64_t lane, int64_t dst, float& out) {
    float acc = 0.f;
    bool any = false;
    int64_t e = a.csc_ptr[dst];
    const int64_t e1 = a.csc_ptr[dst + 1];
    for (; e + 4 <= e1; e += 4) {
        const int r0 = a.csc_src[e], r1 = a.csc_src[e + 1];
        const int r2 = a.csc_src[e + 2], r3 = a.csc_src[e + 3];
        const bool b0 = pulled_active(a, lane, r0), b1 = pulled_active(a, lane, r1);
        const bool b2 = pulled_active(a, lane, r2), b3 = pulled_active(a, lane, r3);
        if (b0) acc = acc + a.csc_w[e];
        if (b1) acc = acc + a.csc_w[e + 1];
        if (b2) acc = acc + a.csc_w[e + 2];
        if (b3) acc = acc + a.csc_w[e + 3];
        any = any || b0 || b1 || b2 || b3;
    }
    for (; e < e1; ++e)
        if (pulled_active(a, lane, a.csc_src[e])) {
            acc = acc + a.csc_w[e];
            any = true;
        }
    out = acc;
    return any;
}

struct ReduceArgs {
    int64_t n;
    int ntiles;
    int cap;
    int sort_max;              // destinations with more terms pull
    int* bin_cnt;
    const int32_t* ev_r;
    const float* ev_w;
    const uint16_t* ev_d;
    float* psc;
    PullArgs pull;
};

// One workgroup per (tile, lane) bin.  Thread tid owns destinations
// d = k*512 + tid (k < 8) of the tile.
__global__ __launch_bounds__(kRedThreads) void k_reduce_bins(ReduceArgs a) {
    __shared__ int s_cnt[kTile];
    __shared__ int s_pos[kTile];
    __shared__ int s_r[kBinCap];
    __shared__ float s_w[kBinCap];
    __shared__ uint16_t s_d[kBinCap];
    __shared__ uint16_t s_ord[kBinCap];
    constexpr int kPer = kTile / kRedThreads;
    const int64_t lane = blockIdx.y;
    const int tile = blockIdx.x;
    const int64_t bin = lane * a.ntiles + tile;
    const int tid = threadIdx.x;
    const int c = a.bin_cnt[bin];
    if (c == 0) return;
    const int64_t dst0 = (int64_t)tile * kTile;
    const int ndst = (int)(a.n - dst0 < kTile ? a.n - dst0 : kTile);
    float* pl = a.psc + lane * a.n + dst0;
    for (int d = tid; d < kTile; d += kRedThreads) s_cnt[d] = 0;
    __syncthreads();
    if (tid == 0) a.bin_cnt[bin] = 0;  // every thread has read c
    if (c > a.cap) {
        // overflowed bin: every destination of the tile pulls
        for (int d = tid; d < ndst; d += kRedThreads) {
            float s;
            if (pull_sum(a.pull, lane, dst0 + d, s)) pl[d] = pl[d] + s;
        }
        return;
    }
    const int64_t q0 = bin * (int64_t)a.cap;
    for (int i = tid; i < c; i += kRedThreads) {
        const int d = a.ev_d[q0 + i] & (kTile - 1);
        s_d[i] = (uint16_t)d;
        s_r[i] = a.ev_r[q0 + i];
        s_w[i] = a.ev_w[q0 + i];
        atomicAdd(&s_cnt[d], 1);
    }
    __syncthreads();
    int v[kPer];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        v[k] = s_cnt[k * kRedThreads + tid];
        sum += v[k];
    }
    int total;
    int base = block_excl_scan<kRedThreads / 64>(sum, total);
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        s_pos[k * kRedThreads + tid] = base;
        base += v[k];
    }
    __syncthreads();
    for (int i = tid; i < c; i += kRedThreads) {
        const int p = atomicAdd(&s_pos[s_d[i]], 1);
        s_ord[p] = (uint16_t)i;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int d = k * kRedThreads + tid;
        const int cd = v[k];
        if (cd == 0 || d >= ndst) continue;
        float s = 0.f;
        if (cd <= a.sort_max) {
            const int end = s_pos[d], start = end - cd;
            // insertion sort of the destination's terms by virtual row
            for (int x = start + 1; x < end; ++x) {
                const int key = s_ord[x];
                const int rk = s_r[key];
                int y = x - 1;
                while (y >= start && s_r[s_ord[y]] > rk) {
                    s_ord[y + 1] = s_ord[y];
                    --y;
                }
                s_ord[y + 1] = (uint16_t)key;
            }
            for (int x = start; x < end; ++x) s = s + s_w[s_ord[x]];
        } else {
            pull_sum(a.pull, lane, dst0 + d, s);
        }
        pl[d] = pl[d] + s;
    }
}

// psc_{T-1} normalised (psc + 0.0f) and recorded.
__global__ __launch_bounds__(256) void k_finish_psc(float* __restrict__ psc, int64_t m,
                                                    float* __restrict__ rec) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const float p = psc[i] + 0.f;
    psc[i] = p;
    if (rec) rec[i] = p;
}

// Standalone exponential-synapse decay psc += k*((-psc)/tau_s) (ode.py:50).
__global__ __launch_bounds__(256) void k_exp_decay(float* __restrict__ psc, int64_t n,
                                                   int64_t m, const float* __restrict__ tsyn,
                                                   int64_t tsyn_s, const float* __restrict__ ksyn,
                                                   int64_t ksyn_s) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const int64_t nidx = i % n;
    const float p = psc[i];
    psc[i] = p + prm(ksyn, ksyn_s, nidx) * ((-p) / prm(tsyn, tsyn_s, nidx));
}

// Per-run tables, written from the launch arguments by one workgroup (a
// run-time index into an array of the argument block would make the compiler
// copy the whole block to scratch memory).
struct RunTables {
    int age_of_sub[BTX_MAX_BINS];
    int bin_delay[BTX_MAX_BINS];
    int bin_sub[BTX_MAX_BINS];
};

__global__ __launch_bounds__(64) void k_set_tables(RunTables t, Ctl* __restrict__ ctl) {
#pragma unroll
    for (int k = 0; k < BTX_MAX_BINS; ++k) {
        if ((int)threadIdx.x == k) {
            ctl->age_of_sub[k] = t.age_of_sub[k];
            ctl->bin_delay[k] = t.bin_delay[k];
            ctl->bin_sub[k] = t.bin_sub[k];
        }
    }
}

static inline int64_t mod(int64_t a, int64_t b) {
    const int64_t r = a % b;
    return r < 0 ? r + b : r;
}

// Engine workspace: control tables, then (ordered) bin counters and terms.
struct EngineWs {
    size_t ctl = 0, bin_cnt = 0, ev_r = 0, ev_w = 0, ev_d = 0, total = 0;
    int64_t ntiles = 0;
};

static EngineWs engine_ws(int64_t lanes, int64_t n, int ordered) {
    EngineWs L;
    L.ntiles = (n + kTile - 1) / kTile;
    size_t off = align256(sizeof(Ctl));
    if (ordered) {
        const int64_t bins = lanes * L.ntiles;
        L.bin_cnt = off;
        off = align256(off + sizeof(int) * bins);
        L.ev_r = off;
        off = align256(off + sizeof(int32_t) * bins * kBinCap);
        L.ev_w = off;
        off = align256(off + sizeof(float) * bins * kBinCap);
        L.ev_d = off;
        off = align256(off + sizeof(uint16_t) * bins * kBinCap);
    }
    L.total = off;
    return L;
}

}  // namespace btx

using namespace btx;

static StepArgs neuron_args(const btx_rsnn_plan* p) {
    StepArgs a{};
    a.n = p->n;
    a.nwords = (p->n + 63) / 64;
    a.nchunks = (p->n + kChunk - 1) / kChunk;
    a.dt = p->dt;
    a.hard = p->hard_reset;
    a.use_ref = p->use_refractory;
    a.tau = p->tau; a.tau_s = p->tau_stride;
    a.vth = p->v_th; a.vth_s = p->v_th_stride;
    a.vr = p->v_reset; a.vr_s = p->v_reset_stride;
    a.cm = p->c_m; a.cm_s = p->c_m_stride;
    a.tref = p->tau_ref; a.tref_s = p->tau_ref_stride;
    a.tsyn = p->tau_syn; a.tsyn_s = p->tau_syn_stride;
    a.ksyn = p->k_syn; a.ksyn_s = p->k_syn_stride;
    a.v = p->v;
    a.ref = p->refractory;
    return a;
}

static int check_neuron(const btx_rsnn_plan* p) {
    if (p->neuron_model != BTX_NEURON_LIF) return fail("unsupported neuron model");
    if (p->lanes > 65535) return fail("lanes exceeds 65535");
    if (p->use_refractory && (p->refractory == nullptr || p->tau_ref == nullptr))
        return fail("refractory state missing");
    return 0;
}

BTX_API int btx_lif_step(const btx_rsnn_plan* p, const float* x, float* z, float* v_pre,
                         void* stream) {
    if (p == nullptr) return fail("btx_lif_step: null plan");
    if (p->lanes <= 0 || p->n <= 0) return 0;
    if (int rc = check_neuron(p)) return rc;
    StepArgs a = neuron_args(p);
    a.first = 1;
    a.x = x;
    a.rec_z = z;
    a.rec_v_pre = v_pre;
    const dim3 grid((unsigned)a.nchunks, (unsigned)p->lanes);
    hipLaunchKernelGGL(k_lif_exp_step, grid, dim3(kChunk), 0, (hipStream_t)stream, a);
    return check_launch("btx_lif_step");
}

BTX_API int btx_exp_decay(float* psc, int64_t lanes, int64_t n, const float* tau_syn,
                          int64_t tau_syn_stride, const float* k_syn, int64_t k_syn_stride,
                          void* stream) {
    const int64_t m = lanes * n;
    if (m <= 0) return 0;
    hipLaunchKernelGGL(k_exp_decay, dim3((unsigned)((m + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, psc, n, m, tau_syn, tau_syn_stride, k_syn,
                       k_syn_stride);
    return check_launch("btx_exp_decay");
}

BTX_API size_t btx_rsnn_workspace_bytes(int64_t lanes, int64_t n, int slots, int ordered) {
    (void)slots;
    return engine_ws(lanes, n, ordered).total;
}

namespace btx {

// Optional per-kernel timing: an event pair around each launch.
struct KernelTimer {
    bool on = false;
    hipStream_t s = nullptr;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[4];
    void begin(int k) {
        if (!on) return;
        hipEvent_t a;
        hipEventCreate(&a);
        hipEventRecord(a, s);
        ev[k].push_back({a, nullptr});
    }
    void end(int k) {
        if (!on) return;
        hipEvent_t b;
        hipEventCreate(&b);
        hipEventRecord(b, s);
        ev[k].back().second = b;
    }
    void collect(double* ms, int64_t* calls, double* per, int64_t cap) {
        hipStreamSynchronize(s);
        for (int k = 0; k < 4; ++k) {
            double tot = 0;
            int64_t idx = 0;
            for (auto& pr : ev[k]) {
                float m = 0;
                hipEventElapsedTime(&m, pr.first, pr.second);
                tot += m;
                if (per != nullptr && idx < cap) per[k * cap + idx] = m;
                ++idx;
                hipEventDestroy(pr.first);
                hipEventDestroy(pr.second);
            }
            ms[k] = tot;
            calls[k] = (int64_t)ev[k].size();
            ev[k].clear();
        }
    }
};

static int rsnn_run(const btx_rsnn_plan* p, int64_t n_steps, hipStream_t s, KernelTimer& kt);

template <bool ORDERED>
static void launch_push_step(const PushStep& ps, int group, dim3 grid, hipStream_t s) {
    switch (group) {
        case 8: hipLaunchKernelGGL((k_push_step<8, ORDERED>), grid, dim3(kPushThreads), 0, s, ps); break;
        case 16: hipLaunchKernelGGL((k_push_step<16, ORDERED>), grid, dim3(kPushThreads), 0, s, ps); break;
        case 32: hipLaunchKernelGGL((k_push_step<32, ORDERED>), grid, dim3(kPushThreads), 0, s, ps); break;
        default: hipLaunchKernelGGL((k_push_step<64, ORDERED>), grid, dim3(kPushThreads), 0, s, ps); break;
    }
}

}  // namespace btx

BTX_API int btx_rsnn_run(const btx_rsnn_plan* p, int64_t n_steps, void* stream) {
    KernelTimer kt;
    return rsnn_run(p, n_steps, (hipStream_t)stream, kt);
}

BTX_API int btx_rsnn_run_timed(const btx_rsnn_plan* p, int64_t n_steps, void* stream,
                               double* ms_out, int64_t* calls_out, double* per_launch_ms,
                               int64_t per_launch_cap) {
    KernelTimer kt;
    kt.on = true;
    kt.s = (hipStream_t)stream;
    const int rc = rsnn_run(p, n_steps, (hipStream_t)stream, kt);
    kt.collect(ms_out, calls_out, per_launch_ms, per_launch_cap);
    return rc;
}

static int btx::rsnn_run(const btx_rsnn_plan* p, int64_t n_steps, hipStream_t s,
                         KernelTimer& kt) {
    if (p == nullptr) return fail("btx_rsnn_run: null plan");
    if (n_steps <= 0 || p->lanes <= 0 || p->n <= 0) return 0;
    if (int rc = check_neuron(p)) return rc;
    if (p->psc == nullptr || p->row_ptr == nullptr || p->raster == nullptr)
        return fail("btx_rsnn_run: psc, connectivity and raster are required");
    if (p->ring_idx == nullptr || p->ring_cnt == nullptr)
        return fail("btx_rsnn_run: the ring lists are required");
    if (p->n * (int64_t)p->csr_d >= ((int64_t)1 << 31))
        return fail("btx_rsnn_run: n*csr_d exceeds int32 row ids");
    if (p->nbins < 1 || p->nbins > BTX_MAX_BINS) return fail("btx_rsnn_run: bad nbins");
    if (p->ring_d < 1) return fail("btx_rsnn_run: ring_d must be >= 1");
    if (p->csr_d < 1 || p->csr_d > BTX_MAX_BINS)
        return fail("btx_rsnn_run: csr_d must be in [1, 64]");
    for (int j = 0; j < p->nbins; ++j) {
        if (p->bin_delay[j] < 0 || p->bin_delay[j] >= p->ring_d)
            return fail("btx_rsnn_run: bin delay outside the ring");
        if (p->bin_sub[j] < 0 || p->bin_sub[j] >= p->csr_d)
            return fail("btx_rsnn_run: bin offset outside csr_d");
    }
    const int64_t n = p->n, lanes = p->lanes, m = n * lanes;
    const int64_t nwords = (n + 63) / 64;
    const int64_t nchunks = (n + kChunk - 1) / kChunk;
    const EngineWs L = engine_ws(lanes, n, p->ordered);
    if (L.ntiles > kMaxTiles) return fail("btx_rsnn_run: n exceeds the tile table");
    if (p->ring_words == nullptr) return fail("btx_rsnn_run: the ring words are required");
    if (p->ordered && (p->csc_ptr == nullptr || p->csc_src == nullptr || p->csc_w == nullptr))
        return fail("btx_rsnn_run: ordered mode needs the dst-major arrays");
    if (p->ordered && (p->slots < 1 || p->slots > kSortMax))
        return fail("btx_rsnn_run: ordered mode needs 1 <= slots <= 32");
    if (p->bin_cap < 0 || p->bin_cap > kBinCap)
        return fail("btx_rsnn_run: bin_cap must be in [0, 6144]");
    if (p->in_prob != nullptr && (p->in_seed == nullptr || p->in_t0 == nullptr))
        return fail("btx_rsnn_run: Poisson drive needs per-lane seeds and clocks");
    if (p->lane_t_end != nullptr && p->in_t0 == nullptr)
        return fail("btx_rsnn_run: lane_t_end needs per-lane clocks");
    if (p->rec_sample != nullptr && (p->n_sample < 1 || p->n_sample > p->n))
        return fail("btx_rsnn_run: n_sample must be in [1, n]");
    if (p->workspace == nullptr || p->workspace_bytes < L.total)
        return fail("btx_rsnn_run: workspace too small");
    char* ws = (char*)p->workspace;
    Ctl* ctl = (Ctl*)ws;

    RunTables tb;
    for (int k = 0; k < BTX_MAX_BINS; ++k) {
        tb.age_of_sub[k] = -1;
        tb.bin_delay[k] = 0;
        tb.bin_sub[k] = 0;
    }
    for (int j = 0; j < p->nbins; ++j) {
        tb.age_of_sub[p->bin_sub[j]] = p->bin_delay[j];
        tb.bin_delay[j] = p->bin_delay[j];
        tb.bin_sub[j] = p->bin_sub[j];
    }
    hipLaunchKernelGGL(k_set_tables, dim3(1), dim3(64), 0, s, tb, ctl);
    if (p->ring_rebuild)
        hipLaunchKernelGGL(k_ring_from_words,
                           dim3((unsigned)nchunks, (unsigned)lanes, (unsigned)p->ring_d),
                           dim3(256), 0, s, p->ring_words, n, nwords, lanes, nchunks,
                           p->ring_idx, p->ring_cnt);

    StepArgs a = neuron_args(p);
    a.psc = p->psc;
    a.in_seed = p->in_seed;
    a.in_prob = p->in_prob;
    a.in_t0 = p->in_t0;
    a.in_kick = p->in_kick;
    a.counts = p->counts;
    a.n_sample = p->n_sample;
    a.lane_t_end = p->lane_t_end;

    PushStep ps{};
    ps.row_ptr = p->row_ptr;
    ps.col = p->col;
    ps.w = p->w;
    ps.n = n;
    ps.lanes = lanes;
    ps.nchunks = nchunks;
    ps.nbins = p->nbins;
    ps.csr_d = p->csr_d;
    ps.ring_d = p->ring_d;
    ps.ctl = ctl;
    ps.ring_idx = p->ring_idx;
    ps.ring_cnt = p->ring_cnt;
    ps.ntiles = (int)L.ntiles;
    ps.cap = p->bin_cap > 0 ? p->bin_cap : kBinCap;
    ps.bin_cnt = p->ordered ? (int*)(ws + L.bin_cnt) : nullptr;
    ps.ev_r = p->ordered ? (int32_t*)(ws + L.ev_r) : nullptr;
    ps.ev_w = p->ordered ? (float*)(ws + L.ev_w) : nullptr;
    ps.ev_d = p->ordered ? (uint16_t*)(ws + L.ev_d) : nullptr;
    ps.psc = p->psc;
    // segments per push workgroup: about 64 workgroups per lane, at most
    // kPushThreads segments (delay bins x segments) each
    int cpb = (int)(nchunks / 64);
    if (cpb < 1) cpb = 1;
    if (cpb * p->nbins > kPushThreads) cpb = kPushThreads / p->nbins;
    ps.cpb = cpb;
    const dim3 pgrid((unsigned)((nchunks + cpb - 1) / cpb), (unsigned)lanes);
    const int group = pick_group((double)p->nnz / (double)(n * p->csr_d));

    ReduceArgs ra{};
    ra.n = n;
    ra.ntiles = (int)L.ntiles;
    ra.cap = ps.cap;
    ra.sort_max = p->slots;
    ra.bin_cnt = ps.bin_cnt;
    ra.ev_r = ps.ev_r;
    ra.ev_w = ps.ev_w;
    ra.ev_d = ps.ev_d;
    ra.psc = p->psc;
    ra.pull.csc_ptr = p->csc_ptr;
    ra.pull.csc_src = p->csc_src;
    ra.pull.csc_w = p->csc_w;
    ra.pull.ring_words = p->ring_words;
    ra.pull.lanes = lanes;
    ra.pull.nwords = nwords;
    ra.pull.ring_d = p->ring_d;
    ra.pull.csr_d = p->csr_d;
    ra.pull.age_of_sub = ctl->age_of_sub;
    const dim3 rgrid((unsigned)L.ntiles, (unsigned)lanes);

    // the vector step needs n % 4 == 0 and 16-byte aligned state arrays
    auto al16 = [](const void* q) { return ((uintptr_t)q & 15) == 0; };
    const bool vec4 = (n % 4 == 0) && al16(p->v) && al16(p->psc) &&
                      (!p->use_refractory || al16(p->refractory)) && (!p->x || al16(p->x));
    const dim3 sgrid((unsigned)nchunks, (unsigned)lanes);
    const int64_t rec_step = m;
    for (int64_t t = 0; t < n_steps; ++t) {
        a.first = (t == 0);
        a.x = p->x ? p->x + t * p->x_step_stride : nullptr;
        a.t = t;
        a.rec_sample = p->rec_sample ? p->rec_sample + t * lanes * p->n_sample : nullptr;
        a.raster = p->raster + t * lanes * nwords;
        a.teacher = p->teacher ? p->teacher + t * lanes * nwords : nullptr;
        const int64_t slot = mod(p->ring_cursor + t, p->ring_d);
        a.ring_idx = p->ring_idx + slot * lanes * n;
        a.ring_cnt = p->ring_cnt + slot * lanes * nchunks;
        a.ring_words = p->ring_words + slot * lanes * nwords;
        a.rec_z = p->rec_z ? p->rec_z + t * rec_step : nullptr;
        a.rec_v = p->rec_v ? p->rec_v + t * rec_step : nullptr;
        a.rec_ref = p->rec_refractory ? p->rec_refractory + t * rec_step : nullptr;
        a.rec_v_pre = p->rec_v_pre ? p->rec_v_pre + t * rec_step : nullptr;
        a.rec_v_prev = p->rec_v_prev ? p->rec_v_prev + t * rec_step : nullptr;
        a.rec_psc_prev = (p->rec_psc && t > 0) ? p->rec_psc + (t - 1) * rec_step : nullptr;
        kt.begin(0);
        if (vec4) hipLaunchKernelGGL(k_lif_exp_step4, sgrid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL(k_lif_exp_step, sgrid, dim3(kChunk), 0, s, a);
        kt.end(0);
        ps.clock = p->ring_cursor + t;
        kt.begin(2);
        if (p->ordered) launch_push_step<true>(ps, group, pgrid, s);
        else launch_push_step<false>(ps, group, pgrid, s);
        kt.end(2);
        if (p->ordered) {
            ra.pull.clock = p->ring_cursor + t;
            kt.begin(1);
            hipLaunchKernelGGL(k_reduce_bins, rgrid, dim3(kRedThreads), 0, s, ra);
            kt.end(1);
        }
    }
    kt.begin(3);
    hipLaunchKernelGGL(k_finish_psc, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s,
                       p->psc, m, p->rec_psc ? p->rec_psc + (n_steps - 1) * rec_step : nullptr);
    kt.end(3);
    return check_launch("btx_rsnn_run");
}
