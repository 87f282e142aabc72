// Fused recurrent time step on gfx950: one elementwise pass over (lane, neuron)
// that finishes the previous step's synaptic input, integrates the neuron,
// fires, resets, decays the synapse and writes the spike bits; then the
// compaction of those bits into the delay ring, then the event-driven push
// of every active (virtual) row the ring exposes this step.
//
// Per step, in stream order:
//   F(t): psc_{t-1} = psc_dec + sum(terms of step t-1)   (skipped at t = 0)
//         I = psc_{t-1} + x_t ; LIF charge/fire/reset ; psc_dec = decay(psc_{t-1})
//         spike bits -> raster[t]
//   C(t): raster[t] -> ascending active list in ring slot (cursor + t) % ring_d
//   S(t): push rows n*csr_d + bin_sub[j] for n in ring slot (cursor + t - bin_delay[j])
// and after the last step FIN: psc_{T-1} = psc_dec + sum(terms of step T-1).
//
// Op order follows the reference's eager expressions exactly (compiled with
// -ffp-contract=off):
//   LIF charge  v + dt*((-(v - v_reset))/tau + I/c_m)   lif.py:144-154, ode.py:53-55
//   fire        ((v - v_th)/(v_th - v_reset)) >= 0       lif.py:160-168, surrogate/base.py:6-8
//   refractory  s * (ref == 0) ; relu((ref + s*tau_ref) - dt)   lif.py:165-167, 184-187
//   reset soft  v - (v_th - v_reset)*s ; hard v - (v - v_reset)*s   lif.py:176-180
//   exp-Euler   psc + k*((-psc)/tau_s), k = expm1(dt*l)/l, l = -1/tau_s  ode.py:50, synapse.py:171-178
//   adaptation  psc + W^T z                                  synapse.py:82-87

#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "btx_common.h"
#include "btx_spmv.h"

namespace btx {

struct StepArgs {
    int64_t n;
    int64_t nwords;
    float dt;
    int hard;
    int use_ref;
    int ordered;
    int K;
    int first;
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
    float* rec_z;          // step t or NULL
    float* rec_v;
    float* rec_ref;
    float* rec_psc_prev;   // step t-1 or NULL
    float* rec_v_pre;
    float* rec_v_prev;
    int lanes;
    int* cnt;
    const int2* slots;     // [K][lanes*n]
    float* acc;
    // pull of a destination that received more than K terms
    int64_t push_step;     // step whose push F consumes (t - 1)
    const int64_t* csc_ptr;
    const int32_t* csc_src;
    const float* csc_w;
    const uint64_t* words;       // pushed spike words of this run [T][lanes][nwords]
    const uint64_t* hist_words;  // words of the ring_d steps before the run [ring_d][lanes][nwords]
    int64_t hist_steps;          // > 0: hist_words is the raster of that many preceding steps
    // in-kernel Poisson drive
    const uint64_t* in_seed;
    const float* in_prob;
    const int64_t* in_t0;
    float in_kick;
    int64_t t;                   // step of this launch within the run
    int* counts;
    float* rec_sample;           // step t, [lanes][n_sample]
    int64_t n_sample;
    const int64_t* lane_t_end;
    int ring_d;
    int cursor;
    int csr_d;
    const int* age_of_sub;  // [csr_d] ring age of virtual-row offset, -1 = not read
};

// Per-run table of ages, written from the launch arguments by one block.
struct AgeTable {
    int v[BTX_MAX_BINS];
};

__global__ __launch_bounds__(64) void k_set_ages(AgeTable t, int* __restrict__ out) {
#pragma unroll
    for (int k = 0; k < BTX_MAX_BINS; ++k)
        if ((int)threadIdx.x == k) out[k] = t.v[k];
}

__device__ __forceinline__ float prm(const float* p, int64_t stride, int64_t n) {
    return p[stride * n];
}

// Exact input of destination `dst` of `lane` from the push of step
// a.push_step, pulled over its incoming edges in the coalesced (dst, src)
// order: a virtual row src*csr_d + sub is active when neuron src spiked
// age_of_sub[sub] steps before that push.
__device__ __forceinline__ bool pulled_active(const StepArgs& a, int64_t lane, int r) {
    const int src = r / a.csr_d;
    const int age = a.age_of_sub[r - src * a.csr_d];
    if (age < 0) return false;
    const int64_t st = a.push_step - age;
    const uint64_t* row;
    if (st >= 0) {
        row = a.words + (st * a.lanes + lane) * a.nwords;
    } else if (a.hist_steps > 0) {
        row = a.hist_words + ((a.hist_steps + st) * a.lanes + lane) * a.nwords;
    } else {
        int64_t slot = (a.cursor + st) % a.ring_d;
        if (slot < 0) slot += a.ring_d;
        row = a.hist_words + (slot * a.lanes + lane) * a.nwords;
    }
    return (row[src >> 6] >> (src & 63)) & 1ull;
}

__device__ float pull_input(const StepArgs& a, int64_t lane, int64_t dst) {
    float acc = 0.f;
    int64_t e = a.csc_ptr[dst];
    const int64_t e1 = a.csc_ptr[dst + 1];
    // four independent activity tests in flight, summed in order
    for (; e + 4 <= e1; e += 4) {
        const int r0 = a.csc_src[e], r1 = a.csc_src[e + 1];
        const int r2 = a.csc_src[e + 2], r3 = a.csc_src[e + 3];
        const bool b0 = pulled_active(a, lane, r0), b1 = pulled_active(a, lane, r1);
        const bool b2 = pulled_active(a, lane, r2), b3 = pulled_active(a, lane, r3);
        if (b0) acc = acc + a.csc_w[e];
        if (b1) acc = acc + a.csc_w[e + 1];
        if (b2) acc = acc + a.csc_w[e + 2];
        if (b3) acc = acc + a.csc_w[e + 3];
    }
    for (; e < e1; ++e)
        if (pulled_active(a, lane, a.csc_src[e])) acc = acc + a.csc_w[e];
    return acc;
}

// Synaptic input delivered by the previous step's push (c terms in the
// ordered mode), leaving the accumulator zeroed for the next push.
__device__ __forceinline__ float take_input_c(const StepArgs& a, int64_t i, int64_t lane,
                                              int64_t nidx, int c) {
    if (c == 0) return 0.f;
    const float r = c <= a.K ? ordered_sum<false>(i, c, a.slots, a.lanes * a.n, nullptr)
                             : pull_input(a, lane, nidx);
    a.cnt[i] = 0;
    return r;
}

__device__ __forceinline__ float take_input(const StepArgs& a, int64_t i, int64_t lane,
                                            int64_t nidx) {
    if (a.ordered) return take_input_c(a, i, lane, nidx, a.cnt[i]);
    const float r = a.acc[i];
    if (r != 0.f) a.acc[i] = 0.f;
    return r;
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

// One neuron of the fused step.  `psc` enters as the decayed current of the
// previous step (without its push), leaves as this step's decayed current.
// Returns the spike (0/1).
__device__ __forceinline__ float lif_exp_neuron(const StepArgs& a, int64_t lane, int64_t nidx,
                                                int64_t i, float& psc, float xin, float& v,
                                                float& ref, float wz) {
    // xin: external input (tensor) of this neuron-lane, before the Poisson drive
    if (!a.first) {
        psc = psc + wz;
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

// Scalar form: one neuron per thread, spike word by wavefront ballot.
__global__ __launch_bounds__(256) void k_lif_exp_step(StepArgs a) {
    const int64_t lane = blockIdx.y;
    const int64_t nidx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int spk = 0;
    if (nidx < a.n) {
        const int64_t i = lane * a.n + nidx;
        float psc = a.psc ? a.psc[i] : 0.f;
        const float wz = a.first ? 0.f : take_input(a, i, lane, nidx);
        const float xin = a.x ? a.x[i] : 0.f;
        float v = a.v[i];
        float ref = a.use_ref ? a.ref[i] : 0.f;
        const float s = lif_exp_neuron(a, lane, nidx, i, psc, xin, v, ref, wz);
        a.v[i] = v;
        if (a.use_ref) a.ref[i] = ref;
        if (a.psc) a.psc[i] = psc;
        spk = s != 0.f;
    }
    const uint64_t m = __ballot(spk);
    if (a.raster != nullptr && (threadIdx.x & 63) == 0) {
        const int64_t word = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
        if (word < a.nwords) a.raster[lane * a.nwords + word] = m;
    }
}

// Vector form (n % 4 == 0, 16-byte aligned state): four consecutive neurons
// per thread with 16-byte loads/stores; the 64-bit spike word of 16 threads
// is assembled by an OR across them.
__global__ __launch_bounds__(256) void k_lif_exp_step4(StepArgs a) {
    const int64_t lane = blockIdx.y;
    const int64_t n0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 4;
    uint64_t nib = 0;
    if (n0 < a.n) {
        const int64_t i0 = lane * a.n + n0;
        float4 v4 = *(const float4*)(a.v + i0);
        float4 p4 = a.psc ? *(const float4*)(a.psc + i0) : make_float4(0.f, 0.f, 0.f, 0.f);
        float4 r4 = a.use_ref ? *(const float4*)(a.ref + i0) : make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 x4 = a.x ? *(const float4*)(a.x + i0) : make_float4(0.f, 0.f, 0.f, 0.f);
        float w[4] = {0.f, 0.f, 0.f, 0.f};
        if (!a.first) {
            if (a.ordered) {
                const int4 c4 = *(const int4*)(a.cnt + i0);
                w[0] = take_input_c(a, i0, lane, n0, c4.x);
                w[1] = take_input_c(a, i0 + 1, lane, n0 + 1, c4.y);
                w[2] = take_input_c(a, i0 + 2, lane, n0 + 2, c4.z);
                w[3] = take_input_c(a, i0 + 3, lane, n0 + 3, c4.w);
            } else {
                const float4 q = *(const float4*)(a.acc + i0);
                w[0] = q.x; w[1] = q.y; w[2] = q.z; w[3] = q.w;
                if (q.x != 0.f || q.y != 0.f || q.z != 0.f || q.w != 0.f)
                    *(float4*)(a.acc + i0) = make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
        float s0 = lif_exp_neuron(a, lane, n0, i0, p4.x, x4.x, v4.x, r4.x, w[0]);
        float s1 = lif_exp_neuron(a, lane, n0 + 1, i0 + 1, p4.y, x4.y, v4.y, r4.y, w[1]);
        float s2 = lif_exp_neuron(a, lane, n0 + 2, i0 + 2, p4.z, x4.z, v4.z, r4.z, w[2]);
        float s3 = lif_exp_neuron(a, lane, n0 + 3, i0 + 3, p4.w, x4.w, v4.w, r4.w, w[3]);
        *(float4*)(a.v + i0) = v4;
        if (a.use_ref) *(float4*)(a.ref + i0) = r4;
        if (a.psc) *(float4*)(a.psc + i0) = p4;
        nib = (uint64_t)(s0 != 0.f) | ((uint64_t)(s1 != 0.f) << 1) |
              ((uint64_t)(s2 != 0.f) << 2) | ((uint64_t)(s3 != 0.f) << 3);
    }
    uint64_t word = nib << (4 * (threadIdx.x & 15));
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) word |= __shfl_xor(word, o, 64);
    if (a.raster != nullptr && (threadIdx.x & 15) == 0) {
        const int64_t wi = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
        if (wi < a.nwords) a.raster[lane * a.nwords + wi] = word;
    }
}

// psc_{T-1} = psc_dec + input of the last push.
__global__ __launch_bounds__(256) void k_finish_psc(StepArgs a, int64_t m) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const float psc = a.psc[i] + take_input(a, i, i / a.n, i % a.n);
    a.psc[i] = psc;
    if (a.rec_psc_prev) a.rec_psc_prev[i] = psc;
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

static inline int64_t mod(int64_t a, int64_t b) {
    const int64_t r = a % b;
    return r < 0 ? r + b : r;
}

}  // namespace btx

using namespace btx;

static StepArgs neuron_args(const btx_rsnn_plan* p) {
    StepArgs a{};
    a.n = p->n;
    a.nwords = (p->n + 63) / 64;
    a.dt = p->dt;
    a.hard = p->hard_reset;
    a.use_ref = p->use_refractory;
    a.ordered = p->ordered;
    a.K = p->slots;
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
    const dim3 grid((unsigned)((p->n + 255) / 256), (unsigned)p->lanes);
    hipLaunchKernelGGL(k_lif_exp_step, grid, dim3(256), 0, (hipStream_t)stream, a);
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
    return ws_layout(lanes, n, slots, ordered).total;
}

namespace btx {

// Optional per-kernel timing: an event pair around each launch.
struct KernelTimer {
    bool on = false;
    hipStream_t s = nullptr;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[4];
    void begin(int k, hipEvent_t& a) {
        if (!on) return;
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
    if (p->n * (int64_t)p->csr_d >= ((int64_t)1 << 31))
        return fail("btx_rsnn_run: n*csr_d exceeds int32 row ids");
    if (p->nbins < 1 || p->nbins > BTX_MAX_BINS) return fail("btx_rsnn_run: bad nbins");
    if (p->ring_d < 1) return fail("btx_rsnn_run: ring_d must be >= 1");
    for (int j = 0; j < p->nbins; ++j) {
        if (p->bin_delay[j] < 0 || p->bin_delay[j] >= p->ring_d)
            return fail("btx_rsnn_run: bin delay outside the ring");
        if (p->bin_sub[j] < 0 || p->bin_sub[j] >= p->csr_d)
            return fail("btx_rsnn_run: bin offset outside csr_d");
    }
    if (p->ordered && (p->slots < 2 || p->slots > kMaxSlots))
        return fail("btx_rsnn_run: ordered mode needs 2 <= slots <= 16");
    if (p->ordered && (p->csc_ptr == nullptr || p->csc_src == nullptr || p->csc_w == nullptr ||
                       p->hist_words == nullptr))
        return fail("btx_rsnn_run: ordered mode needs the dst-major arrays and history words");
    if (p->hist_steps > 0 && p->hist_steps < p->ring_d)
        return fail("btx_rsnn_run: hist_steps must cover the ring");
    if (p->in_prob != nullptr && (p->in_seed == nullptr || p->in_t0 == nullptr))
        return fail("btx_rsnn_run: Poisson drive needs per-lane seeds and clocks");
    if (p->lane_t_end != nullptr && p->in_t0 == nullptr)
        return fail("btx_rsnn_run: lane_t_end needs per-lane clocks");
    if (p->rec_sample != nullptr && (p->n_sample < 1 || p->n_sample > p->n))
        return fail("btx_rsnn_run: n_sample must be in [1, n]");
    if (p->csr_d > BTX_MAX_BINS) return fail("btx_rsnn_run: csr_d exceeds 64");
    const WsLayout L = ws_layout(p->lanes, p->n, p->slots, p->ordered);
    if (p->workspace_bytes < L.total) return fail("btx_rsnn_run: workspace too small");
    char* ws = (char*)p->workspace;

    const int64_t n = p->n, lanes = p->lanes, m = n * lanes;
    const int64_t nwords = (n + 63) / 64;
    StepArgs a = neuron_args(p);
    a.psc = p->psc;
    a.lanes = (int)lanes;
    if (p->ordered) {
        a.cnt = (int*)(ws + L.cnt);
        a.slots = (const int2*)(ws + L.slots);
    } else {
        a.acc = (float*)(ws + L.acc);
    }
    a.csc_ptr = p->csc_ptr;
    a.csc_src = p->csc_src;
    a.csc_w = p->csc_w;
    a.words = p->teacher ? p->teacher : p->raster;
    a.hist_words = p->hist_words;
    a.hist_steps = p->hist_steps;
    a.in_seed = p->in_seed;
    a.in_prob = p->in_prob;
    a.in_t0 = p->in_t0;
    a.in_kick = p->in_kick;
    a.counts = p->counts;
    a.n_sample = p->n_sample;
    a.lane_t_end = p->lane_t_end;
    a.ring_d = p->ring_d;
    a.cursor = p->ring_cursor;
    a.csr_d = p->csr_d;
    AgeTable ages;
    for (int k = 0; k < BTX_MAX_BINS; ++k) ages.v[k] = -1;
    for (int j = 0; j < p->nbins; ++j) ages.v[p->bin_sub[j]] = p->bin_delay[j];
    int* age_dev = ((Ctl*)(ws + L.ctl))->age_of_sub;
    hipLaunchKernelGGL(k_set_ages, dim3(1), dim3(64), 0, s, ages, age_dev);
    a.age_of_sub = age_dev;

    // Per step the push reads, for every delay bin, the compacted spikes of
    // the ring slot that bin's age points at, and streams the bin's rows.
    PushArgs pa{};
    pa.row_ptr = p->row_ptr;
    pa.col = p->col;
    pa.w = p->w;
    pa.n_dst = n;
    pa.ring_idx = p->ring_idx;
    pa.ring_cnt = p->ring_cnt;
    pa.ring_cap = n;
    pa.ring_lanes = lanes;
    pa.lane_base = 0;
    pa.lanes = (int)lanes;
    pa.nbins = p->nbins;
    pa.csr_d = p->csr_d;
    pa.x = nullptr;
    pa.ordered = p->ordered;
    pa.K = p->slots;
    pa.ctl = (Ctl*)(ws + L.ctl);
    pa.cnt = a.cnt;
    pa.slots = (int2*)a.slots;
    pa.acc = a.acc;

    // push group width from the mean row length of the bins read
    const int group = pick_group((double)p->nnz / (double)(n * p->csr_d));
    const int blocks = push_blocks();
    // the vector step needs n % 4 == 0 and 16-byte aligned state arrays
    auto al16 = [](const void* q) { return ((uintptr_t)q & 15) == 0; };
    const bool vec4 = (n % 4 == 0) && al16(p->v) && al16(p->psc) &&
                      (!p->use_refractory || al16(p->refractory)) && (!p->x || al16(p->x)) &&
                      al16(ws) && (!p->rec_psc || al16(p->rec_psc));
    const dim3 sgrid(vec4 ? (unsigned)((n + 1023) / 1024) : (unsigned)((n + 255) / 256),
                     (unsigned)lanes);
    hipEvent_t e0;
    const int64_t rec_step = m;
    for (int64_t t = 0; t < n_steps; ++t) {
        a.first = (t == 0);
        a.push_step = t - 1;
        a.x = p->x ? p->x + t * p->x_step_stride : nullptr;
        a.t = t;
        a.rec_sample = p->rec_sample ? p->rec_sample + t * lanes * p->n_sample : nullptr;
        a.raster = p->raster + t * lanes * nwords;
        a.rec_z = p->rec_z ? p->rec_z + t * rec_step : nullptr;
        a.rec_v = p->rec_v ? p->rec_v + t * rec_step : nullptr;
        a.rec_ref = p->rec_refractory ? p->rec_refractory + t * rec_step : nullptr;
        a.rec_v_pre = p->rec_v_pre ? p->rec_v_pre + t * rec_step : nullptr;
        a.rec_v_prev = p->rec_v_prev ? p->rec_v_prev + t * rec_step : nullptr;
        a.rec_psc_prev = (p->rec_psc && t > 0) ? p->rec_psc + (t - 1) * rec_step : nullptr;
        kt.begin(0, e0);
        if (vec4) hipLaunchKernelGGL(k_lif_exp_step4, sgrid, dim3(256), 0, s, a);
        else hipLaunchKernelGGL(k_lif_exp_step, sgrid, dim3(256), 0, s, a);
        kt.end(0);
        const int64_t slot = mod(p->ring_cursor + t, p->ring_d);
        const uint64_t* pushed = p->teacher ? p->teacher + t * lanes * nwords : a.raster;
        kt.begin(1, e0);
        hipLaunchKernelGGL(k_compact_bits, dim3((unsigned)lanes), dim3(1024), 0, s, pushed,
                           n, nwords, nwords, p->ring_idx + slot * lanes * n, n,
                           p->ring_cnt + slot * lanes);
        kt.end(1);
        for (int j = 0; j < p->nbins; ++j) {
            pa.bin_slot[j] = (int)mod(p->ring_cursor + t - p->bin_delay[j], p->ring_d);
            pa.bin_sub[j] = p->bin_sub[j];
        }
        kt.begin(2, e0);
        int rc = launch_push(pa, group, blocks, s);
        kt.end(2);
        if (rc) return rc;
    }
    a.first = 0;
    a.push_step = n_steps - 1;
    a.rec_psc_prev = p->rec_psc ? p->rec_psc + (n_steps - 1) * rec_step : nullptr;
    kt.begin(3, e0);
    hipLaunchKernelGGL(k_finish_psc, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, a, m);
    kt.end(3);
    return check_launch("btx_rsnn_run");
}

