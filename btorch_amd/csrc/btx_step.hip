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
    int* cnt;
    const int2* slots;
    const int4* ovf;
    int ovf_cap;
    const Ctl* ctl;
    float* acc;
};

__device__ __forceinline__ float prm(const float* p, int64_t stride, int64_t n) {
    return p[stride * n];
}

// Synaptic input delivered by the previous step's push, leaving the
// accumulator zeroed for the next push.
__device__ __forceinline__ float take_input(const StepArgs& a, int64_t i) {
    if (a.ordered) {
        const int c = a.cnt[i];
        if (c == 0) return 0.f;
        const int novf = a.ctl->ovf_count < a.ovf_cap ? a.ctl->ovf_count : a.ovf_cap;
        const float r = ordered_sum<false>(i, c, a.slots, a.K, a.ovf, novf, nullptr);
        a.cnt[i] = 0;
        return r;
    }
    const float r = a.acc[i];
    if (r != 0.f) a.acc[i] = 0.f;
    return r;
}

__global__ __launch_bounds__(256) void k_lif_exp_step(StepArgs a) {
    const int64_t lane = blockIdx.y;
    const int64_t nidx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int spk = 0;
    if (nidx < a.n) {
        const int64_t i = lane * a.n + nidx;
        float psc = a.psc ? a.psc[i] : 0.f;
        if (!a.first) {
            psc = psc + take_input(a, i);
            if (a.rec_psc_prev) a.rec_psc_prev[i] = psc;
        }
        const float xin = a.x ? a.x[i] : 0.f;
        const float I = a.psc ? psc + xin : xin;
        const float tau = prm(a.tau, a.tau_s, nidx);
        const float vth = prm(a.vth, a.vth_s, nidx);
        const float vr = prm(a.vr, a.vr_s, nidx);
        const float cm = prm(a.cm, a.cm_s, nidx);
        const float v = a.v[i];
        if (a.rec_v_prev) a.rec_v_prev[i] = v;
        const float dv = (-(v - vr)) / tau + I / cm;
        const float v1 = v + a.dt * dv;
        if (a.rec_v_pre) a.rec_v_pre[i] = v1;
        const float u = (v1 - vth) / (vth - vr);
        float s = (u >= 0.f) ? 1.f : 0.f;
        if (a.use_ref) {
            const float ref = a.ref[i];
            s = s * ((ref == 0.f) ? 1.f : 0.f);
            float r2 = (ref + s * prm(a.tref, a.tref_s, nidx)) - a.dt;
            r2 = (r2 < 0.f) ? 0.f : r2;
            a.ref[i] = r2;
            if (a.rec_ref) a.rec_ref[i] = r2;
        }
        const float v2 = a.hard ? (v1 - (v1 - vr) * s) : (v1 - (vth - vr) * s);
        a.v[i] = v2;
        if (a.rec_v) a.rec_v[i] = v2;
        if (a.rec_z) a.rec_z[i] = s;
        if (a.psc) {
            const float tsyn = prm(a.tsyn, a.tsyn_s, nidx);
            const float k = prm(a.ksyn, a.ksyn_s, nidx);
            a.psc[i] = psc + k * ((-psc) / tsyn);
        }
        spk = s != 0.f;
    }
    const uint64_t m = __ballot(spk);
    if (a.raster != nullptr && (threadIdx.x & 63) == 0) {
        const int64_t word = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
        if (word < a.nwords) a.raster[lane * a.nwords + word] = m;
    }
}

// psc_{T-1} = psc_dec + input of the last push.
__global__ __launch_bounds__(256) void k_finish_psc(StepArgs a, int64_t m) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const float psc = a.psc[i] + take_input(a, i);
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
    void collect(double* ms, int64_t* calls) {
        hipStreamSynchronize(s);
        for (int k = 0; k < 4; ++k) {
            double tot = 0;
            for (auto& pr : ev[k]) {
                float m = 0;
                hipEventElapsedTime(&m, pr.first, pr.second);
                tot += m;
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
                               double* ms_out, int64_t* calls_out) {
    KernelTimer kt;
    kt.on = true;
    kt.s = (hipStream_t)stream;
    const int rc = rsnn_run(p, n_steps, (hipStream_t)stream, kt);
    kt.collect(ms_out, calls_out);
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
    if (p->ordered && p->slots < 2) return fail("btx_rsnn_run: ordered mode needs slots >= 2");
    const WsLayout L = ws_layout(p->lanes, p->n, p->slots, p->ordered);
    if (p->workspace_bytes < L.total) return fail("btx_rsnn_run: workspace too small");
    char* ws = (char*)p->workspace;

    const int64_t n = p->n, lanes = p->lanes, m = n * lanes;
    const int64_t nwords = (n + 63) / 64;
    StepArgs a = neuron_args(p);
    a.psc = p->psc;
    a.ctl = (const Ctl*)(ws + L.ctl);
    a.ovf_cap = (int)L.ovf_cap;
    if (p->ordered) {
        a.cnt = (int*)(ws + L.cnt);
        a.slots = (const int2*)(ws + L.slots);
        a.ovf = (const int4*)(ws + L.ovf);
    } else {
        a.acc = (float*)(ws + L.acc);
    }

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
    pa.ovf_cap = (int)L.ovf_cap;
    pa.ctl = (Ctl*)(ws + L.ctl);
    pa.cnt = a.cnt;
    pa.slots = (int2*)a.slots;
    pa.ovf = (int4*)a.ovf;
    pa.acc = a.acc;

    // Average row length over the bins read, for the push group width.
    const int group = pick_group((double)p->nnz / (double)(n * p->csr_d));
    const int blocks = push_blocks();
    const dim3 sgrid((unsigned)((n + 255) / 256), (unsigned)lanes);
    const int64_t rec_step = m;
    for (int64_t t = 0; t < n_steps; ++t) {
        a.first = (t == 0);
        a.x = p->x ? p->x + t * p->x_step_stride : nullptr;
        a.raster = p->raster + t * lanes * nwords;
        a.rec_z = p->rec_z ? p->rec_z + t * rec_step : nullptr;
        a.rec_v = p->rec_v ? p->rec_v + t * rec_step : nullptr;
        a.rec_ref = p->rec_refractory ? p->rec_refractory + t * rec_step : nullptr;
        a.rec_v_pre = p->rec_v_pre ? p->rec_v_pre + t * rec_step : nullptr;
        a.rec_v_prev = p->rec_v_prev ? p->rec_v_prev + t * rec_step : nullptr;
        a.rec_psc_prev = (p->rec_psc && t > 0) ? p->rec_psc + (t - 1) * rec_step : nullptr;
        hipEvent_t e0;
        kt.begin(0, e0);
        hipLaunchKernelGGL(k_lif_exp_step, sgrid, dim3(256), 0, s, a);
        kt.end(0);
        const int64_t slot = mod(p->ring_cursor + t, p->ring_d);
        const uint64_t* pushed = p->teacher ? p->teacher + t * lanes * nwords : a.raster;
        kt.begin(1, e0);
        hipLaunchKernelGGL(k_compact_bits, dim3((unsigned)lanes), dim3(1024), 0, s, pushed,
                           n, nwords, nwords, p->ring_idx + slot * lanes * n, n,
                           p->ring_cnt + slot * lanes, (int*)(ws + L.ctl));
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
    a.rec_psc_prev = p->rec_psc ? p->rec_psc + (n_steps - 1) * rec_step : nullptr;
    hipEvent_t e1;
    kt.begin(3, e1);
    hipLaunchKernelGGL(k_finish_psc, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, a, m);
    kt.end(3);
    return check_launch("btx_rsnn_run");
}

BTX_API int btx_rsnn_check_overflow(void* workspace, void* stream) {
    Ctl c{};
    if (hipMemcpyAsync(&c, workspace, sizeof(Ctl), hipMemcpyDeviceToHost, (hipStream_t)stream) !=
        hipSuccess)
        return fail("btx_rsnn_check_overflow: copy failed");
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess)
        return fail("btx_rsnn_check_overflow: sync failed");
    return c.ovf_error != 0;
}
