// Fused recurrent time step on gfx950.
//
// Every step is one launch whose workgroups play two roles:
//   F (lane, 1024-neuron segment c): first reduces the previous step's push
//     into its own neurons -- the terms binned for (lane, c), sorted in LDS by
//     neuron and, per neuron, by virtual source row, summed from +0.0f in that
//     order -- then integrates the step: psc_{t-1} = psc_dec + sum, LIF
//     charge/fire/refractory/reset, exponential-synapse decay psc_dec, spike
//     words -> raster, and the pushed spikes (own or teacher's) as the
//     segment's ascending list and words into ring slot (cursor + t) % R.
//   P (lane, range of segments): the event-driven push of step t -- for every
//     delay bin j, the neurons the ring lists for step t - bin_delay[j], each
//     streaming its CSR virtual row src*csr_d + bin_sub[j]; every term
//     (virtual row, weight) is binned by (lane, destination segment) with one
//     global reservation per workgroup and segment (ordered mode), or added
//     to a float accumulator with an atomic (fast mode).
// The push of step t only reads ring slots of earlier steps when no delay
// bin has age 0, so both roles share the launch; otherwise the F and P
// workgroups of a step are two launches.  Term bins and accumulators are
// double-buffered by step parity.  The ring has R = ring_d + 1 slots so that
// the slot a step writes is never one the concurrent reduction still reads.
// After the last step FIN reduces the last push into psc.
//
// A neuron with more than `slots` terms, or every neuron of a bin that
// overflowed, takes its input by pulling its incoming edges in the
// destination-major (reference coalesced) order instead: the same terms in
// the same order, so the same bits.
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
// GLIF3 (glif.py:264-342) and the two-compartment neuron
// (two_compartment.py:208-292) replace the LIF update with their own, in
// their reference op order; their transcendental coefficients come from the
// host.

#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "btx_common.h"
#include "btx_spmv.h"

namespace btx {

constexpr int kChunk = BTX_RING_CHUNK;  // neurons per F workgroup = ring segment = bin
constexpr int kChunkShift = 10;
constexpr int kThreads = 256;           // 4 neurons per thread in the F role
constexpr int kBinCap = 1024;           // terms per bin sorted in LDS
constexpr int kBinCapMax = 8192;        // terms per bin in the workspace (larger bins are
                                        // sorted through a global order array)
constexpr int kSortMax = 32;            // most terms a neuron sorts in LDS
constexpr int kMaxChunks = 1024;        // n <= 1,048,576 (push reservation table)
constexpr int kStage = 1024;            // terms a push workgroup stages in LDS
constexpr int kPullWgs = 512;           // workgroups a dense pull spreads over, at least

static_assert(kChunk == (1 << kChunkShift) && kChunk == 4 * kThreads, "segment layout");

// ---------------------------------------------------------------------------
// Arguments
// ---------------------------------------------------------------------------

struct PullArgs {
    const int64_t* csc_ptr;
    const int32_t* csc_src;
    const float* csc_w;
    const uint64_t* ring_words;  // [R][lanes][nwords], ring-slot order
    int64_t lanes;
    int64_t nwords;
    int R;
    int64_t clock;               // cursor + step of the push being reduced
    int slot0;                   // clock mod R
    int csr_d;
    const int* age_of_sub;       // [csr_d] ring age by virtual-row offset, -1 = not read
};

// Reduction of one step's push into the neurons of one segment.
struct ReduceIn {
    int ordered;
    int cap;
    int sort_max;
    int* bin_cnt;                // [lanes][nchunks] of the buffer being consumed
    const int32_t* ev_r;         // [lanes][nchunks][kBinCapMax]
    const float* ev_w;
    const uint16_t* ev_d;
    uint16_t* ev_o;              // order of a bin's terms by neuron (bins past kBinCap)
    float* acc;                  // fast mode: [lanes][n] of the buffer being consumed
    const int* pull_mode;        // ordered mode: 1 = the push was a dense pull (Ctl::push_mode) ...
    const float* pull_acc;       // ... whose sums are here, [lanes][n]
    PullArgs pull;
};

constexpr int kMaxAsc = 4;

// Neuron kinds of the F role.
constexpr int kLif = 0;
constexpr int kGlif3 = 1;
constexpr int kTwoComp = 2;
// Several neuron models on one axis: each neuron's model (and reset mode)
// comes from StepArgs::kinds; the state arrays cover every neuron.
constexpr int kMixed = 3;
// Flag on a kind: the synapse has kinetics other than the exponential one
// (a second state h, or two conductances), chosen by StepArgs::syn.
constexpr int kSynX = 4;

struct StepArgs {
    int64_t n;
    int64_t nwords;
    int64_t nchunks;
    float dt;
    int hard;
    int use_ref;
    int kind;              // kLif, kGlif3, kTwoComp or kMixed
    const uint8_t* kinds;  // kMixed: per-neuron model code (BTX_MIXED_*)
    int n_asc;
    const float* vrest;    int64_t vrest_s;
    const float* kv;       int64_t kv_s;
    const float* asc_k;    // (n, n_asc)
    const float* asc_kk;
    const float* asc_amp;
    float* iasc;           // (lanes, n, n_asc)
    int prev;                    // 1 = the previous step's push is pending
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
    uint64_t* raster;      // step t, (lanes, nwords) or NULL
    const uint64_t* teacher;  // step t, (lanes, nwords) or NULL
    int32_t* ring_idx;     // this step's ring slot, [lanes][n], or NULL
    int32_t* ring_cnt;     // this step's ring slot, [lanes][nchunks]
    uint64_t* ring_words;  // this step's ring slot, [lanes][nwords]
    uint64_t* lane_bits;   // ordered mode: this step's slot of the lane-bit ring, [n][lw],
    int lw;                // marked when *lane_bits_on (set one launch ahead)
    const int* lane_bits_on;
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
    int vec4;              // 16-byte aligned state and n % 4 == 0
    int rec_any;           // 1 = some record, count or sample pointer is set
    int shared;            // 1 = v_th, v_reset, c_m, tau_ref, tau_syn, k_syn all stride 0
    // two-compartment neuron
    const float* xa;       // step t apical input (lanes, n) or NULL
    float* tia;            // state i_a, i_bap, theta_th (lanes, n)
    float* tibap;
    float* tth;
    const float* tcp[BTX_TC_NPARAM];
    int64_t tcs[BTX_TC_NPARAM];
    float* rec_ia;         // step t or NULL
    float* rec_ibap;
    float* rec_th;
    // synapse kinetics other than exponential (BTX_SYN_*)
    int syn;
    float* sh1;            // h (alpha) or g_rise (dual exponential), (lanes, n)
    float* sh2;            // g_decay (dual exponential)
    const float* sg;       int64_t sg_s;    // gain on the summed push
    const float* tsyn2;    int64_t tsyn2_s;
    const float* ksyn2;    int64_t ksyn2_s;
    const float* sa;       int64_t sa_s;    // psc = sa*(g_decay - g_rise)
    float* rec_sh1_prev;   // step t-1 or NULL
    float* rec_sh2_prev;
    ReduceIn red;          // the previous step's push
};

struct PushStep {
    const int64_t* row_ptr;
    const int32_t* col;
    const float* w;
    int64_t n;
    int64_t lanes;
    int64_t nchunks;
    int cpb;                     // segments per P workgroup
    int nbins;
    int csr_d;
    int R;
    int64_t clock;               // cursor + t
    const Ctl* ctl;              // bin_delay / bin_sub tables
    const int32_t* ring_idx;     // [R][lanes][n]
    const int32_t* ring_cnt;     // [R][lanes][nchunks]
    int cap;
    int* bin_cnt;                // [lanes][nchunks] of this step's buffer
    int32_t* ev_r;               // [lanes][nchunks][kBinCapMax]
    float* ev_w;
    uint16_t* ev_d;
    float* acc;                  // fast mode: [lanes][n] of this step's buffer
    const int64_t* lane_t_end;   // lanes past their trial end push nothing
    const int64_t* in_t0;
    int64_t t;
    int affine;                  // bin j is (delay0 + j, sub0 + j)
    // ordered mode, dense pull (chosen per step when the push would bin more
    // than a threshold of terms per (lane, segment)): sums of every destination's
    // active incoming edges in the coalesced order into pull_acc
    uint64_t* lane_bits;         // [R][n][lw]: bit b of word (slot, j, b/64) = lane b pushed j
    int lw;
    float* pull_acc;             // [lanes][n] of this step's buffer
    const int64_t* csc_ptr;
    const int32_t* csc_src;
    const float* csc_w;
    int delay0;
    int sub0;
    int nh;                      // pull-only workgroups per lane row (see k_step)
    int pull_sep;                // the dense pull runs in k_pull, not in the P role
};

// Arguments of the control workgroup, a block of their own: were they read
// from the push's block, the compiler would hold the shared values for the
// whole kernel, reloading them from register lanes thousands of times on the
// step's other paths.
struct ControlArgs {
    const int32_t* ring_cnt;     // [R][lanes][nchunks]
    const Ctl* ctl;              // bin delays
    Ctl* ctl_rw;                 // written: slot totals, next push mode, lane-bit marking
    int64_t m;                   // lanes * nchunks
    int64_t clock;               // cursor + t
    int R;
    int nbins;
    int affine;                  // bin j has delay delay0 + j
    int delay0;
    float terms_per_row;         // nnz / (n * csr_d)
    float above;                 // terms per bin above which a step pulls, x lanes x nchunks
};

// LDS of one workgroup, shared by the two roles.
struct RedLds {
    int cnt[kChunk];             // terms per neuron, then (same words) their positions
    int r[kBinCap];
    float w[kBinCap];
    uint16_t d[kBinCap];
    uint16_t ord[kBinCap];
};
struct PushLds {
    int off[kThreads + 1];
    int slot[BTX_MAX_BINS];
    int sub[BTX_MAX_BINS];
    int tile[kMaxChunks];
    uint16_t base[kMaxChunks];   // reserved position, clamped to the bin capacity
    int nev;                     // terms met in the first pass
    int er[kStage];              // staged terms: virtual row, weight, destination
    float ew[kStage];
    int ed[kStage];
};
union StepLds {
    RedLds red;
    PushLds push;
};

// ---------------------------------------------------------------------------
// Neuron update
// ---------------------------------------------------------------------------

// A value that no workgroup of the launch writes before reading it (launch
// arguments, parameters, the previous step's bin count), read through the
// constant address space: a uniform address then becomes a scalar load,
// whose wait does not wait for the vector loads issued around it.
template <typename T>
__device__ __forceinline__ T ld_const(const T* p) {
    return *(const __attribute__((address_space(4))) T*)p;
}

// Parameter of neuron n: per-neuron array (stride 1) or one shared value
// (stride 0); the stride is uniform, so this is a select, not a multiply.
__device__ __forceinline__ float prm(const float* p, int64_t stride, int64_t n) {
    return stride ? p[n] : p[0];
}

// Neuron parameters requested at the start of the F role, so that their
// loads overlap the reduction: the shared value of each (used when its
// stride is 0).
struct PreParams {
    float vth, vr, cm, tref, tsyn, ksyn;
};

__device__ __forceinline__ float prm_or(const float* p, int64_t stride, int64_t n, float shared) {
    return stride ? p[n] : shared;
}

// SHR: the launch's parameters other than tau are all shared (stride 0), so
// the values requested at the start are used and no stride is consulted.
template <bool SHR>
__device__ __forceinline__ float prm_sh(const float* p, int64_t stride, int64_t n, float shared) {
    return SHR ? shared : prm_or(p, stride, n, shared);
}

// (x >= 0) of the quotient u = a / d (d = v_th - v_reset) as the
// reference's surrogate Heaviside sees it (surrogate/base.py:6-8).  For
// finite d > 0 the sign of u is the sign of a, except that a tiny negative a
// can round to -0.0, which also counts; only then is the division made.
__device__ __forceinline__ bool fires(float a, float d) {
    if (d > 0.f && d <= 3.4028234e38f) {
        if (a >= 0.f) return true;
        if (!(a > -d * 1e-43f)) return false;  // |a/d| stays above the underflow range
    }
    return (a / d) >= 0.f;
}

// 16-byte state accesses marked non-temporal: the state of a step is touched
// once per step, so it streams through the caches instead of leaving dirty
// lines behind for the end-of-kernel cache write-back.
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void load4(const float* p, float* o) {
    const f4v q = __builtin_nontemporal_load((const f4v*)p);
    o[0] = q.x; o[1] = q.y; o[2] = q.z; o[3] = q.w;
}

__device__ __forceinline__ void store4(float* p, const float* v) {
    f4v q;
    q.x = v[0]; q.y = v[1]; q.z = v[2]; q.w = v[3];
    __builtin_nontemporal_store(q, (f4v*)p);
}

// SplitMix64 finaliser (Steele, Lea, Flood 2014): a bijective 64-bit mix.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// In-kernel Poisson drive of neurons n0 .. n0+3 (n0 even) of `lane` at this
// step, as a 4-bit mask (bit k = neuron n0+k gets the kick): a SplitMix64
// stream per (lane, local step), stream = mix64(seed ^ step*0x100000001b3);
// output q = mix64(stream + (q+1)*gamma) gives neurons 2q (bits 40..63) and
// 2q+1 (bits 16..39) a uniform in [0, 1) each; a neuron gets the kick when
// its uniform falls below the lane's rate.
struct DriveArgs {
    uint64_t seed;
    int64_t t0;
    float prob;
};

// The lane's seed, clock and rate (loaded ahead of the state).
__device__ __forceinline__ DriveArgs drive_args(const StepArgs& a, int64_t lane) {
    DriveArgs d{0, 0, 0.f};
    if (a.in_prob == nullptr) return d;
    d.seed = (uint64_t)ld_const(a.in_seed + lane);
    d.t0 = ld_const(a.in_t0 + lane);
    d.prob = ld_const(a.in_prob + lane);
    return d;
}

__device__ __forceinline__ uint32_t poisson_drive4(const StepArgs& a, const DriveArgs& d,
                                                   int64_t n0) {
    if (a.in_prob == nullptr) return 0u;
    const uint64_t step = (uint64_t)(d.t0 + a.t);
    const uint64_t stream = mix64(d.seed ^ (step * 0x100000001b3ull));
    const float pr = d.prob;
    uint32_t m = 0;
#pragma unroll
    for (int h2 = 0; h2 < 2; ++h2) {
        const uint64_t q = (uint64_t)((n0 >> 1) + h2);
        const uint64_t h = mix64(stream + (q + 1) * 0x9e3779b97f4a7c15ull);
        const float u0 = (float)(h >> 40) * (1.0f / 16777216.0f);
        const float u1 = (float)((h >> 16) & 0xffffffull) * (1.0f / 16777216.0f);
        m |= (u0 < pr ? 1u : 0u) << (2 * h2);
        m |= (u1 < pr ? 1u : 0u) << (2 * h2 + 1);
    }
    return m;
}

// Somatic input of a neuron: the stored psc (decayed current of the previous
// step) plus the previous step's summed push `wz`, plus x and the drive.
template <bool REC>
__device__ __forceinline__ float soma_input(const StepArgs& a, int64_t i, float& psc, float wz,
                                           float xin, float drive) {
    if (a.prev) {
        psc = psc + wz;
        if (REC && a.rec_psc_prev) a.rec_psc_prev[i] = psc;
    }
    xin = xin + drive;
    return a.psc ? psc + xin : xin;
}

// Records, spike counts and membrane samples of a neuron after its update
// (REC: the launch writes any of them), then the synapse's exponential decay
// of psc.
template <bool REC>
__device__ __forceinline__ void step_records(const StepArgs& a, int64_t lane, int64_t nidx,
                                             int64_t i, float s, float v) {
    if (REC) {
        if (a.rec_v) a.rec_v[i] = v;
        const bool live = a.lane_t_end == nullptr || a.in_t0[lane] + a.t < a.lane_t_end[lane];
        if (a.counts && s != 0.f && live) a.counts[i] += 1;
        if (a.rec_sample && nidx < a.n_sample && live) a.rec_sample[lane * a.n_sample + nidx] = v;
        if (a.rec_z) a.rec_z[i] = s;
    }
}

template <bool REC, bool SHR = false>
__device__ __forceinline__ void step_outputs(const StepArgs& a, int64_t lane, int64_t nidx,
                                             int64_t i, float s, float v, float& psc,
                                             const PreParams& q) {
    step_records<REC>(a, lane, nidx, i, s, v);
    if (a.psc) {
        const float tsyn = prm_sh<SHR>(a.tsyn, a.tsyn_s, nidx, q.tsyn);
        const float k = prm_sh<SHR>(a.ksyn, a.ksyn_s, nidx, q.ksyn);
        psc = psc + k * ((-psc) / tsyn);
    }
}

// Synapse states besides psc of one neuron: h (alpha kinds) or the rise and
// decay conductances (dual exponential).
struct Syn2 {
    float h1, h2;
};

// Somatic input with alpha or dual-exponential kinetics.  The previous
// step's summed push W z completes that step's synapse update
// (synapse.py:243-253, 307-317, 394-406): alpha kinds h = h + gain*(W z),
// the dual exponential g_r + W z, g_d + W z and psc = a*(g_d - g_r).  psc
// then holds the current the neuron receives.
template <bool REC>
__device__ __forceinline__ float syn_x_input(const StepArgs& a, int64_t i, int64_t nidx,
                                            float& psc, Syn2& s2, float wz, float xin,
                                            float drive) {
    if (a.prev) {
        if (a.syn == BTX_SYN_DUAL_EXP) {
            s2.h1 = s2.h1 + wz;
            s2.h2 = s2.h2 + wz;
            psc = prm(a.sa, a.sa_s, nidx) * (s2.h2 - s2.h1);
        } else {
            s2.h1 = s2.h1 + prm(a.sg, a.sg_s, nidx) * wz;
        }
        if (REC) {
            if (a.rec_psc_prev) a.rec_psc_prev[i] = psc;
            if (a.rec_sh1_prev) a.rec_sh1_prev[i] = s2.h1;
            if (a.rec_sh2_prev) a.rec_sh2_prev[i] = s2.h2;
        }
    }
    xin = xin + drive;
    return psc + xin;
}

// The decay part of the next step's synapse update (conductance_charge,
// synapse.py:239-241, 304-305, 390-392, through exp_euler_step, ode.py:50);
// the push W z of this step completes it at the next step.
__device__ __forceinline__ void syn_x_decay(const StepArgs& a, int64_t nidx, float& psc,
                                           Syn2& s2) {
    const float k = prm(a.ksyn, a.ksyn_s, nidx);
    if (a.syn == BTX_SYN_ALPHA) {
        const float ts = prm(a.tsyn, a.tsyn_s, nidx);
        const float h = s2.h1;
        psc = psc + k * ((-psc) / ts + h / ts);
        s2.h1 = h + k * ((-h) / ts);
    } else if (a.syn == BTX_SYN_ALPHA_BILLEH) {
        const float h = s2.h1;
        psc = k * psc + k * h;
        s2.h1 = k * h;
    } else {
        const float tr = prm(a.tsyn, a.tsyn_s, nidx);
        const float td = prm(a.tsyn2, a.tsyn2_s, nidx);
        const float kd = prm(a.ksyn2, a.ksyn2_s, nidx);
        s2.h1 = s2.h1 + k * ((-s2.h1) / tr);
        s2.h2 = s2.h2 + kd * ((-s2.h2) / td);
    }
}

// One LIF or GLIF3 neuron of the fused step.  `psc` enters as the stored psc
// and leaves as this step's decayed current.  `ia` holds the neuron's
// after-spike currents (GLIF3).  Returns the spike (0/1).
template <int KIND, bool REC, bool SHR, bool SX = false>
__device__ __forceinline__ float lif_exp_neuron(const StepArgs& a, int64_t lane, int64_t nidx,
                                                int64_t i, float& psc, float wz, float xin,
                                                float drive, float& v, float& ref, float* ia,
                                                const PreParams& q, float tau, Syn2* s2 = nullptr,
                                                int hard = -1) {
    const float I = SX ? syn_x_input<REC>(a, i, nidx, psc, *s2, wz, xin, drive)
                       : soma_input<REC>(a, i, psc, wz, xin, drive);
    const float vth = prm_sh<SHR>(a.vth, a.vth_s, nidx, q.vth);
    const float vr = prm_sh<SHR>(a.vr, a.vr_s, nidx, q.vr);
    const float cm = prm_sh<SHR>(a.cm, a.cm_s, nidx, q.cm);
    if (REC && a.rec_v_prev) a.rec_v_prev[i] = v;
    float v1;
    if (KIND == kGlif3) {
        // charge (exponential Euler, Iasc of the previous step)
        float isum = 0.f;
#pragma unroll
        for (int j = 0; j < kMaxAsc; ++j)
            if (j < a.n_asc) isum = isum + ia[j];
        const float dv = (-(v - prm(a.vrest, a.vrest_s, nidx))) / tau +
                         (I + isum) / cm;
        v1 = v + prm(a.kv, a.kv_s, nidx) * dv;
        // adaptation
        const int64_t qa = nidx * a.n_asc;
#pragma unroll
        for (int j = 0; j < kMaxAsc; ++j)
            if (j < a.n_asc) ia[j] = ia[j] + a.asc_kk[qa + j] * ((-a.asc_k[qa + j]) * ia[j]);
    } else {
        // SHR launches with c_m == 1 skip the division (x / 1 == x)
        const float dv = (-(v - vr)) / tau + (SHR ? I : I / cm);
        v1 = v + a.dt * dv;
    }
    if (REC && a.rec_v_pre) a.rec_v_pre[i] = v1;
    float s = fires(v1 - vth, vth - vr) ? 1.f : 0.f;
    if (a.use_ref) {
        s = s * ((ref == 0.f) ? 1.f : 0.f);
        float r2 = (ref + s * prm_sh<SHR>(a.tref, a.tref_s, nidx, q.tref)) - a.dt;
        ref = (r2 < 0.f) ? 0.f : r2;
        if (REC && a.rec_ref) a.rec_ref[i] = ref;
    }
    const bool hr = hard < 0 ? a.hard != 0 : hard != 0;
    v = hr ? (v1 - (v1 - vr) * s) : (v1 - (vth - vr) * s);
    if (KIND == kGlif3) {
        const int64_t qa = nidx * a.n_asc;
#pragma unroll
        for (int j = 0; j < kMaxAsc; ++j)
            if (j < a.n_asc) ia[j] = ia[j] + a.asc_amp[qa + j] * s;
    }
    if (SX) {
        step_records<REC>(a, lane, nidx, i, s, v);
        syn_x_decay(a, nidx, psc, *s2);
    } else {
        step_outputs<REC, SHR>(a, lane, nidx, i, s, v, psc, q);
    }
    return s;
}

// One two-compartment neuron (two_compartment.py:208-292): semi-implicit
// apical current with the calcium plateau, semi-implicit soma with the
// exponential initiation term, Heaviside spike at the adaptive threshold,
// hard reset, bAP current for the next step, threshold decay and increment.
// `xa` is the apical input.  Returns the spike (0/1).
template <bool REC>
__device__ __forceinline__ float two_comp_neuron(const StepArgs& a, int64_t lane, int64_t nidx,
                                                 int64_t i, float& psc, float wz, float xin,
                                                 float drive, float xa, float& v, float& ia,
                                                 float& ibap, float& th, const PreParams& q) {
#define TCP(k) prm(a.tcp[k], a.tcs[k], nidx)
    const float I = soma_input<REC>(a, i, psc, wz, xin, drive);
    if (REC && a.rec_v_prev) a.rec_v_prev[i] = v;
    // apical current: w_Ca * sigmoid(i_a - theta_Ca) is +-0 for w_Ca == 0
    // whenever the sigmoid is a number, so it is only evaluated otherwise
    const float wca = TCP(BTX_TC_W_CA);
    const float thca = TCP(BTX_TC_THETA_CA);
    const float z = ia - thca;
    const float plateau = (wca == 0.f && z == z) ? wca * 0.5f
                                                 : wca * (1.f / (1.f + expf(-z)));
    const float ia1 = (ia + TCP(BTX_TC_DA) * ((xa + plateau) + ibap)) / TCP(BTX_TC_DA1);
    // soma
    const float thr = prm_or(a.vth, a.vth_s, nidx, q.vth) + th;
    const float total = I + TCP(BTX_TC_W_AS) * ia1;
    const float dT = TCP(BTX_TC_DT);
    float ex = 0.f;
    if (dT > 0.f) {
        float e = (v - thr) / TCP(BTX_TC_DTC);
        e = e > 10.f ? 10.f : e;  // keeps a NaN, as torch.clamp does
        ex = dT * expf(e);
    }
    const float vp =
        (v + TCP(BTX_TC_DS) * ((TCP(BTX_TC_E_L) + TCP(BTX_TC_R_S) * total) + ex)) /
        TCP(BTX_TC_DS1);
    if (REC && a.rec_v_pre) a.rec_v_pre[i] = vp;
    const float s = (vp - thr) >= 0.f ? 1.f : 0.f;
    ia = ia1;
    v = vp - (vp - prm_or(a.vr, a.vr_s, nidx, q.vr)) * s;
    ibap = TCP(BTX_TC_W_SA) * s;
    th = th / TCP(BTX_TC_DTH1) + TCP(BTX_TC_DELTA_TH) * s;
#undef TCP
    if (REC) {
        if (a.rec_ia) a.rec_ia[i] = ia;
        if (a.rec_ibap) a.rec_ibap[i] = ibap;
        if (a.rec_th) a.rec_th[i] = th;
    }
    step_outputs<REC>(a, lane, nidx, i, s, v, psc, q);
    return s;
}

// ---------------------------------------------------------------------------
// Reduction of a push into one segment
// ---------------------------------------------------------------------------

// Virtual row r = src*csr_d + sub is active in the push being reduced when
// neuron src spiked age_of_sub[sub] steps before it: ring slot
// (clock - age) mod R = slot0 - age (+R), as 0 <= age < R.  `lane_words`
// is the lane's row of slot 0.
__device__ __forceinline__ bool pulled_active(const PullArgs& a, const uint64_t* lane_words,
                                              int64_t slot_stride, int r) {
    const int src = r / a.csr_d;
    const int age = a.age_of_sub[r - src * a.csr_d];
    if (age < 0) return false;
    int slot = a.slot0 - age;
    slot += slot < 0 ? a.R : 0;
    const uint64_t* row = lane_words + slot * slot_stride;
    return (row[src >> 6] >> (src & 63)) & 1ull;
}

// Input of neuron `dst` of `lane`: its active incoming edges summed in the
// coalesced (dst, src) order from +0.0f.  Eight edges are requested at a
// time; the sum stays sequential in edge order.
__device__ float pull_sum(const PullArgs& a, int64_t lane, int64_t dst) {
    float acc = 0.f;
    int64_t e = a.csc_ptr[dst];
    const int64_t e1 = a.csc_ptr[dst + 1];
    const uint64_t* lw = a.ring_words + lane * a.nwords;
    const int64_t ss = a.lanes * a.nwords;
    constexpr int U = 8;
    for (; e + U <= e1; e += U) {
        int r[U];
        float w[U];
        bool b[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            r[u] = a.csc_src[e + u];
            w[u] = a.csc_w[e + u];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) b[u] = pulled_active(a, lw, ss, r[u]);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (b[u]) acc = acc + w[u];
    }
    for (; e < e1; ++e)
        if (pulled_active(a, lw, ss, a.csc_src[e])) acc = acc + a.csc_w[e];
    return acc;
}

// The requests that open the reduction of a segment: the bin's term count
// and its first kThreads terms (slots past the count hold stale terms and
// are ignored), and how the push was delivered.  Issued together, ahead of
// the state loads, so that waiting for them does not wait for the state;
// after a dense pull the bin head is stale and unused.
struct RedHead {
    int c;
    int d, r;
    float w;
    int pull;
};

__device__ __forceinline__ RedHead reduce_head(const ReduceIn& a, int64_t nchunks, int64_t lane,
                                               int64_t seg) {
    RedHead h{0, 0, 0, 0.f, 0};
    if (!a.ordered) return h;
    const int64_t q0 = (lane * nchunks + seg) * (int64_t)kBinCapMax + threadIdx.x;
    h.d = a.ev_d[q0];
    h.r = a.ev_r[q0];
    h.w = a.ev_w[q0];
    h.c = ld_const(a.bin_cnt + lane * nchunks + seg);
    // a vector load, waited for with the bin head rather than with the
    // scalar loads of the launch's parameters
    h.pull = a.pull_mode != nullptr
                 ? __hip_atomic_load(a.pull_mode, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                 : 0;
    return h;
}

// A bin of more terms than LDS holds (kBinCap < c <= cap): the same counting
// sort by neuron, with the order kept in the bin's global order array, then
// each neuron sums its terms in ascending virtual row by repeated selection
// of the next larger row (its few terms stay in L2).  Neurons with more
// than sort_max terms pull instead.  L.cnt is zero on entry (set by the
// caller, before its barrier).
__device__ void reduce_big_bin(const ReduceIn& a, RedLds& L, int64_t n, int64_t lane, int64_t q0,
                               int c, int64_t n0, float s4[4]) {
    const int tid = threadIdx.x;
    for (int i = tid; i < c; i += kThreads) atomicAdd(&L.cnt[a.ev_d[q0 + i] & (kChunk - 1)], 1);
    __syncthreads();
    int v[4];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = L.cnt[tid * 4 + k];
        sum += v[k];
    }
    int total;
    int base = block_excl_scan<kThreads / 64>(sum, total);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        L.cnt[tid * 4 + k] = base;
        base += v[k];
    }
    __syncthreads();
    uint16_t* ord = a.ev_o + q0;
    for (int i = tid; i < c; i += kThreads) {
        const int p = atomicAdd(&L.cnt[a.ev_d[q0 + i] & (kChunk - 1)], 1);
        ord[p] = (uint16_t)i;
    }
    __threadfence_block();
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int cd = v[k];
        if (cd == 0 || n0 + k >= n) continue;
        if (cd > a.sort_max) {
            s4[k] = pull_sum(a.pull, lane, n0 + k);
            continue;
        }
        const int end = L.cnt[tid * 4 + k], start = end - cd;
        float s = 0.f;
        int last = -1;
        for (int m = 0; m < cd; ++m) {
            int best = 0x7fffffff;
            float bw = 0.f;
            for (int x = start; x < end; ++x) {
                const int t = ord[x];
                const int r = a.ev_r[q0 + t];
                if (r > last && r < best) {
                    best = r;
                    bw = a.ev_w[q0 + t];
                }
            }
            s = s + bw;
            last = best;
        }
        s4[k] = s;
    }
}

// s4[k] = summed input of neuron seg*1024 + tid*4 + k of `lane` (+0.0f when
// it received none).  Every thread of the workgroup calls it.
__device__ __forceinline__ void reduce_segment(const ReduceIn& a, RedLds& L, int64_t n, int64_t nchunks,
                               int64_t lane, int64_t seg, float s4[4], const RedHead& h) {
    const int tid = threadIdx.x;
    const int64_t n0 = seg * kChunk + tid * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) s4[k] = 0.f;
    if (!a.ordered) {
        if (n0 < n) {
            float* q = a.acc + lane * n + n0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (n0 + k < n) {
                    s4[k] = q[k];
                    if (s4[k] != 0.f) q[k] = 0.f;
                }
            }
        }
        return;
    }
    if (h.pull) {
        // the push was delivered as dense sums: the bins stayed empty
        if (n0 < n) {
            const float* q = a.pull_acc + lane * n + n0;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (n0 + k < n) s4[k] = q[k];
        }
        return;
    }
    int* cntp = a.bin_cnt + lane * nchunks + seg;
    const int64_t q0 = (lane * nchunks + seg) * (int64_t)kBinCapMax;
    const int d_pre = h.d, r_pre = h.r;
    const float w_pre = h.w;
    const int c = h.c;
    if (c == 0) return;  // uniform across the workgroup
#pragma unroll
    for (int k = 0; k < 4; ++k) L.cnt[tid * 4 + k] = 0;
    __syncthreads();
    if (tid == 0) *cntp = 0;  // every thread has read c
    if (c < 0 || c > a.cap) {
        // overflowed bin: every neuron of the segment pulls
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (n0 + k < n) s4[k] = pull_sum(a.pull, lane, n0 + k);
        return;
    }
    if (c > kBinCap) {
        reduce_big_bin(a, L, n, lane, q0, c, n0, s4);
        return;
    }
    for (int i = tid; i < c; i += kThreads) {
        const bool pre = i == tid;
        const int d = (pre ? d_pre : (int)a.ev_d[q0 + i]) & (kChunk - 1);
        L.d[i] = (uint16_t)d;
        L.r[i] = pre ? r_pre : a.ev_r[q0 + i];
        L.w[i] = pre ? w_pre : a.ev_w[q0 + i];
        atomicAdd(&L.cnt[d], 1);
    }
    __syncthreads();
    int v[4];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = L.cnt[tid * 4 + k];
        sum += v[k];
    }
    int total;
    int base = block_excl_scan<kThreads / 64>(sum, total);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        L.cnt[tid * 4 + k] = base;  // a thread rewrites only the counts it read
        base += v[k];
    }
    __syncthreads();
    for (int i = tid; i < c; i += kThreads) {
        const int p = atomicAdd(&L.cnt[L.d[i]], 1);
        L.ord[p] = (uint16_t)i;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int cd = v[k];
        if (cd == 0 || n0 + k >= n) continue;
        if (cd <= a.sort_max) {
            const int end = L.cnt[tid * 4 + k], start = end - cd;
            // insertion sort of the neuron's terms by virtual row
            for (int x = start + 1; x < end; ++x) {
                const int key = L.ord[x];
                const int rk = L.r[key];
                int y = x - 1;
                while (y >= start && L.r[L.ord[y]] > rk) {
                    L.ord[y + 1] = L.ord[y];
                    --y;
                }
                L.ord[y + 1] = (uint16_t)key;
            }
            float s = 0.f;
            for (int x = start; x < end; ++x) s = s + L.w[L.ord[x]];
            s4[k] = s;
        } else {
            s4[k] = pull_sum(a.pull, lane, n0 + k);
        }
    }
}

__device__ __forceinline__ void reduce_segment(const ReduceIn& a, RedLds& L, int64_t n,
                                               int64_t nchunks, int64_t lane, int64_t seg,
                                               float s4[4]) {
    reduce_segment(a, L, n, nchunks, lane, seg, s4, reduce_head(a, nchunks, lane, seg));
}

// ---------------------------------------------------------------------------
// F role
// ---------------------------------------------------------------------------

// Ascending list of one segment: thread tid holds the pushed bits `bits` of
// neurons seg_base + tid*4 + b.  Every thread of the workgroup calls it.
__device__ __forceinline__ int write_segment(uint32_t bits, int64_t seg_base,
                                             int32_t* __restrict__ idx,
                                             int32_t* __restrict__ cnt) {
    int total;
    int pos = block_excl_scan<kThreads / 64>(__popc(bits), total);
    while (bits) {
        const int b = __ffs(bits) - 1;
        idx[seg_base + pos++] = (int32_t)(seg_base + threadIdx.x * 4 + b);
        bits &= bits - 1;
    }
    if (threadIdx.x == 0) *cnt = total;
    return total;
}

// The lane-bit ring of the dense pull: lane `lane` pushed neurons n0 + k
// (bit k of `bits`) into the slot whose rows start at `slot_bits` (zeroed one
// launch earlier).
__device__ __forceinline__ void mark_lane_bits(uint64_t* __restrict__ slot_bits, int lw,
                                               int64_t n0, int64_t lane, uint32_t bits) {
    const int64_t g = lane >> 6;
    const unsigned long long m = 1ull << (lane & 63);
    while (bits) {
        const int b = __ffs(bits) - 1;
        atomicOr((unsigned long long*)(slot_bits + (n0 + b) * lw + g), m);
        bits &= bits - 1;
    }
}

// The lane-bit ring of the dense pull, transposed from the ring's spike
// words without atomics (when the ring is rebuilt): one wavefront turns word `wi` of the 64 lanes of
// group g (one load per lane) into the 64 neurons' lane words, bit b of
// neuron 64*wi + k = bit k of lane 64*g + b, one ballot per neuron.  Every
// lane of the wavefront calls it with the same (wi, g).
__device__ __forceinline__ void transpose_lane_bits(const uint64_t* __restrict__ slot_words,
                                                    uint64_t* __restrict__ slot_bits,
                                                    int64_t lanes, int64_t nwords, int64_t n,
                                                    int lw, int64_t wi, int g) {
    const int l = (int)(threadIdx.x & 63);
    const int64_t lane = 64 * (int64_t)g + l;
    const uint64_t word = lane < lanes ? slot_words[lane * nwords + wi] : 0ull;
    uint64_t mine = 0ull;
#pragma unroll 8
    for (int k = 0; k < 64; ++k) {
        const uint64_t m = __ballot((word >> k) & 1ull);
        if (l == k) mine = m;
    }
    const int64_t j = 64 * wi + l;
    if (j < n) slot_bits[j * lw + g] = mine;
}

// State of neurons n0 .. n0+3 of a lane (`room` of them exist): one 16-byte
// access when `vec` (uniform), else one per neuron.  Neurons past n read the
// lane's first element `safe` instead (their values are never used), so no
// lane branches and the loads of several arrays stay in flight together.
__device__ __forceinline__ void load_lane4(const float* p, int64_t i0, int64_t room, int64_t safe,
                                           bool vec, float* o) {
    if (vec) {
        load4(p + (room > 0 ? i0 : safe), o);
        return;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) o[k] = p[k < room ? i0 + k : safe];
}

__device__ __forceinline__ void store_lane4(float* p, int64_t i0, int64_t room, bool vec,
                                            const float* v) {
    if (vec && room >= 4) {
        store4(p + i0, v);
        return;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (k < room) p[i0 + k] = v[k];
}

// Whether any of four values has a bit set (-0.0 counts as set).
__device__ __forceinline__ bool any4(const float* a) {
    return (__float_as_uint(a[0]) | __float_as_uint(a[1]) | __float_as_uint(a[2]) |
            __float_as_uint(a[3])) != 0u;
}

template <int KIND_SX>
__device__ __forceinline__ void f_role(const StepArgs& a, StepLds& lds, int64_t seg, int64_t lane) {
    constexpr int KIND = KIND_SX & ~kSynX;
    constexpr bool SX = (KIND_SX & kSynX) != 0;
    if (a.lane_t_end != nullptr && a.in_t0[lane] + a.t >= a.lane_t_end[lane]) {
        // no trial on this lane: consume what the previous push left and keep
        // the state (the lane is reset when a job is admitted to it)
        if (a.prev) {
            float wz[4];
            reduce_segment(a.red, lds.red, a.n, a.nchunks, lane, seg, wz);
        }
        // and push nothing: an empty ring entry, so that the spike totals and
        // the lane bits count only lanes with a trial
        if (a.ring_idx != nullptr) {
            const int64_t wi = (seg * kChunk >> 6) + threadIdx.x;
            if (threadIdx.x < kChunk / 64 && wi < a.nwords) a.ring_words[lane * a.nwords + wi] = 0ull;
            if (threadIdx.x == 0) a.ring_cnt[lane * a.nchunks + seg] = 0;
        }
        return;
    }
    constexpr bool MIX = KIND == kMixed;
    constexpr bool GLIF = KIND == kGlif3 || MIX;
    constexpr bool TC = KIND == kTwoComp || MIX;
    const int tid = threadIdx.x;
    const int64_t seg_base = seg * kChunk;
    const int64_t n0 = seg_base + tid * 4;
    const int64_t room = a.n - n0;
    const bool vec = a.vec4 != 0;
    const int64_t i0 = lane * a.n + n0;
    // requests in the order they are waited for: the bin head, the shared
    // parameters and drive arguments, then the state
    RedHead head{0, 0, 0, 0.f};
    if (a.prev) head = reduce_head(a.red, a.nchunks, lane, seg);
    PreParams q;
    q.vth = ld_const(a.vth);
    q.vr = ld_const(a.vr);
    q.cm = KIND == kTwoComp ? 1.f : ld_const(a.cm);
    q.tref = a.use_ref ? ld_const(a.tref) : 0.f;
    q.tsyn = a.psc ? ld_const(a.tsyn) : 1.f;
    q.ksyn = a.psc ? ld_const(a.ksyn) : 0.f;
    const DriveArgs dargs = drive_args(a, lane);
    const bool mark = a.lane_bits != nullptr && ld_const(a.lane_bits_on) != 0;
    const int64_t safe = lane * a.n;
    float v[4] = {0.f, 0.f, 0.f, 0.f}, p[4] = {0.f, 0.f, 0.f, 0.f};
    float r[4] = {0.f, 0.f, 0.f, 0.f}, x[4] = {0.f, 0.f, 0.f, 0.f};
    float ta[4] = {0.f, 0.f, 0.f, 0.f}, tb[4] = {0.f, 0.f, 0.f, 0.f};
    float tt[4] = {0.f, 0.f, 0.f, 0.f}, xa[4] = {0.f, 0.f, 0.f, 0.f};
    load_lane4(a.v, i0, room, safe, vec, v);
    if (a.psc) load_lane4(a.psc, i0, room, safe, vec, p);
    float h1[4] = {0.f, 0.f, 0.f, 0.f}, h2[4] = {0.f, 0.f, 0.f, 0.f};
    if (SX) {
        load_lane4(a.sh1, i0, room, safe, vec, h1);
        if (a.syn == BTX_SYN_DUAL_EXP) load_lane4(a.sh2, i0, room, safe, vec, h2);
    }
    if (a.use_ref) load_lane4(a.ref, i0, room, safe, vec, r);
    if (a.x) load_lane4(a.x, i0, room, safe, vec, x);
    if (TC && (!MIX || a.tia != nullptr)) {
        load_lane4(a.tia, i0, room, safe, vec, ta);
        load_lane4(a.tibap, i0, room, safe, vec, tb);
        load_lane4(a.tth, i0, room, safe, vec, tt);
        if (a.xa) load_lane4(a.xa, i0, room, safe, vec, xa);
    }
    uint32_t kind4 = 0;  // kMixed: model codes of the four neurons, a byte each
    if (MIX) {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (k < room) kind4 |= (uint32_t)a.kinds[n0 + k] << (8 * k);
    }
    // the membrane time constants, per neuron (shared by the lanes) or one value
    float tau4[4] = {1.f, 1.f, 1.f, 1.f};
    if (!TC || MIX) {
        if (a.tau_s) {
            load_lane4(a.tau, n0, room, 0, vec, tau4);
        } else {
            const float t0 = ld_const(a.tau);
            tau4[0] = tau4[1] = tau4[2] = tau4[3] = t0;
        }
    }
    float ia[4][GLIF ? kMaxAsc : 1];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int j = 0; j < (GLIF ? kMaxAsc : 1); ++j)
            ia[k][j] = (GLIF && a.iasc != nullptr && j < a.n_asc && k < room)
                           ? a.iasc[(i0 + k) * a.n_asc + j] : 0.f;
    // the drive's hashing runs while the loads are in flight
    const uint32_t kicks = poisson_drive4(a, dargs, n0);
    float wz[4] = {0.f, 0.f, 0.f, 0.f};
    if (a.prev) reduce_segment(a.red, lds.red, a.n, a.nchunks, lane, seg, wz, head);
    // the refractory counter and the bAP current are mostly 0 before and
    // after the step: they are written back only where either is not
    const bool ref_live = any4(r);
    const bool bap_live = TC && any4(tb);
    uint32_t nib = 0;
    // variants: with records (always for GLIF3 and the two-compartment
    // neuron, to hold their register count), and for LIF without records,
    // general or with every parameter but tau shared and c_m == 1 (fewer
    // uniform values live through the loop)
    const bool shr = a.shared && q.cm == 1.f;
    if (MIX) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k < room) {
                const float drv = ((kicks >> k) & 1u) ? a.in_kick : 0.f;
                const uint32_t code = (kind4 >> (8 * k)) & 0xffu;
                const uint32_t model = code & BTX_MIXED_MODEL_MASK;
                const int hard = (code & BTX_MIXED_HARD_RESET) ? 1 : 0;
                float s;
                if (model == BTX_NEURON_TWO_COMPARTMENT)
                    s = two_comp_neuron<true>(a, lane, n0 + k, i0 + k, p[k], wz[k], x[k], drv,
                                              xa[k], v[k], ta[k], tb[k], tt[k], q);
                else if (model == BTX_NEURON_GLIF3)
                    s = lif_exp_neuron<kGlif3, true, false>(a, lane, n0 + k, i0 + k, p[k], wz[k],
                                                            x[k], drv, v[k], r[k], ia[k], q,
                                                            tau4[k], nullptr, hard);
                else
                    s = lif_exp_neuron<kLif, true, false>(a, lane, n0 + k, i0 + k, p[k], wz[k],
                                                          x[k], drv, v[k], r[k], ia[k], q,
                                                          tau4[k], nullptr, hard);
                nib |= (uint32_t)(s != 0.f) << k;
            }
        }
    } else if (SX) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k < room) {
                const float drv = ((kicks >> k) & 1u) ? a.in_kick : 0.f;
                Syn2 s2{h1[k], h2[k]};
                const float s = lif_exp_neuron<KIND, true, false, true>(
                    a, lane, n0 + k, i0 + k, p[k], wz[k], x[k], drv, v[k], r[k], ia[k], q,
                    tau4[k], &s2);
                h1[k] = s2.h1;
                h2[k] = s2.h2;
                nib |= (uint32_t)(s != 0.f) << k;
            }
        }
    } else if (KIND != kLif || a.rec_any) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k < room) {
                const float drv = ((kicks >> k) & 1u) ? a.in_kick : 0.f;
                const float s =
                    TC ? two_comp_neuron<true>(a, lane, n0 + k, i0 + k, p[k], wz[k], x[k], drv,
                                               xa[k], v[k], ta[k], tb[k], tt[k], q)
                       : lif_exp_neuron<KIND, true, false>(a, lane, n0 + k, i0 + k, p[k], wz[k],
                                                           x[k], drv, v[k], r[k], ia[k], q,
                                                           tau4[k]);
                nib |= (uint32_t)(s != 0.f) << k;
            }
        }
    } else if (shr) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k < room) {
                const float drv = ((kicks >> k) & 1u) ? a.in_kick : 0.f;
                const float s = lif_exp_neuron<KIND, false, true>(
                    a, lane, n0 + k, i0 + k, p[k], wz[k], x[k], drv, v[k], r[k], ia[k], q,
                    tau4[k]);
                nib |= (uint32_t)(s != 0.f) << k;
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k < room) {
                const float drv = ((kicks >> k) & 1u) ? a.in_kick : 0.f;
                const float s = lif_exp_neuron<KIND, false, false>(
                    a, lane, n0 + k, i0 + k, p[k], wz[k], x[k], drv, v[k], r[k], ia[k], q,
                    tau4[k]);
                nib |= (uint32_t)(s != 0.f) << k;
            }
        }
    }
    if (GLIF && (!MIX || a.iasc != nullptr)) {
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int j = 0; j < (GLIF ? kMaxAsc : 1); ++j)
                if (j < a.n_asc && k < room) a.iasc[(i0 + k) * a.n_asc + j] = ia[k][j];
    }
    if (room > 0) {
        store_lane4(a.v, i0, room, vec, v);
        if (a.use_ref && (ref_live || any4(r))) store_lane4(a.ref, i0, room, vec, r);
        if (a.psc) store_lane4(a.psc, i0, room, vec, p);
        if (SX) {
            store_lane4(a.sh1, i0, room, vec, h1);
            if (a.syn == BTX_SYN_DUAL_EXP) store_lane4(a.sh2, i0, room, vec, h2);
        }
        if (TC && (!MIX || a.tia != nullptr)) {
            store_lane4(a.tia, i0, room, vec, ta);
            if (bap_live || any4(tb)) store_lane4(a.tibap, i0, room, vec, tb);
            store_lane4(a.tth, i0, room, vec, tt);
        }
    }
    // spike word of 16 threads (64 neurons), by an OR across them
    const int64_t wi = n0 >> 6;
    uint64_t word = (uint64_t)nib << (4 * (tid & 15));
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) word |= __shfl_xor(word, o, 64);
    const bool lead = (tid & 15) == 0 && wi < a.nwords;
    if (a.raster != nullptr && lead) a.raster[lane * a.nwords + wi] = word;
    if (a.ring_idx != nullptr) {
        uint32_t pushed = nib;
        uint64_t pw = word;
        if (a.teacher != nullptr) {
            pw = wi < a.nwords ? a.teacher[lane * a.nwords + wi] : 0ull;
            pushed = (uint32_t)((pw >> (4 * (tid & 15))) & 15ull);
            const int64_t room = a.n - n0;  // bits past n never count
            if (room < 4) pushed &= room > 0 ? (1u << room) - 1u : 0u;
        }
        if (lead) a.ring_words[lane * a.nwords + wi] = pw;
        if (mark) mark_lane_bits(a.lane_bits, a.lw, n0, lane, pushed);
        write_segment(pushed, seg_base, a.ring_idx + lane * a.n,
                      a.ring_cnt + lane * a.nchunks + seg);
    }
}

// ---------------------------------------------------------------------------
// P role
// ---------------------------------------------------------------------------

__device__ __forceinline__ int ring_slot(int64_t clock, int delay, int R) {
    const int64_t s = (clock - delay) % R;
    return (int)(s < 0 ? s + R : s);
}

// Ordered mode, one workgroup per launch: the spike total of the ring slot
// written one launch earlier (complete now; summed from its segment counts,
// no atomics), then how the next launch delivers its push.  That push reads
// the slots at ages dl - 1 from this launch's clock for its bins' delays dl:
// the totals of ages >= 1 are known, and a slot younger than that (written
// in this launch, or the next) is taken to hold as many spikes as the
// youngest known one.  The pull is taken when the push would bin more than
// a threshold of terms per (lane, segment) on average (set by the host) and every slot it reads
// (ages up to ring_d - 1) had its lane bits marked when it was written; the
// marking is kept on while the push would bin more than an eighth of that.
// Either way the sums are the same bits, so an estimate only ever costs
// time.
__device__ void control_role(const ControlArgs& a) {
    const int tid = threadIdx.x;
    const int s1 = ring_slot(a.clock, 1, a.R);
    const int32_t* cnt = a.ring_cnt + (int64_t)s1 * a.m;
    int sum = 0;
    for (int64_t i = tid; i < a.m; i += kThreads) sum += cnt[i];
    int total;
    block_excl_scan<kThreads / 64>(sum, total);
    if (tid != 0) return;
    a.ctl_rw->slot_total[s1] = total;
    float rows = 0.f;
    for (int j = 0; j < a.nbins; ++j) {
        const int dl = a.affine ? a.delay0 + j : a.ctl->bin_delay[j];
        rows += dl >= 3 ? (float)a.ctl_rw->slot_total[ring_slot(a.clock, dl - 1, a.R)]
                        : (float)total;
    }
    const float terms = rows * a.terms_per_row;
    int run = a.ctl_rw->lane_bits_run;
    run = a.ctl_rw->lane_bits_mode[a.clock & 3] ? (run < (1 << 20) ? run + 1 : run) : 0;
    a.ctl_rw->lane_bits_run = run;
    const bool pull = terms > a.above && run >= a.R;
    a.ctl_rw->push_mode[(a.clock + 1) & 3] = pull ? 1 : 0;
    a.ctl_rw->lane_bits_mode[(a.clock + 1) & 3] = (pull || terms > 0.125f * a.above) ? 1 : 0;
}

// Division of 0 <= x < 2^31 by a run-time 1 <= d <= 64 without a divide:
// q = (umulhi(x, m) + x) >> l with l = ceil(log2 d), m = 2^32 (2^l - d) / d + 1
// (the round-up multiplier of Granlund and Montgomery, "Division by
// invariant integers using multiplication", 1994; checked against x / d for
// every d on [0, 2^20) and a stride of the rest of [0, 2^31)).
struct FastDiv {
    uint32_t m;
    int l;
};

__device__ __forceinline__ FastDiv fast_div(int d) {
    int l = 0;
    while ((1 << l) < d) ++l;
    const uint64_t m = ((uint64_t)1 << 32) * (((uint64_t)1 << l) - (uint64_t)d) / (uint64_t)d + 1;
    return FastDiv{(uint32_t)m, l};
}

__device__ __forceinline__ uint32_t div_u31(const FastDiv& f, uint32_t x) {
    return (__umulhi(x, f.m) + x) >> f.l;
}

// The dense pull of one step: workgroup `wid` of `nwg` sums, for a range of
// destinations and every lane, the weights of the active incoming edges in
// the coalesced (dst, virtual source) order from +0.0f -- the terms and the
// order of the ordered push, so the same bits.  Virtual row r = src*csr_d +
// sub is active for lane b when bit b of the lane-bit word of src in the
// slot that sub's age points at is set (every slot the step reads was marked
// when it was written: the pull is only taken then).  Groups of Gp threads
// (one per lane of a 64-lane group) walk one destination's edges together.
// WAVE (k_pull only): a 64-lane group takes 64 of a destination's edges at a
// time, lane k reading edge c + k and its source's lane-bit word, and the
// wavefront walks them in order out of register lanes (three requests per
// 64 edges instead of three per edge).  Inside k_step this form put the
// kernel at 300-400 B of scratch per lane next to the other, so the step
// keeps the 12-edge rounds.
template <bool WAVE>
__device__ void pull_role(const PushStep& a, PushLds& L, int64_t wid, int64_t nwg) {
    const int tid = threadIdx.x;
    const FastDiv fd = fast_div(a.csr_d);
    for (int k = tid; k < a.csr_d; k += kThreads) {
        const int age = a.ctl->age_of_sub[k];
        L.slot[k] = age < 0 ? -1 : ring_slot(a.clock, age, a.R);
    }
    __syncthreads();
    const int64_t per = (a.n + nwg - 1) / nwg;
    const int64_t d0 = wid * per;
    const int64_t d1 = d0 + per < a.n ? d0 + per : a.n;
    for (int g = 0; g < a.lw; ++g) {
        const int lanes_g = (int)(a.lanes - 64 * (int64_t)g < 64 ? a.lanes - 64 * (int64_t)g : 64);
        int Gp = 1;
        while (Gp < lanes_g) Gp <<= 1;
        const int gl = tid % Gp;
        const int64_t lane = 64 * (int64_t)g + gl;
        const uint64_t* bits = a.lane_bits + g;
        for (int64_t d = d0 + tid / Gp; d < d1; d += kThreads / Gp) {
            float acc = 0.f;
            int64_t e = a.csc_ptr[d];
            const int64_t e1 = a.csc_ptr[d + 1];
            if (WAVE && Gp == 64) {
                // one destination per wavefront: e and e1 are uniform.  The
                // next 64 edges and their lane-bit words are requested before
                // the current ones are summed
                const int sh = gl & 31;
                auto edges = [&](int64_t c, int& wi) -> uint64_t {
                    const int m = (int)(e1 - c < 64 ? e1 - c : 64);
                    const bool in = c < e1 && gl < m;
                    const int r = in ? a.csc_src[c + gl] : 0;
                    wi = in ? __float_as_int(a.csc_w[c + gl]) : 0;
                    const int src = (int)div_u31(fd, (uint32_t)r);
                    const int slot = L.slot[r - src * a.csr_d];
                    return (in && slot >= 0) ? bits[((int64_t)slot * a.n + src) * a.lw] : 0ull;
                };
                int wi_next = 0;
                uint64_t wb_next = e < e1 ? edges(e, wi_next) : 0ull;
                for (int64_t c = e; c < e1; c += 64) {
                    const int m =
                        __builtin_amdgcn_readfirstlane((int)(e1 - c < 64 ? e1 - c : 64));
                    const int wi = wi_next;
                    const uint64_t wb = wb_next;
                    if (c + 64 < e1) wb_next = edges(c + 64, wi_next);
                    const int lo = (int)(uint32_t)wb, hi = (int)(uint32_t)(wb >> 32);
                    for (int k = 0; k < m; ++k) {
                        const uint32_t wl = (uint32_t)__builtin_amdgcn_readlane(lo, k);
                        const uint32_t wh = (uint32_t)__builtin_amdgcn_readlane(hi, k);
                        if (((gl < 32 ? wl : wh) >> sh) & 1u)
                            acc = acc + __int_as_float(__builtin_amdgcn_readlane(wi, k));
                    }
                }
                e = e1;
            }
            // 12 edges in flight per round of dependent loads (edges, then
            // lane bits): the most that keeps the step within its registers
            constexpr int U = 12;
            for (; e < e1; e += U) {
                int r[U];
                float w[U];
                uint64_t wb[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const bool in = e + u < e1;
                    r[u] = in ? a.csc_src[e + u] : 0;
                    w[u] = in ? a.csc_w[e + u] : 0.f;
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int src = (int)div_u31(fd, (uint32_t)r[u]);
                    const int slot = L.slot[r[u] - src * a.csr_d];
                    wb[u] = (e + u < e1 && slot >= 0)
                                ? bits[((int64_t)slot * a.n + src) * a.lw] : 0ull;
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if ((wb[u] >> gl) & 1ull) acc = acc + w[u];
            }
            if (gl < lanes_g) {
                const bool live = a.lane_t_end == nullptr ||
                                  a.in_t0[lane] + a.t < a.lane_t_end[lane];
                a.pull_acc[lane * a.n + d] = live ? acc : 0.f;
            }
        }
    }
}

template <int G, bool ORDERED>
__device__ void p_role(const PushStep& a, PushLds& L, int64_t px, int64_t lane) {
    // workgroups px >= np only join a dense pull; a lane past its trial end
    // pushes nothing (a dense pull still covers its share of the
    // destinations)
    const int64_t np = (a.nchunks + a.cpb - 1) / a.cpb;
    const bool helper = px >= np;
    const bool idle = helper || (a.lane_t_end != nullptr &&
                                 a.in_t0[lane] + a.t >= a.lane_t_end[lane]);
    const int64_t nwg = np * a.lanes, wid = lane * np + px;
    int pull = 0;
    if (ORDERED && a.lane_bits != nullptr) {
        // how the previous launch decided this push is delivered (looked at
        // once the segment counts are in, so that their requests do not wait
        // for it).  While this launch marks lane bits, those of the slot the
        // next launch writes start empty (spread over the P workgroups; no
        // workgroup of this launch reads that slot: its age would be
        // ring_d).  The first marking launch after a pause marks a slot left
        // unzeroed, which no pull reads: the pull needs R marking launches
        // in a row, by when that slot is written again.
        pull = ld_const(&a.ctl->push_mode[a.clock & 3]);
        const int64_t zn =
            !helper && ld_const(&a.ctl->lane_bits_mode[a.clock & 3]) ? a.n * a.lw : 0;
        uint64_t* nb = a.lane_bits + (int64_t)ring_slot(a.clock + 1, 0, a.R) * a.n * a.lw;
        for (int64_t i = wid * kThreads + threadIdx.x; i < zn; i += nwg * kThreads) nb[i] = 0ull;
    }
    const int64_t c0 = helper ? 0 : px * a.cpb;
    const int ncb = (int)(a.nchunks - c0 < a.cpb ? a.nchunks - c0 : a.cpb);
    const int tid = threadIdx.x;
    if (tid < a.nbins) {
        // bins (delay0 + j, sub0 + j) need no table
        const int dl = a.affine ? a.delay0 + tid : a.ctl->bin_delay[tid];
        L.slot[tid] = ring_slot(a.clock, dl, a.R);
        L.sub[tid] = a.affine ? a.sub0 + tid : a.ctl->bin_sub[tid];
    }
    if (ORDERED) {
        for (int t = tid; t < a.nchunks; t += kThreads) L.tile[t] = 0;
        if (tid == 0) L.nev = 0;
    }
    __syncthreads();
    const int nseg = a.nbins * ncb;  // <= kThreads (host check)
    int c = 0;
    if (tid < nseg && !idle) {
        const int j = tid / ncb, cc = tid - j * ncb;
        c = a.ring_cnt[((int64_t)L.slot[j] * a.lanes + lane) * a.nchunks + c0 + cc];
        // a segment never lists more neurons than it has
        const int64_t room = a.n - (c0 + cc) * kChunk;
        const int cap = room < kChunk ? (int)room : kChunk;
        c = c < 0 ? 0 : (c > cap ? cap : c);
    }
    int total;
    const int off = block_excl_scan<kThreads / 64>(c, total);
    if (tid < nseg) L.off[tid] = off;
    if (tid == 0) L.off[nseg] = total;
    __syncthreads();
    if (ORDERED && pull) {
        if (!a.pull_sep) {
            const int64_t npp = np + a.nh;
            pull_role<false>(a, L, lane * npp + px, npp * a.lanes);
        }
        return;
    }
    if (total == 0) return;

    constexpr int gpb = kThreads / G;
    const int gl = tid % G;
    const int64_t key0 = lane * a.n;
    // pass 0: every term of the workgroup's rows, counted per destination
    // segment and staged in LDS (fast mode: added to the accumulator);
    // pass 1 (only when the stage overflowed) streams the rows again
    int staged = 0;
    for (int pass = 0; pass < (ORDERED ? 2 : 1); ++pass) {
        if (pass == 1 && staged <= kStage) break;
        for (int item = tid / G; item < total; item += gpb) {
            int lo = 0, hi = nseg;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (L.off[mid] <= item) lo = mid;
                else hi = mid;
            }
            const int j = lo / ncb, cc = lo - j * ncb;
            const int k = item - L.off[lo];
            const int src = a.ring_idx[((int64_t)L.slot[j] * a.lanes + lane) * a.n +
                                       (c0 + cc) * kChunk + k];
            if ((uint32_t)src >= (uint32_t)a.n) continue;  // never indexes past the CSR
            const int r = src * a.csr_d + L.sub[j];
            const int64_t beg = a.row_ptr[r], end = a.row_ptr[r + 1];
            for (int64_t e = beg + gl; e < end; e += G) {
                const int dst = a.col[e];
                const float wv = a.w[e];
                if (!ORDERED) {
                    atomicAdd(&a.acc[key0 + dst], wv);
                } else if (pass == 0) {
                    atomicAdd(&L.tile[dst >> kChunkShift], 1);
                    const int q = atomicAdd(&L.nev, 1);
                    if (q < kStage) {
                        L.er[q] = r;
                        L.ew[q] = wv;
                        L.ed[q] = dst;
                    }
                } else {
                    const int tl = dst >> kChunkShift;
                    const int pos = L.base[tl] + atomicAdd(&L.tile[tl], 1);
                    if (pos < a.cap) {
                        const int64_t q = (lane * a.nchunks + tl) * (int64_t)kBinCapMax + pos;
                        a.ev_r[q] = r;
                        a.ev_w[q] = wv;
                        a.ev_d[q] = (uint16_t)(dst & (kChunk - 1));
                    }
                }
            }
        }
        if (ORDERED && pass == 0) {
            // one reservation per destination segment this workgroup touches
            __syncthreads();
            staged = L.nev;
            for (int t = tid; t < a.nchunks; t += kThreads) {
                const int ct = L.tile[t];
                if (ct) {
                    const int b = atomicAdd(&a.bin_cnt[lane * a.nchunks + t], ct);
                    L.base[t] = (uint16_t)(b < a.cap ? b : a.cap);
                    L.tile[t] = 0;
                }
            }
            __syncthreads();
            if (staged <= kStage) {
                for (int i = tid; i < staged; i += kThreads) {
                    const int dst = L.ed[i];
                    const int tl = dst >> kChunkShift;
                    const int pos = L.base[tl] + atomicAdd(&L.tile[tl], 1);
                    if (pos < a.cap) {
                        const int64_t q = (lane * a.nchunks + tl) * (int64_t)kBinCapMax + pos;
                        a.ev_r[q] = L.er[i];
                        a.ev_w[q] = L.ew[i];
                        a.ev_d[q] = (uint16_t)(dst & (kChunk - 1));
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// The roles are inlined (the F role always, the P role by its call for the
// larger neuron kinds): a role left out of line (the compiler's choice for
// those kinds) takes its argument block by reference, and the whole block
// (~1 KB) is then copied into scratch memory for every lane, which slows the
// launch several times.
// A row of workgroups per lane: [0, nc) is the control column (nc = 0 or 1;
// lane 0's sums the slot totals and decides the next push, the others end at
// once), [nc, nc + np) play P, the next nh (= p.nh; a scalar argument of its
// own, like np, so that the role choice reads no argument block) only join a
// dense pull (a launch of few lanes and segments has too few P workgroups to
// pull every destination; they end at once after a push), the rest F.
template <int G, bool ORDERED, int KIND, bool TR>
__global__ __launch_bounds__(kThreads, KIND == kLif ? 7 : 1) void k_step(StepArgs a, PushStep p,
                                                                          ControlArgs ca, int np,
                                                                          int nc, int nh) {
    // KIND: kLif, kGlif3, kTwoComp or kMixed, with kSynX for the alpha and
    // dual-exponential synapses
    __shared__ StepLds lds;
    // Row positions of a lane: the control workgroup (lane 0 of position 0
    // when nc = 1), the P workgroups, then the F workgroups.  TR (a grid the
    // GPU holds at once): the grid is (lanes, positions) with the positions
    // reversed, so, dispatched lane fastest, every lane's F workgroups go
    // out first.  Otherwise (positions, lanes): one lane's row after
    // another, so the P and F work of the lanes dispatched first overlaps
    // (measured on 256-lane pools: lane-fastest dispatch cost 25% of the
    // pool's rate).  A compile-time choice: picking blockIdx.x or .y at run
    // time made the compiler spill scalar registers to scratch.
    int64_t lane;
    int x;
    if constexpr (TR) {
        lane = blockIdx.x;
        x = (int)(gridDim.y - 1 - blockIdx.y) - nc;
    } else {
        // workgroup i of the dispatch runs on XCD i mod 8: with lanes a
        // multiple of 8, XCD k takes the rows of lanes k, k + 8, ... in
        // order, so a lane's state, bins and ring stay in one XCD's L2
        const uint32_t cols = gridDim.x;
        if ((gridDim.y & 7u) == 0) {
            const uint32_t i = blockIdx.x + blockIdx.y * cols;
            const uint32_t slot = i >> 3;
            const uint32_t q = slot / cols;
            lane = (int64_t)(q * 8u + (i & 7u));
            x = (int)(slot - q * cols) - nc;
        } else {
            lane = blockIdx.y;
            x = (int)blockIdx.x - nc;
        }
    }
    if (x < 0) {
        if (ORDERED && lane == 0) control_role(ca);
    } else if (x < np + nh) {
        if constexpr (KIND == kLif) {
            p_role<G, ORDERED>(p, lds.push, x, lane);
        } else {
            // the larger kinds: the compiler would leave the P role out of line
            // and copy its argument block to scratch (see f_role above)
            [[clang::always_inline]] p_role<G, ORDERED>(p, lds.push, x, lane);
        }
    } else {
        f_role<KIND>(a, lds, (int64_t)x - np - nh, lane);
    }
}

// The dense pull of a step as a launch of its own, after the step's launch
// (batches of more than 32 lanes): one wavefront per destination and 64-lane
// group, four destinations per workgroup.  Ends at once unless the previous
// launch chose the pull for this step.
__global__ __launch_bounds__(kThreads) void k_pull(PushStep p) {
    __shared__ PushLds L;
    if (ld_const(&p.ctl->push_mode[p.clock & 3]) == 0) return;
    pull_role<true>(p, L, (int64_t)blockIdx.x, (int64_t)gridDim.x);
}

// FIN: the last push reduced into the synapse (and recorded: psc in `rec`,
// the other synapse states in a.rec_sh1_prev / a.rec_sh2_prev).
__global__ __launch_bounds__(kThreads) void k_finish(StepArgs a, float* __restrict__ rec) {
    __shared__ StepLds lds;
    const int64_t lane = blockIdx.y, seg = blockIdx.x;
    float wz[4];
    reduce_segment(a.red, lds.red, a.n, a.nchunks, lane, seg, wz);
    const int64_t n0 = seg * kChunk + threadIdx.x * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (n0 + k < a.n) {
            const int64_t nidx = n0 + k;
            const int64_t i = lane * a.n + nidx;
            if (a.syn == BTX_SYN_EXP) {
                const float q = a.psc[i] + wz[k];
                a.psc[i] = q;
                if (rec) rec[i] = q;
            } else if (a.syn == BTX_SYN_DUAL_EXP) {
                const float g1 = a.sh1[i] + wz[k];
                const float g2 = a.sh2[i] + wz[k];
                const float q = prm(a.sa, a.sa_s, nidx) * (g2 - g1);
                a.sh1[i] = g1;
                a.sh2[i] = g2;
                a.psc[i] = q;
                if (rec) rec[i] = q;
                if (a.rec_sh1_prev) a.rec_sh1_prev[i] = g1;
                if (a.rec_sh2_prev) a.rec_sh2_prev[i] = g2;
            } else {
                const float h = a.sh1[i] + prm(a.sg, a.sg_s, nidx) * wz[k];
                a.sh1[i] = h;
                if (rec) rec[i] = a.psc[i];
                if (a.rec_sh1_prev) a.rec_sh1_prev[i] = h;
            }
        }
    }
}

// Ring lists of all ring slots from spike words [R][lanes][nwords]
// (grid: segments x lanes x slots).
__global__ __launch_bounds__(kThreads) void k_ring_from_words(const uint64_t* __restrict__ words,
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
    write_segment(nib, seg_base, ring_idx + (slot * lanes + lane) * n,
                  ring_cnt + (slot * lanes + lane) * nchunks + seg);
}

// The lane bits of every ring slot, transposed from its spike words: one
// wavefront per (word, 64-lane group, slot).
__global__ __launch_bounds__(64) void k_ring_lane_bits(const uint64_t* __restrict__ words,
                                                       uint64_t* __restrict__ lane_bits,
                                                       int64_t lanes, int64_t nwords, int64_t n,
                                                       int lw) {
    const int64_t wi = blockIdx.x, slot = blockIdx.z;
    transpose_lane_bits(words + slot * lanes * nwords, lane_bits + slot * n * lw, lanes, nwords,
                        n, lw, wi, (int)blockIdx.y);
}

// The spike total of every ring slot from its segment counts (one workgroup
// per slot), every push mode back to the term bins, and the lane bits (just
// transposed for every slot) current and kept.
__global__ __launch_bounds__(kThreads) void k_ring_totals(const int32_t* __restrict__ ring_cnt,
                                                          int64_t m, Ctl* __restrict__ ctl) {
    const int64_t slot = blockIdx.x;
    int sum = 0;
    for (int64_t i = threadIdx.x; i < m; i += kThreads) sum += ring_cnt[slot * m + i];
    int total;
    block_excl_scan<kThreads / 64>(sum, total);
    if (threadIdx.x == 0) ctl->slot_total[slot] = total;
    if (slot == 0 && threadIdx.x < 4) {
        ctl->push_mode[threadIdx.x] = 0;
        ctl->lane_bits_mode[threadIdx.x] = 1;
    }
    if (slot == 0 && threadIdx.x == 0) ctl->lane_bits_run = 1 << 20;
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

// Standalone decay of an alpha or dual-exponential synapse: the same update
// as the fused step's (psc from psc and h, then h; or both conductances).
__global__ __launch_bounds__(256) void k_syn_decay(StepArgs a, int64_t m) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const int64_t nidx = i % a.n;
    const bool dual = a.syn == BTX_SYN_DUAL_EXP;
    float psc = a.psc[i];
    Syn2 s2{a.sh1[i], dual ? a.sh2[i] : 0.f};
    syn_x_decay(a, nidx, psc, s2);
    a.sh1[i] = s2.h1;
    if (dual) a.sh2[i] = s2.h2;
    else a.psc[i] = psc;
}

// The in-kernel drive as a dense tensor: one thread per (lane, neuron pair)
// of step blockIdx.y, the same numbers poisson_drive4 draws.
__global__ __launch_bounds__(256) void k_poisson_fill(const uint64_t* __restrict__ seed,
                                                      const float* __restrict__ prob,
                                                      const int64_t* __restrict__ t0, float kick,
                                                      int64_t lanes, int64_t n,
                                                      float* __restrict__ x) {
    const int64_t t = blockIdx.y;
    const int64_t npair = (n + 1) >> 1;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= lanes * npair) return;
    const int64_t lane = g / npair, q = g - lane * npair;
    const uint64_t step = (uint64_t)(t0[lane] + t);
    const uint64_t stream = mix64(seed[lane] ^ (step * 0x100000001b3ull));
    const uint64_t h = mix64(stream + (uint64_t)(q + 1) * 0x9e3779b97f4a7c15ull);
    const float pr = prob[lane];
    const float u0 = (float)(h >> 40) * (1.0f / 16777216.0f);
    const float u1 = (float)((h >> 16) & 0xffffffull) * (1.0f / 16777216.0f);
    float* row = x + (t * lanes + lane) * n;
    const int64_t j = 2 * q;
    row[j] = u0 < pr ? kick : 0.f;
    if (j + 1 < n) row[j + 1] = u1 < pr ? kick : 0.f;
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

// Engine workspace: control tables, then two buffers of term bins (ordered)
// or of accumulators (fast).
struct EngineWs {
    size_t bin_cnt = 0, ev_r = 0, ev_w = 0, ev_d = 0, ev_o = 0, acc = 0, total = 0;
    int64_t bins = 0;  // per buffer
    int64_t m = 0;     // accumulators per buffer
};

static EngineWs engine_ws(int64_t lanes, int64_t n, int ordered) {
    EngineWs L;
    const int64_t nchunks = (n + kChunk - 1) / kChunk;
    size_t off = align256(sizeof(Ctl));
    if (ordered) {
        L.bins = lanes * nchunks;
        L.bin_cnt = off;
        off = align256(off + sizeof(int) * 2 * L.bins);
        L.ev_r = off;
        off = align256(off + sizeof(int32_t) * 2 * L.bins * kBinCapMax);
        L.ev_w = off;
        off = align256(off + sizeof(float) * 2 * L.bins * kBinCapMax);
        L.ev_d = off;
        off = align256(off + sizeof(uint16_t) * 2 * L.bins * kBinCapMax);
        L.ev_o = off;
        off = align256(off + sizeof(uint16_t) * 2 * L.bins * kBinCapMax);
        L.m = lanes * n;  // dense pull sums, two buffers
        L.acc = off;
        off = align256(off + sizeof(float) * 2 * L.m);
    } else {
        L.m = lanes * n;
        L.acc = off;
        off = align256(off + sizeof(float) * 2 * L.m);
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
    a.shared = (a.vth_s == 0 && a.vr_s == 0 && a.cm_s == 0 && a.tref_s == 0 && a.tsyn_s == 0 &&
                a.ksyn_s == 0) ? 1 : 0;
    a.v = p->v;
    a.ref = p->refractory;
    a.kind = p->neuron_model == BTX_NEURON_GLIF3             ? kGlif3
             : p->neuron_model == BTX_NEURON_TWO_COMPARTMENT ? kTwoComp
             : p->neuron_model == BTX_NEURON_MIXED           ? kMixed
                                                             : kLif;
    const bool mixed = a.kind == kMixed;
    if (mixed) a.kinds = p->neuron_kinds;
    if (a.kind == kTwoComp ||
        (mixed && (p->mixed_models & (1 << BTX_NEURON_TWO_COMPARTMENT)))) {
        a.tia = p->i_a;
        a.tibap = p->i_bap;
        a.tth = p->theta_th;
        for (int k = 0; k < BTX_TC_NPARAM; ++k) {
            a.tcp[k] = p->tc[k];
            a.tcs[k] = p->tc_stride[k];
        }
    }
    if (a.kind == kGlif3 || (mixed && (p->mixed_models & (1 << BTX_NEURON_GLIF3)))) {
        a.n_asc = p->n_asc;
        a.vrest = p->v_rest; a.vrest_s = p->v_rest_stride;
        a.kv = p->k_v; a.kv_s = p->k_v_stride;
        a.asc_k = p->asc_k;
        a.asc_kk = p->asc_kk;
        a.asc_amp = p->asc_amp;
        a.iasc = p->iasc;
    }
    return a;
}

static bool aligned16(const void* q) { return ((uintptr_t)q & 15) == 0; }

static int check_neuron(const btx_rsnn_plan* p) {
    if (p->neuron_model != BTX_NEURON_LIF && p->neuron_model != BTX_NEURON_GLIF3 &&
        p->neuron_model != BTX_NEURON_TWO_COMPARTMENT && p->neuron_model != BTX_NEURON_MIXED)
        return fail("unsupported neuron model");
    const bool mixed = p->neuron_model == BTX_NEURON_MIXED;
    const int models = mixed ? p->mixed_models : (1 << p->neuron_model);
    if (mixed) {
        if (p->neuron_kinds == nullptr) return fail("a mixed population needs neuron_kinds");
        if (models == 0 || (models & ~7) != 0)
            return fail("mixed_models must name LIF, GLIF3 and/or two-compartment neurons");
        if (p->syn_kind != BTX_SYN_EXP)
            return fail("a mixed population runs with the exponential synapse");
        if (p->tau == nullptr || p->c_m == nullptr || p->v_th == nullptr ||
            p->v_reset == nullptr)
            return fail("a mixed population needs tau, c_m, v_th and v_reset for every neuron");
    }
    if (models & (1 << BTX_NEURON_TWO_COMPARTMENT)) {
        if (p->i_a == nullptr || p->i_bap == nullptr || p->theta_th == nullptr)
            return fail("two-compartment neuron needs i_a, i_bap and theta_th state");
        for (int k = 0; k < BTX_TC_NPARAM; ++k)
            if (p->tc[k] == nullptr || p->tc_stride[k] < 0 || p->tc_stride[k] > 1)
                return fail("two-compartment neuron needs every coefficient (stride 0 or 1)");
        if (p->use_refractory && !mixed)
            return fail("two-compartment neuron has no refractory state");
    }
    if ((models & (1 << BTX_NEURON_GLIF3)) &&
        (p->n_asc < 1 || p->n_asc > kMaxAsc || p->v_rest == nullptr || p->k_v == nullptr ||
         p->asc_k == nullptr || p->asc_kk == nullptr || p->asc_amp == nullptr ||
         p->iasc == nullptr))
        return fail("GLIF3 needs 1..4 after-spike currents, their parameters and state");
    if (p->lanes > 65535) return fail("lanes exceeds 65535");
    if (p->use_refractory && (p->refractory == nullptr || p->tau_ref == nullptr))
        return fail("refractory state missing");
    return 0;
}

static int set_rec_any(const StepArgs& a) {
    return (a.rec_z || a.rec_v || a.rec_ref || a.rec_psc_prev || a.rec_v_pre || a.rec_v_prev ||
            a.counts || a.rec_sample || a.rec_ia || a.rec_ibap || a.rec_th || a.rec_sh1_prev ||
            a.rec_sh2_prev)
               ? 1 : 0;
}

// 16-byte accesses are possible when every state array is aligned and n % 4 == 0.
static int vec4_ok(const btx_rsnn_plan* p, const float* x) {
    bool ok = (p->n % 4 == 0) && aligned16(p->v) && (!p->psc || aligned16(p->psc)) &&
              (!p->use_refractory || aligned16(p->refractory)) && (!x || aligned16(x)) &&
              (p->tau == nullptr || p->tau_stride == 0 || aligned16(p->tau));
    if (p->neuron_model == BTX_NEURON_TWO_COMPARTMENT ||
        (p->neuron_model == BTX_NEURON_MIXED &&
         (p->mixed_models & (1 << BTX_NEURON_TWO_COMPARTMENT))))
        ok = ok && aligned16(p->i_a) && aligned16(p->i_bap) && aligned16(p->theta_th) &&
             (!p->x_apical || aligned16(p->x_apical)) &&
             (p->x_apical_step_stride % 4 == 0);
    if (p->syn_kind != BTX_SYN_EXP)
        ok = ok && aligned16(p->syn_h1) &&
             (p->syn_kind != BTX_SYN_DUAL_EXP || aligned16(p->syn_h2));
    ok = ok && (p->x_step_stride % 4 == 0);
    return ok ? 1 : 0;
}

template <int KIND>
static void launch_neuron_only(const StepArgs& a, dim3 grid, hipStream_t s) {
    PushStep none{};
    hipLaunchKernelGGL((k_step<16, true, KIND, false>), grid, dim3(kThreads), 0, s, a, none,
                       ControlArgs{}, 0, 0, 0);
}

BTX_API int btx_lif_step(const btx_rsnn_plan* p, const float* x, float* z, float* v_pre,
                         void* stream) {
    if (p == nullptr) return fail("btx_lif_step: null plan");
    if (p->lanes <= 0 || p->n <= 0) return 0;
    if (int rc = check_neuron(p)) return rc;
    StepArgs a = neuron_args(p);
    a.x = x;
    a.rec_z = z;
    a.rec_v_pre = v_pre;
    a.xa = (a.kind == kTwoComp || a.kind == kMixed) ? p->x_apical : nullptr;
    a.vec4 = vec4_ok(p, x);
    a.rec_any = set_rec_any(a);
    const dim3 grid((unsigned)a.nchunks, (unsigned)p->lanes);  // (positions, lanes)
    const hipStream_t s = (hipStream_t)stream;
    if (a.kind == kGlif3) launch_neuron_only<kGlif3>(a, grid, s);
    else if (a.kind == kTwoComp) launch_neuron_only<kTwoComp>(a, grid, s);
    else if (a.kind == kMixed) launch_neuron_only<kMixed>(a, grid, s);
    else launch_neuron_only<kLif>(a, grid, s);
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

BTX_API int btx_syn_decay(const btx_rsnn_plan* p, void* stream) {
    if (p == nullptr) return fail("btx_syn_decay: null plan");
    const int64_t m = p->lanes * p->n;
    if (m <= 0) return 0;
    if (p->syn_kind <= BTX_SYN_EXP || p->syn_kind > BTX_SYN_DUAL_EXP)
        return fail("btx_syn_decay: syn_kind must be an alpha or dual-exponential kind");
    if (p->psc == nullptr || p->syn_h1 == nullptr || p->tau_syn == nullptr ||
        p->k_syn == nullptr)
        return fail("btx_syn_decay: psc, syn_h1, tau_syn and k_syn are required");
    if (p->syn_kind == BTX_SYN_DUAL_EXP &&
        (p->syn_h2 == nullptr || p->tau_syn2 == nullptr || p->k_syn2 == nullptr))
        return fail("btx_syn_decay: the dual exponential synapse needs syn_h2, tau_syn2, k_syn2");
    StepArgs a{};
    a.n = p->n;
    a.psc = p->psc;
    a.syn = p->syn_kind;
    a.sh1 = p->syn_h1;
    a.sh2 = p->syn_h2;
    a.tsyn = p->tau_syn; a.tsyn_s = p->tau_syn_stride;
    a.ksyn = p->k_syn; a.ksyn_s = p->k_syn_stride;
    a.tsyn2 = p->tau_syn2; a.tsyn2_s = p->tau_syn2_stride;
    a.ksyn2 = p->k_syn2; a.ksyn2_s = p->k_syn2_stride;
    hipLaunchKernelGGL(k_syn_decay, dim3((unsigned)((m + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, a, m);
    return check_launch("btx_syn_decay");
}

BTX_API int btx_poisson_fill(const uint64_t* seed, const float* prob, const int64_t* t0,
                             float kick, int64_t T, int64_t lanes, int64_t n, float* x,
                             void* stream) {
    if (T <= 0 || lanes <= 0 || n <= 0) return 0;
    if (seed == nullptr || prob == nullptr || t0 == nullptr || x == nullptr)
        return fail("btx_poisson_fill: null pointer");
    if (T > 65535) return fail("btx_poisson_fill: T exceeds 65535 steps per call");
    const int64_t work = lanes * ((n + 1) >> 1);
    if ((work + 255) / 256 > 0x7fffffff) return fail("btx_poisson_fill: too many lanes x neurons");
    hipLaunchKernelGGL(k_poisson_fill, dim3((unsigned)((work + 255) / 256), (unsigned)T),
                       dim3(256), 0, (hipStream_t)stream, seed, prob, t0, kick, lanes, n, x);
    return check_launch("btx_poisson_fill");
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

// Workgroups of k_step<G, ORDERED, KIND, true> the device holds at once
// (its occupancy per CU times the CU count), queried on first use.
template <int G, bool ORDERED, int KIND>
static int64_t resident_wgs() {
    static const int64_t n = [] {
        int blocks = 0, dev = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &blocks, reinterpret_cast<const void*>(&k_step<G, ORDERED, KIND, true>), kThreads,
                0) != hipSuccess)
            blocks = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            cus = 0;
        return (int64_t)blocks * cus;
    }();
    return n;
}

// `grid` is (lanes, row positions).  The LIF step launches it as it is
// (positions reversed) when every workgroup is resident at once; otherwise,
// and for the other kinds, it is launched transposed.
template <int G, bool ORDERED, int KIND>
static void launch_step_one(const StepArgs& a, const PushStep& ps, const ControlArgs& ca, int np,
                            int nc, dim3 grid, hipStream_t s) {
    if constexpr (KIND == kLif) {
        if ((int64_t)grid.x * grid.y <= resident_wgs<G, ORDERED, KIND>()) {
            hipLaunchKernelGGL((k_step<G, ORDERED, KIND, true>), grid, dim3(kThreads), 0, s, a,
                               ps, ca, np, nc, ps.nh);
            return;
        }
    }
    hipLaunchKernelGGL((k_step<G, ORDERED, KIND, false>), dim3(grid.y, grid.x), dim3(kThreads), 0,
                       s, a, ps, ca, np, nc, ps.nh);
}

template <bool ORDERED, int KIND>
static void launch_step_g(const StepArgs& a, const PushStep& ps, const ControlArgs& ca, int np,
                          int nc, int group, dim3 grid, hipStream_t s) {
    switch (group) {
        case 8: launch_step_one<8, ORDERED, KIND>(a, ps, ca, np, nc, grid, s); break;
        case 16: launch_step_one<16, ORDERED, KIND>(a, ps, ca, np, nc, grid, s); break;
        case 32: launch_step_one<32, ORDERED, KIND>(a, ps, ca, np, nc, grid, s); break;
        default: launch_step_one<64, ORDERED, KIND>(a, ps, ca, np, nc, grid, s); break;
    }
}

template <bool ORDERED>
static void launch_step_k(const StepArgs& a, const PushStep& ps, const ControlArgs& ca, int np,
                          int nc, int group, dim3 grid, hipStream_t s) {
    if (a.syn != BTX_SYN_EXP) {
        if (a.kind == kGlif3) launch_step_g<ORDERED, kGlif3 | kSynX>(a, ps, ca, np, nc, group, grid, s);
        else launch_step_g<ORDERED, kLif | kSynX>(a, ps, ca, np, nc, group, grid, s);
        return;
    }
    if (a.kind == kGlif3) launch_step_g<ORDERED, kGlif3>(a, ps, ca, np, nc, group, grid, s);
    else if (a.kind == kTwoComp) launch_step_g<ORDERED, kTwoComp>(a, ps, ca, np, nc, group, grid, s);
    else if (a.kind == kMixed) launch_step_g<ORDERED, kMixed>(a, ps, ca, np, nc, group, grid, s);
    else launch_step_g<ORDERED, kLif>(a, ps, ca, np, nc, group, grid, s);
}

static void launch_step(int ordered, const StepArgs& a, const PushStep& ps,
                        const ControlArgs& ca, int np, int nc, int group, dim3 grid,
                        hipStream_t s) {
    if (ordered) launch_step_k<true>(a, ps, ca, np, nc, group, grid, s);
    else launch_step_k<false>(a, ps, ca, np, nc, group, grid, s);
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
    if (p->ring_idx == nullptr || p->ring_cnt == nullptr || p->ring_words == nullptr)
        return fail("btx_rsnn_run: the ring lists and words are required");
    if (p->tau_syn == nullptr || p->k_syn == nullptr)
        return fail("btx_rsnn_run: tau_syn and k_syn are required");
    if (p->syn_kind < BTX_SYN_EXP || p->syn_kind > BTX_SYN_DUAL_EXP)
        return fail("btx_rsnn_run: unknown syn_kind");
    if (p->syn_kind != BTX_SYN_EXP) {
        if (p->neuron_model == BTX_NEURON_TWO_COMPARTMENT)
            return fail("btx_rsnn_run: the two-compartment neuron runs with BTX_SYN_EXP only");
        if (p->syn_h1 == nullptr)
            return fail("btx_rsnn_run: this synapse kind needs the syn_h1 state");
        if (p->syn_kind == BTX_SYN_DUAL_EXP &&
            (p->syn_h2 == nullptr || p->tau_syn2 == nullptr || p->k_syn2 == nullptr ||
             p->syn_a == nullptr))
            return fail("btx_rsnn_run: the dual exponential synapse needs syn_h2, tau_syn2, "
                        "k_syn2 and syn_a");
        if (p->syn_kind != BTX_SYN_DUAL_EXP && p->syn_gain == nullptr)
            return fail("btx_rsnn_run: the alpha synapses need syn_gain");
    }
    if (p->n * (int64_t)p->csr_d >= ((int64_t)1 << 31))
        return fail("btx_rsnn_run: n*csr_d exceeds int32 row ids");
    if (p->nbins < 1 || p->nbins > BTX_MAX_BINS) return fail("btx_rsnn_run: bad nbins");
    if (p->ring_d < 1 || p->ring_d >= BTX_MAX_BINS)
        return fail("btx_rsnn_run: ring_d must be in [1, 63]");
    if (p->csr_d < 1 || p->csr_d > BTX_MAX_BINS)
        return fail("btx_rsnn_run: csr_d must be in [1, 64]");
    int min_delay = p->ring_d;
    for (int j = 0; j < p->nbins; ++j) {
        if (p->bin_delay[j] < 0 || p->bin_delay[j] >= p->ring_d)
            return fail("btx_rsnn_run: bin delay outside the ring");
        if (p->bin_sub[j] < 0 || p->bin_sub[j] >= p->csr_d)
            return fail("btx_rsnn_run: bin offset outside csr_d");
        if (p->bin_delay[j] < min_delay) min_delay = p->bin_delay[j];
    }
    const int64_t n = p->n, lanes = p->lanes, m = n * lanes;
    const int64_t nwords = (n + 63) / 64;
    const int64_t nchunks = (n + kChunk - 1) / kChunk;
    const int R = p->ring_d + 1;
    if (nchunks > kMaxChunks) return fail("btx_rsnn_run: n exceeds 1024 ring segments");
    if (p->ordered && (p->csc_ptr == nullptr || p->csc_src == nullptr || p->csc_w == nullptr))
        return fail("btx_rsnn_run: ordered mode needs the dst-major arrays");
    if (p->ordered && (p->slots < 1 || p->slots > kSortMax))
        return fail("btx_rsnn_run: ordered mode needs 1 <= slots <= 32");
    if (p->bin_cap < 0 || p->bin_cap > kBinCapMax)
        return fail("btx_rsnn_run: bin_cap must be in [0, 8192]");
    if (p->in_prob != nullptr && (p->in_seed == nullptr || p->in_t0 == nullptr))
        return fail("btx_rsnn_run: Poisson drive needs per-lane seeds and clocks");
    if (p->lane_t_end != nullptr && p->in_t0 == nullptr)
        return fail("btx_rsnn_run: lane_t_end needs per-lane clocks");
    if (p->rec_sample != nullptr && (p->n_sample < 1 || p->n_sample > p->n))
        return fail("btx_rsnn_run: n_sample must be in [1, n]");
    const EngineWs L = engine_ws(lanes, n, p->ordered);
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
    if (!p->tables_ready) hipLaunchKernelGGL(k_set_tables, dim3(1), dim3(64), 0, s, tb, ctl);
    // lane-bit ring of the dense pull (ordered mode with the plan's ring of
    // lane bits); 64-lane words per neuron
    uint64_t* lane_bits = p->ordered ? p->ring_lane_bits : nullptr;
    const int lw = (int)((lanes + 63) / 64);
    if (p->ring_rebuild) {
        hipLaunchKernelGGL(k_ring_from_words,
                           dim3((unsigned)nchunks, (unsigned)lanes, (unsigned)R),
                           dim3(kThreads), 0, s, p->ring_words, n, nwords, lanes, nchunks,
                           p->ring_idx, p->ring_cnt);
        if (lane_bits != nullptr) {
            hipLaunchKernelGGL(k_ring_lane_bits, dim3((unsigned)nwords, (unsigned)lw, (unsigned)R),
                               dim3(64), 0, s, p->ring_words, lane_bits, lanes, nwords, n, lw);
            hipLaunchKernelGGL(k_ring_totals, dim3((unsigned)R), dim3(kThreads), 0, s, p->ring_cnt,
                               lanes * nchunks, ctl);
        }
    }

    const int cap = p->bin_cap > 0 ? p->bin_cap : kBinCapMax;
    // buffer b of the term bins / accumulators, for the push of the launch
    // at `clock`
    auto bins_of = [&](int b, int64_t clock, ReduceIn& r) {
        r.ordered = p->ordered;
        r.cap = cap;
        r.sort_max = p->slots;
        if (p->ordered) {
            r.bin_cnt = (int*)(ws + L.bin_cnt) + b * L.bins;
            r.ev_r = (const int32_t*)(ws + L.ev_r) + b * L.bins * kBinCapMax;
            r.ev_w = (const float*)(ws + L.ev_w) + b * L.bins * kBinCapMax;
            r.ev_d = (const uint16_t*)(ws + L.ev_d) + b * L.bins * kBinCapMax;
            r.ev_o = (uint16_t*)(ws + L.ev_o) + b * L.bins * kBinCapMax;
            r.pull_mode = lane_bits ? &ctl->push_mode[clock & 3] : nullptr;
            r.pull_acc = (float*)(ws + L.acc) + b * L.m;
        } else {
            r.acc = (float*)(ws + L.acc) + b * L.m;
        }
    };
    StepArgs a = neuron_args(p);
    a.psc = p->psc;
    a.syn = p->syn_kind;
    if (a.syn != BTX_SYN_EXP) {
        a.sh1 = p->syn_h1;
        a.sh2 = p->syn_h2;
        a.sg = p->syn_gain; a.sg_s = p->syn_gain_stride;
        a.tsyn2 = p->tau_syn2; a.tsyn2_s = p->tau_syn2_stride;
        a.ksyn2 = p->k_syn2; a.ksyn2_s = p->k_syn2_stride;
        a.sa = p->syn_a; a.sa_s = p->syn_a_stride;
    }
    a.in_seed = p->in_seed;
    a.in_prob = p->in_prob;
    a.in_t0 = p->in_t0;
    a.in_kick = p->in_kick;
    a.counts = p->counts;
    a.n_sample = p->n_sample;
    a.lane_t_end = p->lane_t_end;
    a.vec4 = vec4_ok(p, p->x);
    a.red.pull.csc_ptr = p->csc_ptr;
    a.red.pull.csc_src = p->csc_src;
    a.red.pull.csc_w = p->csc_w;
    a.red.pull.ring_words = p->ring_words;
    a.red.pull.lanes = lanes;
    a.red.pull.nwords = nwords;
    a.red.pull.R = R;
    a.red.pull.csr_d = p->csr_d;
    a.red.pull.age_of_sub = ctl->age_of_sub;

    PushStep ps{};
    ps.row_ptr = p->row_ptr;
    ps.col = p->col;
    ps.w = p->w;
    ps.n = n;
    ps.lanes = lanes;
    ps.nchunks = nchunks;
    ps.nbins = p->nbins;
    ps.csr_d = p->csr_d;
    ps.R = R;
    ps.ctl = ctl;
    ps.ring_idx = p->ring_idx;
    ps.ring_cnt = p->ring_cnt;
    ps.cap = cap;
    ps.lane_t_end = p->lane_t_end;
    ps.in_t0 = p->in_t0;
    ps.lane_bits = lane_bits;
    ps.lw = lw;
    ps.csc_ptr = p->csc_ptr;
    ps.csc_src = p->csc_src;
    ps.csc_w = p->csc_w;
    ps.affine = 1;
    ps.delay0 = p->bin_delay[0];
    ps.sub0 = p->bin_sub[0];
    for (int j = 1; j < p->nbins; ++j)
        if (p->bin_delay[j] != ps.delay0 + j || p->bin_sub[j] != ps.sub0 + j) ps.affine = 0;
    // segments per P workgroup: one while the launch has a few workgroups
    // per CU (shorter, better balanced push chains: 4% faster than two at 8
    // lanes of the 140k connectome), two once lanes x segments reaches 16 per
    // CU of the 256 (7% faster at 64 lanes, 9% at 256, equal at 32); at most
    // kThreads (delay bin, segment) pairs each
    int cpb = lanes * nchunks >= 4096 ? 2 : 1;
    if (cpb * p->nbins > kThreads) cpb = kThreads / p->nbins;
    ps.cpb = cpb;
    const int npb = (int)((nchunks + cpb - 1) / cpb);
    const int group = pick_group((double)p->nnz / (double)(n * p->csr_d));
    // one launch per step when no bin reads the step's own spikes
    const bool fused = min_delay >= 1;
    const int nf = (int)nchunks;
    // the control column: the slot totals and the next launch's push mode
    const int nc = lane_bits != nullptr ? 1 : 0;
    // pull-only workgroups: a dense pull spreads over at least kPullWgs
    ps.nh = 0;
    // batches of more than 32 lanes pull in k_pull (a wavefront per 64-lane
    // group and destination)
    ps.pull_sep = lane_bits != nullptr && lanes > 32 ? 1 : 0;
    if (lane_bits != nullptr && !ps.pull_sep) {
        const int64_t want = (kPullWgs + lanes - 1) / lanes;
        ps.nh = want > npb ? (int)(want - npb) : 0;
    }
    // a wavefront per destination, at most 4096 workgroups: a launch that
    // only ends at once (the step pushed) costs ~8 us at 35k workgroups, ~4
    // at 2048, and the 16k-neuron pull 235 us at 4096 against 258 at 2048
    const int64_t pull_wgs = (n + kThreads / 64 - 1) / (kThreads / 64);
    const dim3 pull_grid((unsigned)(pull_wgs < 4096 ? pull_wgs : 4096));
    ControlArgs ca{};
    ca.ring_cnt = p->ring_cnt;
    ca.ctl = ctl;
    ca.ctl_rw = ctl;
    ca.m = lanes * nchunks;
    ca.R = R;
    ca.nbins = p->nbins;
    ca.affine = ps.affine;
    ca.delay0 = ps.delay0;
    ca.terms_per_row = (float)((double)p->nnz / (double)(n * p->csr_d));
    // a dense pull once the push would bin more than half the bin capacity
    // per (lane, segment) on average; in a launch of fewer bins than CUs
    // (each bin one workgroup's serial sort) already once the bins outgrow
    // what LDS sorts (kBinCap: the larger ones take the slower global order).
    // Measured at the 140k connectome, 8 lanes (1096 bins): the pull reads
    // all 50M edges, ~1.3 ms, against 0.1-0.4 ms for pushes of 1-4M terms;
    // at 21.7k neurons, one lane (22 bins): ~65 us against 137 us for a
    // push of 71k terms.
    const int above = (lanes * nchunks < 256 && kBinCap < cap / 2) ? kBinCap : cap / 2;
    ca.above = (float)above * (float)(lanes * nchunks);

    const int64_t rec_step = m;
    for (int64_t t = 0; t < n_steps; ++t) {
        a.prev = t > 0;
        if (a.prev) {
            bins_of((int)((t - 1) & 1), p->ring_cursor + t - 1, a.red);
            a.red.pull.clock = p->ring_cursor + t - 1;
            a.red.pull.slot0 = (int)mod(a.red.pull.clock, R);
        }
        a.x = p->x ? p->x + t * p->x_step_stride : nullptr;
        a.xa = ((a.kind == kTwoComp || a.kind == kMixed) && p->x_apical)
                   ? p->x_apical + t * p->x_apical_step_stride
                                                   : nullptr;
        a.t = t;
        a.rec_sample = p->rec_sample ? p->rec_sample + t * lanes * p->n_sample : nullptr;
        a.raster = p->raster + t * lanes * nwords;
        a.teacher = p->teacher ? p->teacher + t * lanes * nwords : nullptr;
        const int64_t slot = mod(p->ring_cursor + t, R);
        a.ring_idx = p->ring_idx + slot * lanes * n;
        a.ring_cnt = p->ring_cnt + slot * lanes * nchunks;
        a.ring_words = p->ring_words + slot * lanes * nwords;
        a.lane_bits = lane_bits ? lane_bits + slot * n * lw : nullptr;
        a.lw = lw;
        a.lane_bits_on = lane_bits ? &ctl->lane_bits_mode[(p->ring_cursor + t) & 3] : nullptr;
        a.rec_z = p->rec_z ? p->rec_z + t * rec_step : nullptr;
        a.rec_v = p->rec_v ? p->rec_v + t * rec_step : nullptr;
        a.rec_ref = p->rec_refractory ? p->rec_refractory + t * rec_step : nullptr;
        a.rec_v_pre = p->rec_v_pre ? p->rec_v_pre + t * rec_step : nullptr;
        a.rec_v_prev = p->rec_v_prev ? p->rec_v_prev + t * rec_step : nullptr;
        if (a.kind == kTwoComp || a.kind == kMixed) {
            a.rec_ia = p->rec_i_a ? p->rec_i_a + t * rec_step : nullptr;
            a.rec_ibap = p->rec_i_bap ? p->rec_i_bap + t * rec_step : nullptr;
            a.rec_th = p->rec_theta_th ? p->rec_theta_th + t * rec_step : nullptr;
        }
        a.rec_psc_prev = (p->rec_psc && t > 0) ? p->rec_psc + (t - 1) * rec_step : nullptr;
        a.rec_sh1_prev = (p->rec_syn_h1 && t > 0) ? p->rec_syn_h1 + (t - 1) * rec_step : nullptr;
        a.rec_sh2_prev = (p->rec_syn_h2 && t > 0) ? p->rec_syn_h2 + (t - 1) * rec_step : nullptr;
        a.rec_any = set_rec_any(a);
        ps.clock = p->ring_cursor + t;
        ca.clock = ps.clock;
        ps.t = t;
        ReduceIn cur{};
        bins_of((int)(t & 1), p->ring_cursor + t, cur);
        ps.bin_cnt = cur.bin_cnt;
        ps.ev_r = (int32_t*)cur.ev_r;
        ps.ev_w = (float*)cur.ev_w;
        ps.ev_d = (uint16_t*)cur.ev_d;
        ps.acc = cur.acc;
        ps.pull_acc = cur.pull_acc ? (float*)cur.pull_acc : nullptr;
        if (fused) {
            kt.begin(0);
            const dim3 grid((unsigned)lanes, (unsigned)(nc + npb + ps.nh + nf));
            launch_step(p->ordered, a, ps, ca, npb, nc, group, grid, s);
            kt.end(0);
            if (ps.pull_sep) {
                kt.begin(2);
                hipLaunchKernelGGL(k_pull, pull_grid, dim3(kThreads), 0, s, ps);
                kt.end(2);
            }
        } else {
            kt.begin(0);
            const dim3 fgrid((unsigned)lanes, (unsigned)nf);
            PushStep fonly = ps;  // the F launch has no pull-only workgroups
            fonly.nh = 0;
            launch_step(p->ordered, a, fonly, ca, 0, 0, group, fgrid, s);
            kt.end(0);
            kt.begin(2);
            const dim3 pgrid((unsigned)lanes, (unsigned)(nc + npb + ps.nh));
            launch_step(p->ordered, a, ps, ca, npb, nc, group, pgrid, s);
            if (ps.pull_sep) hipLaunchKernelGGL(k_pull, pull_grid, dim3(kThreads), 0, s, ps);
            kt.end(2);
        }
    }
    bins_of((int)((n_steps - 1) & 1), p->ring_cursor + n_steps - 1, a.red);
    a.rec_sh1_prev = p->rec_syn_h1 ? p->rec_syn_h1 + (n_steps - 1) * rec_step : nullptr;
    a.rec_sh2_prev = p->rec_syn_h2 ? p->rec_syn_h2 + (n_steps - 1) * rec_step : nullptr;
    a.red.pull.clock = p->ring_cursor + n_steps - 1;
    a.red.pull.slot0 = (int)mod(a.red.pull.clock, R);
    kt.begin(3);
    hipLaunchKernelGGL(k_finish, dim3((unsigned)nchunks, (unsigned)lanes), dim3(kThreads), 0, s,
                       a, p->rec_psc ? p->rec_psc + (n_steps - 1) * rec_step : nullptr);
    kt.end(3);
    return check_launch("btx_rsnn_run");
}
