/*
 * btorch_amd.h -- C ABI of libbtx.so, the gfx950 HIP kernels behind the
 * recurrent spiking-network time step (spike x sparse connectome, fused
 * neuron/synapse state update, delay ring, surrogate-gradient adjoint).
 *
 * Conventions
 *  - Every pointer is a device pointer unless stated otherwise.
 *  - No entry point allocates device memory: callers pass outputs and a
 *    workspace sized by btx_*_workspace_bytes().
 *  - Work is enqueued on `stream` (a hipStream_t passed as void*); nothing
 *    synchronises unless stated.
 *  - Return 0 on success, otherwise a nonzero status; btx_last_error() returns
 *    a thread-local message describing the last failure.
 *  - All floating-point state is IEEE binary32.  Kernels are compiled with
 *    -ffp-contract=off so each expression rounds exactly where the reference's
 *    eager op sequence rounds.
 */
#ifndef BTORCH_AMD_H
#define BTORCH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BTX_MAX_BINS 64

/* Neurons per segment of the delay ring's compacted spike lists: each
 * segment of BTX_RING_CHUNK consecutive neurons holds its own ascending list
 * and count, so the concatenation of a lane's segments in order is the
 * lane's ascending active list. */
#define BTX_RING_CHUNK 1024

/* Library version (major*10000 + minor*100 + patch). */
int btx_version(void);

/* Thread-local description of the last nonzero status. */
const char* btx_last_error(void);

/* ------------------------------------------------------------------------ */
/* Spike compaction                                                          */
/* ------------------------------------------------------------------------ */

/* Ascending indices of the nonzero entries of each lane of a dense fp32
 * matrix x[lanes, n] (row stride x_lane_stride).  idx[lane*idx_lane_stride + k]
 * receives the k-th nonzero column, counts[lane] the number of them.
 * tile_counts is scratch of lanes * btx_compact_tiles(n) int32 (any content
 * on entry).  Each lane is cut into tiles of 4096 columns: one launch counts
 * every tile with wavefront ballots over coalesced loads, a second writes
 * each tile's indices at the offset of the earlier tiles.  Replaces the "any
 * source of this task spiked?" scan of the Triton backend
 * (btorch/models/_sparse_triton.py:20-57) and equals numpy.nonzero per lane. */
int64_t btx_compact_tiles(int64_t n);
int btx_compact_dense(const float* x, int64_t lanes, int64_t n, int64_t x_lane_stride,
                      int32_t* idx, int64_t idx_lane_stride, int32_t* counts,
                      int32_t* tile_counts, void* stream);

/* Same, from a spike bitmask words[lanes, ceil(n/64)] (bit j of word w is
 * column 64*w + j).  Used by the time-step engine on the raster it writes. */
int btx_compact_bits(const uint64_t* words, int64_t lanes, int64_t n,
                     int64_t words_lane_stride, int32_t* idx, int64_t idx_lane_stride,
                     int32_t* counts, void* stream);

/* Spike words of each row of a dense fp32 matrix (bit j of word w of a row
 * = x[row, 64w + j] != 0): words[rows][ceil(n/64)].  The delay ring's
 * dense history (btorch/models/history.py:28-358) to bit form. */
int btx_pack_bits(const float* x, int64_t rows, int64_t n, int64_t x_row_stride,
                  uint64_t* words, void* stream);

/* Inverse of btx_pack_bits: out[row*out_row_stride + j] = bit j, as 0.0f/1.0f. */
int btx_unpack_bits(const uint64_t* words, int64_t rows, int64_t n, float* out,
                    int64_t out_row_stride, void* stream);

/* ------------------------------------------------------------------------ */
/* Event-driven sparse product (SparseConn forward)                          */
/* ------------------------------------------------------------------------ */

/* Bytes of workspace an event-driven product with `lanes` lanes over
 * `n_dst` destinations needs.  `slots` is the per-destination slot count of
 * the ordered mode (ignored when ordered == 0).  The workspace must be
 * zero-filled once after allocation; the kernels leave it zeroed. */
size_t btx_spmv_workspace_bytes(int64_t lanes, int64_t n_dst, int slots, int ordered);

/* out[lane, :] = sum over active sources s of x[lane, s] * W[s, :], with W
 * stored source-major (CSR: row_ptr[n_src+1] int64, col[E] int32 destination,
 * w[E] fp32, E = nnz).  active/counts are the compaction of x (btx_compact_dense).
 * x may be NULL, meaning every active entry is 1.0 (spikes).
 * csc_ptr[n_dst+1] / csc_src[E] / csc_w[E] are the same matrix destination-major
 * (the reference's coalesced (dst, src) order: indices/magnitude); the ordered
 * mode reduces a destination with more than `slots` terms by pulling over
 * them, so its result never depends on a bounded overflow list.
 * ordered != 0: each destination sums its terms in ascending source order
 * with a fresh fp32 accumulator (fmaf(w, x, acc)), which is bit-identical to
 * torch.sparse.mm on the coalesced (dst, src) COO on CPU -- the reference's
 * native branch, btorch/models/linear.py:474-483.  ordered == 0 accumulates
 * with float atomics (the reference's Triton order-free semantics,
 * btorch/models/_sparse_triton.py:59-96).
 * Replaces triton_sparse_mm (btorch/models/_sparse_triton.py:418-444). */
int btx_spmv_event(const int64_t* row_ptr, const int32_t* col, const float* w,
                   int64_t n_src, int64_t n_dst, int64_t nnz,
                   const int64_t* csc_ptr, const int32_t* csc_src, const float* csc_w,
                   const int32_t* active, const int32_t* counts, int64_t lanes,
                   int64_t active_lane_stride,
                   const float* x, int64_t x_lane_stride,
                   float* out, int ordered, int slots,
                   void* workspace, size_t workspace_bytes, void* stream);

/* btx_spmv_event without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference: its
 * Triton product always takes a workspace.  The scratch is a file-scope
 * device array of btx_spmv_event_noalloc_max_bytes() bytes (32 MiB) that
 * each GPU holds until the process ends; a call whose
 * btx_spmv_workspace_bytes(lanes, n_dst, slots, ordered) exceeds it fails
 * with a message and launches nothing.  The entry point allocates nothing
 * and clears the part of the scratch that must start at zero on `stream`.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (different streams, or host threads) race; serialise them. */
size_t btx_spmv_event_noalloc_max_bytes(void);
int btx_spmv_event_noalloc(const int64_t* row_ptr, const int32_t* col, const float* w,
                           int64_t n_src, int64_t n_dst, int64_t nnz,
                           const int64_t* csc_ptr, const int32_t* csc_src, const float* csc_w,
                           const int32_t* active, const int32_t* counts, int64_t lanes,
                           int64_t active_lane_stride,
                           const float* x, int64_t x_lane_stride,
                           float* out, int ordered, int slots, void* stream);

/* Transposed product for the backward of SparseConn:
 * gx[lane, s] = sum_e w[e] * g[lane, col[e]] over the CSR row s
 * (every row; g is dense).  Replaces the grad_x index_add of
 * btorch/models/_sparse_triton.py:404-410. */
int btx_spmv_transpose(const int64_t* row_ptr, const int32_t* col, const float* w,
                       int64_t n_src, int64_t n_dst, const float* g, int64_t lanes,
                       float* gx, void* stream);

/* Event-driven weight gradient: gw[e] += sum over lanes with x[lane, s] != 0
 * of x[lane, s] * g[lane, col[e]] for every edge e of every active row s.
 * gw is indexed in CSR order.  Replaces the dense grad_weight of
 * btorch/models/_sparse_triton.py:411-414. */
int btx_spmv_weight_grad(const int64_t* row_ptr, const int32_t* col, int64_t n_src,
                         int64_t n_dst, const int32_t* active, const int32_t* counts,
                         int64_t lanes, int64_t active_lane_stride, const float* x,
                         int64_t x_lane_stride, const float* g, float* gw,
                         void* stream);

/* ------------------------------------------------------------------------ */
/* Recorded spikes in coordinate form                                        */
/* ------------------------------------------------------------------------ */

/* counts[i] = set bits of spike word i of rows x ceil(n/64) words (bits of a
 * row's last word past n are not counted). */
int btx_word_popcount(const uint64_t* words, int64_t rows, int64_t n, int64_t* counts,
                      void* stream);

/* The nonzeros of a 0/1 spike array given as words (rows, ceil(n/64)), in
 * row-major order, as coordinates: for the k-th nonzero, row r and neuron j,
 * idx_a[k] = r / row_div, idx_b[k] = r % row_div, idx_n[k] = j.  offsets is
 * the exclusive prefix sum of btx_word_popcount's counts.  With rows = T*B
 * and row_div = B these are the time, batch and neuron index arrays of
 * to_sparse_repr on the dense (T, B, N) spikes
 * (btorch/io/serialization.py:80-147, np.nonzero order). */
int btx_raster_coo(const uint64_t* words, int64_t rows, int64_t n, int64_t row_div,
                   const int64_t* offsets, int32_t* idx_a, int32_t* idx_b, int32_t* idx_n,
                   void* stream);

/* ------------------------------------------------------------------------ */
/* Connectivity built on the device                                          */
/* ------------------------------------------------------------------------ */

/* Workspace of btx_conn_build for nnz edges. */
size_t btx_conn_build_workspace_bytes(int64_t nnz);

/* A delay-expanded connection from nnz edges (pre, post, delay, w), as
 * expand_conn_for_delays (btorch/connectome/connection.py:973-1027) followed
 * by BaseSparseConn's transpose and scipy sum_duplicates
 * (btorch/models/linear.py:276-291) and this library's CSR layout:
 *   virtual source row  vsrc = pre*n_delay_bins + clip(delay, 0, D-1)
 *                       (delay NULL = 0);
 *   out_dst/out_src/out_w  the unique (post, vsrc) pairs in ascending
 *                       (dst, src) order, duplicate weights summed left to
 *                       right in input order; *nnz_out (device) = their count;
 *   row_ptr (n_pre*D + 1), col, perm  source-major CSR, destinations
 *                       ascending per row, perm = position in the coalesced
 *                       order;
 *   csc_ptr (n_post + 1), csc_src  the coalesced order's row pointer and
 *                       sources.
 * Every output array holds nnz entries (the first *nnz_out are valid).
 * Indices must lie in [0, n_pre) and [0, n_post) (not checked on the device;
 * out-of-range input gives wrong output, never a stray access).  Stream
 * ordered; nnz < 2^31, n_pre*D < 2^31. */
int btx_conn_build(const int64_t* pre, const int64_t* post, const int64_t* delay,
                   const float* w, int64_t nnz, int64_t n_pre, int64_t n_post,
                   int n_delay_bins, int64_t* out_dst, int64_t* out_src, float* out_w,
                   int64_t* nnz_out, int64_t* row_ptr, int32_t* col, int64_t* perm,
                   int64_t* csc_ptr, int32_t* csc_src, void* workspace,
                   size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Fused time-step engine                                                    */
/* ------------------------------------------------------------------------ */

/* Neuron models of the fused step. */
#define BTX_NEURON_LIF 0      /* btorch/models/neurons/lif.py:144-187 */
#define BTX_NEURON_GLIF3 1    /* btorch/models/neurons/glif.py:264-342 */
#define BTX_NEURON_TWO_COMPARTMENT 2  /* btorch/models/neurons/two_compartment.py:208-292 */
#define BTX_NEURON_MIXED 3    /* MixedNeuronPopulation, btorch/models/neurons/mixed.py:16-193 */

/* Per-neuron model code of a BTX_NEURON_MIXED plan (plan.neuron_kinds):
 * the low bits name the neuron's model (BTX_NEURON_LIF, _GLIF3 or
 * _TWO_COMPARTMENT), BTX_MIXED_HARD_RESET selects the hard reset of a LIF
 * or GLIF3 neuron (the two-compartment neuron always resets hard). */
#define BTX_MIXED_MODEL_MASK 3
#define BTX_MIXED_HARD_RESET 4

/* Synapse kinetics of the fused step (plan.syn_kind). */
#define BTX_SYN_EXP 0          /* ExponentialPSC, btorch/models/synapse.py:138-187 */
#define BTX_SYN_ALPHA 1        /* AlphaPSC, synapse.py:274-334 */
#define BTX_SYN_ALPHA_BILLEH 2 /* AlphaPSCBilleh, synapse.py:201-268 */
#define BTX_SYN_DUAL_EXP 3     /* DualExponentialPSC, synapse.py:337-418 */

/* Coefficients of the two-compartment neuron, indices into plan.tc[]. */
#define BTX_TC_DA 0        /* dt / max(tau_a, 1e-4) */
#define BTX_TC_DA1 1       /* 1 + DA */
#define BTX_TC_DS 2        /* dt / max(tau_s, 1e-4) */
#define BTX_TC_DS1 3       /* 1 + DS */
#define BTX_TC_DTH1 4      /* 1 + dt / max(tau_th, 1e-4) */
#define BTX_TC_R_S 5       /* max(R_s, 1e-4) */
#define BTX_TC_E_L 6       /* E_L */
#define BTX_TC_DT 7        /* max(delta_T, 0) */
#define BTX_TC_DTC 8       /* max(DT, 1e-4) */
#define BTX_TC_W_CA 9      /* w_Ca */
#define BTX_TC_THETA_CA 10 /* theta_Ca */
#define BTX_TC_W_SA 11     /* w_sa */
#define BTX_TC_W_AS 12     /* w_as */
#define BTX_TC_DELTA_TH 13 /* delta_th */
#define BTX_TC_NPARAM 14

/* Surrogate functions (forward Heaviside, backward derivative). */
#define BTX_SURR_SIGMOID 0    /* btorch/models/surrogate/sigmoid.py:16-29 */
#define BTX_SURR_SUPERSPIKE 1 /* btorch/models/surrogate/superspike.py:16-31 */
#define BTX_SURR_ERF 2        /* btorch/models/surrogate/erf.py:22-28 */

/* One recurrent population: neuron (LIF) + exponential current synapse +
 * sparse recurrent connectivity, optionally through an integer-delay ring.
 * Follows RecurrentNN.single_step_forward (btorch/models/rnn.py:517-528):
 *   I_t   = psc_{t-1} + x_t
 *   z_t   = neuron(I_t)                         (charge -> fire -> reset)
 *   psc_t = psc_{t-1} + k*((-psc_{t-1})/tau_s)  (exp-Euler decay)
 *           + W_exp^T * flatten(history after pushing z_t)
 * Each step, the fused kernel writes the spikes it pushes (its own, or the
 * teacher's) as segmented ascending lists into the ring slot of this step;
 * the push reads the lists of the ring slot each delay bin points at and
 * streams that bin's virtual rows.
 * Ordered mode: the push bins every term (virtual row, weight) by
 * (lane, destination tile) with one atomic reservation per workgroup and
 * tile; a reduction kernel then sorts each tile's terms in LDS by
 * destination and, per destination, by virtual row, and adds the ascending
 * sum to psc.  Destinations with many terms, and tiles whose terms overflow
 * the bin, pull their input over the destination-major edges instead (same
 * order, same bits).  Fast mode adds every term to psc with a float atomic.
 * Per-neuron parameters are fp32 arrays with a stride (0 = one shared value).
 * Every pointer refers to device memory that outlives the enqueued work. */
typedef struct btx_rsnn_plan {
    int64_t n;            /* neurons */
    int64_t lanes;        /* batch lanes */
    int neuron_model;     /* BTX_NEURON_* */
    int hard_reset;       /* 0 = soft (v - (vth - vr) s), 1 = hard (v - (v - vr) s) */
    int use_refractory;   /* 1 = refractory counter enabled */
    int ordered;          /* 1 = ascending-source bit-exact accumulation */
    int slots;            /* ordered mode: a destination with more terms in one
                           * step pulls its input (1..32) */
    float dt;             /* environ dt as fp32 */
    /* neuron parameters, (n,) or stride 0 */
    const float* tau;      int64_t tau_stride;
    const float* v_th;     int64_t v_th_stride;
    const float* v_reset;  int64_t v_reset_stride;
    const float* c_m;      int64_t c_m_stride;
    const float* tau_ref;  int64_t tau_ref_stride;
    /* synapse: psc_dec = psc + k*((-psc)/tau_s) ; k = expm1(dt*l)/l, l=-1/tau_s */
    const float* tau_syn;  int64_t tau_syn_stride;
    const float* k_syn;    int64_t k_syn_stride;
    /* state, (lanes, n) contiguous, updated in place */
    float* v;
    float* refractory;    /* may be NULL when use_refractory == 0 */
    float* psc;
    /* external input x[t, lane, n], NULL = none; x_step_stride elements per step */
    const float* x;       int64_t x_step_stride;
    /* in-kernel Poisson drive, added to x: neuron n of lane b receives in_kick
     * at its local step s = in_t0[b] + t when its uniform falls below
     * in_prob[b]; NULL in_prob = none.  Uniforms come from a SplitMix64
     * stream per (lane, step): stream = mix64(in_seed[b] ^ s*0x100000001b3),
     * h_q = mix64(stream + (q+1)*0x9e3779b97f4a7c15), neuron 2q takes bits
     * 40..63 of h_q and neuron 2q+1 bits 16..39, each / 2^24.  The drive of a
     * lane depends only on its own seed and clock, so a job gives the same
     * result alone or packed with others. */
    const uint64_t* in_seed;
    const float* in_prob;
    const int64_t* in_t0;
    float in_kick;
    /* connectivity: CSR over virtual rows n*csr_d + bin (int64 row_ptr) */
    const int64_t* row_ptr;
    const int32_t* col;
    const float* w;
    int64_t nnz;          /* edges (sizes the push group width) */
    int csr_d;            /* virtual rows per neuron (delay-expanded bins) */
    /* delay ring of compacted spike lists: ring_d slots x lanes x n, in
     * segments of BTX_RING_CHUNK neurons (segment c of a lane: entries
     * [c*BTX_RING_CHUNK, c*BTX_RING_CHUNK + count)) */
    int ring_d;
    int32_t* ring_idx;    /* [ring_d][lanes][n] */
    int32_t* ring_cnt;    /* [ring_d][lanes][ceil(n/BTX_RING_CHUNK)] */
    int ring_cursor;      /* ring slot that receives z of the first step */
    int nbins;            /* delay bins read each step */
    int bin_delay[BTX_MAX_BINS];  /* ring age of bin j (0 = this step) */
    int bin_sub[BTX_MAX_BINS];    /* virtual-row offset of bin j */
    /* destination-major copy of the connectivity (the reference's coalesced
     * (dst, src) order: indices/magnitude); with it the ordered mode reduces
     * a destination with many terms exactly, by pulling its incoming edges */
    const int64_t* csc_ptr;   /* (n+1) */
    const int32_t* csc_src;   /* (nnz) virtual source rows, ascending per destination */
    const float* csc_w;       /* (nnz) */
    /* spike words pushed into the ring, [ring_d][lanes][ceil(n/64)] in
     * ring-slot order: on entry slot (cursor - 1 - k) mod ring_d holds the
     * step k + 1 steps before the run; the run keeps them current (each
     * step writes its slot), so they carry over to the next run */
    uint64_t* ring_words;
    /* teacher forcing: spike words [T][lanes][ceil(n/64)] pushed into the ring
     * instead of the produced spikes (NULL = free running) */
    const uint64_t* teacher;
    /* outputs */
    uint64_t* raster;     /* [T][lanes][ceil(n/64)] spike bits, required */
    float* rec_z;         /* [T][lanes][n] or NULL */
    float* rec_v;         /* [T][lanes][n] or NULL */
    float* rec_refractory;/* [T][lanes][n] or NULL */
    float* rec_psc;       /* [T][lanes][n] or NULL */
    float* rec_v_pre;     /* [T][lanes][n] pre-reset membrane or NULL */
    float* rec_v_prev;    /* [T][lanes][n] membrane entering the step or NULL */
    int32_t* counts;      /* [lanes][n] spike counts, accumulated, or NULL */
    float* rec_sample;    /* [T][lanes][n_sample] membrane of neurons 0..n_sample-1, or NULL */
    int64_t n_sample;
    /* per-lane end of trial (local step, with in_t0): from there on the
     * lane is idle -- no counts, samples, state update or push, and its state
     * is left as it was (a caller resets the lane before reusing it); NULL =
     * no end */
    const int64_t* lane_t_end;
    /* GLIF3 (neuron_model == BTX_NEURON_GLIF3, glif.py:264-342), in the order
     * charge -> adaptation -> fire -> reset of base.py:1108-1116:
     *   v    += k_v * ((-(v - v_rest))/tau + (I + sum_j Iasc_j)/c_m)
     *   Iasc_j += asc_kk_j * ((-asc_k_j) * Iasc_j)
     *   fire and reset as LIF; Iasc_j += asc_amp_j * s
     * with the exponential-Euler coefficients k_v = expm1(dt*l)/l, l = -1/tau,
     * and asc_kk = expm1(dt*(-asc_k))/(-asc_k) evaluated by the caller
     * (ode.py:20-50); sum_j runs from +0.0f in ascending j. */
    int n_asc;                 /* after-spike currents per neuron, 1..4 */
    const float* v_rest;   int64_t v_rest_stride;
    const float* k_v;      int64_t k_v_stride;
    const float* asc_k;        /* (n, n_asc) */
    const float* asc_kk;       /* (n, n_asc) */
    const float* asc_amp;      /* (n, n_asc) */
    float* iasc;               /* state (lanes, n, n_asc), updated in place */
    /* 1 = build the ring lists from ring_words before the first step;
     * 0 = the ring lists are already current (a resident ring carried over
     * from the previous run) */
    int ring_rebuild;
    /* ordered mode: terms per (lane, destination segment) bin, 0 = the
     * maximum (8192); a bin that receives more is reduced by pulling.  Bins
     * of up to 1024 terms are sorted in LDS, larger ones through an order
     * array in the workspace */
    int bin_cap;
    /* two-compartment GLIF (neuron_model == BTX_NEURON_TWO_COMPARTMENT,
     * TwoCompartmentGLIF.single_step_forward, two_compartment.py:208-292),
     * somatic input I = psc + x, apical input A = x_apical (0 when NULL):
     *   i_a'  = (i_a + DA*((A + W_CA*sigmoid(i_a - THETA_CA)) + i_bap)) / DA1
     *   thr   = v_th + theta_th
     *   ex    = DT > 0 ? DT*exp(min((v - thr)/DTC, 10)) : 0
     *   v_pre = (v + DS*((E_L + R_S*(I + W_AS*i_a')) + ex)) / DS1
     *   s     = (v_pre - thr) >= 0 ; v = v_pre - (v_pre - v_reset)*s
     *   i_bap = W_SA*s ; theta_th = theta_th/DTH1 + DELTA_TH*s
     * with the coefficients BTX_TC_* evaluated by the caller in the
     * reference's ops.  exp and the sigmoid are evaluated on the device, so
     * a neuron with W_CA != 0 or DT > 0 agrees with torch to rounding, every
     * other neuron bit for bit.  use_refractory must be 0. */
    const float* x_apical;  int64_t x_apical_step_stride;
    float* i_a;           /* state (lanes, n), updated in place */
    float* i_bap;
    float* theta_th;
    const float* tc[BTX_TC_NPARAM];      /* (n,) or one value (stride 0) */
    int64_t tc_stride[BTX_TC_NPARAM];
    float* rec_i_a;       /* [T][lanes][n] or NULL */
    float* rec_i_bap;     /* [T][lanes][n] or NULL */
    float* rec_theta_th;  /* [T][lanes][n] or NULL */
    /* synapse kinetics (BTX_SYN_*), with W z the step's summed push; the
     * coefficients are (n,) or one value (stride 0), k = expm1(dt*l)/l with
     * l = -1/tau evaluated by the caller:
     *   EXP        psc = psc + k_syn*((-psc)/tau_syn) ; psc = psc + W z
     *   ALPHA      psc = psc + k_syn*((-psc)/tau_syn + h/tau_syn) ;
     *              h = (h + k_syn*((-h)/tau_syn)) + syn_gain*(W z)
     *   ALPHA_BILLEH  psc = c*psc + c*h ; h = c*h + syn_gain*(W z),
     *              c = k_syn (exp(-1/tau)), syn_gain = e/tau
     *   DUAL_EXP   g_r = g_r + k_syn*((-g_r)/tau_syn) ;
     *              g_d = g_d + k_syn2*((-g_d)/tau_syn2) ;
     *              g_r = g_r + W z ; g_d = g_d + W z ; psc = syn_a*(g_d - g_r)
     * h (or g_r) is syn_h1, g_d is syn_h2: state (lanes, n), updated in
     * place; rec_syn_h1 / rec_syn_h2 record them like rec_psc.  The neuron
     * input of a step is the psc the synapse left the step before.  LIF and
     * GLIF3 neurons run with every kind, the two-compartment neuron with EXP. */
    int syn_kind;
    float* syn_h1;
    float* syn_h2;
    const float* syn_gain;  int64_t syn_gain_stride;
    const float* tau_syn2;  int64_t tau_syn2_stride;
    const float* k_syn2;    int64_t k_syn2_stride;
    const float* syn_a;     int64_t syn_a_stride;
    float* rec_syn_h1;    /* [T][lanes][n] or NULL */
    float* rec_syn_h2;    /* [T][lanes][n] or NULL */
    /* workspace from btx_rsnn_workspace_bytes(), zero-filled once */
    void* workspace;      size_t workspace_bytes;
    /* 1 = the workspace already holds this plan's delay-bin tables (an
     * earlier run with the same workspace, nbins, bin_delay and bin_sub
     * wrote them), so the run skips writing them; 0 = write them first */
    int tables_ready;
    /* BTX_NEURON_MIXED (MixedNeuronPopulation.single_step_forward,
     * mixed.py:113-141: each group's own step on its range of the axis):
     * neuron j runs the model neuron_kinds[j] names, with every parameter
     * given per neuron (stride 1, or 0 for a value all groups share) and
     * every state array over the whole axis; mixed_models is the set of
     * models present (bit 1 << BTX_NEURON_*).  A plan with the
     * two-compartment model gives i_a, i_bap, theta_th and every tc
     * coefficient (any value for the other neurons); one with GLIF3 gives
     * n_asc, iasc, v_rest, k_v and the after-spike parameters (zeros for
     * the other neurons).  Neurons of a LIF or GLIF3 group without a
     * refractory period run with tau_ref 0 and a zero counter, which leaves
     * their update unchanged.  Exponential synapse only. */
    const uint8_t* neuron_kinds;  /* (n,) */
    int mixed_models;
    /* ordered mode, optional (NULL = off): the ring of lane bits, uint64
     * [ring_d + 1][n][ceil(lanes/64)] kept beside ring_words -- bit b of word
     * (slot, j, b/64) set when lane b pushed neuron j into that slot.  With
     * it, a step whose push would bin more than bin_cap/2 terms per (lane,
     * segment) on average (more than 1024, the terms a bin sorts in LDS, when
     * lanes x segments < 256) is delivered as a dense pull instead: every
     * destination sums its active incoming edges in the coalesced order
     * (the same terms and order, so the same bits).  The engine keeps it:
     * while activity nears the pull, each step marks the slot it writes, and
     * ring_rebuild rebuilds every slot from ring_words; a caller that empties
     * lanes of the ring clears their bits too.  The decision is made one
     * step ahead from the ring's spike totals (the newest slot estimated by
     * the one before).  Batches of more than 32 lanes pull in a launch of
     * their own queued after each step's (it ends at once on a push step). */
    uint64_t* ring_lane_bits;
} btx_rsnn_plan;

size_t btx_rsnn_workspace_bytes(int64_t lanes, int64_t n, int slots, int ordered);

/* The in-kernel Poisson drive of btx_rsnn_plan as a dense input tensor:
 * x[t, b, j] = kick when the uniform of neuron j of lane b at local step
 * t0[b] + t falls below prob[b], else 0 (the same SplitMix64 numbers the
 * fused step draws), for t in [0, T).  A run given this x and no drive
 * produces the same bits as a run given the drive.  Input generator of the
 * benchmark workloads (Poisson background as in btorch/datasets/noise.py);
 * it has no counterpart on the reference's step path. */
int btx_poisson_fill(const uint64_t* seed, const float* prob, const int64_t* t0, float kick,
                     int64_t T, int64_t lanes, int64_t n, float* x, void* stream);

/* Run steps [0, n_steps) of the plan.  Records index t in [0, n_steps).
 * State (v, refractory, psc, ring) is final on return (stream order). */
int btx_rsnn_run(const btx_rsnn_plan* plan, int64_t n_steps, void* stream);

/* One neuron step without synapse (LIF.single_step_forward, lif.py:144-187,
 * base.py:1108-1116; GLIF3 and the two-compartment neuron likewise): reads
 * the plan's neuron parameters and state, input x[lanes, n] (and, for the
 * two-compartment neuron, x_apical[lanes, n] of the plan); writes spikes
 * z[lanes, n] (fp32 0/1) and, if v_pre is not NULL, the membrane before
 * reset.  Synapse, ring and connectivity fields of the plan are ignored. */
int btx_lif_step(const btx_rsnn_plan* plan, const float* x, float* z, float* v_pre,
                 void* stream);

/* psc[lanes, n] += k*((-psc)/tau_s) in place: ExponentialPSC.conductance_charge
 * (btorch/models/synapse.py:171-178 with exp_euler_step, btorch/models/ode.py:20-50).
 * k = expm1(dt*l)/l with l = -1/tau_s, evaluated by the caller. */
int btx_exp_decay(float* psc, int64_t lanes, int64_t n, const float* tau_syn,
                  int64_t tau_syn_stride, const float* k_syn, int64_t k_syn_stride,
                  void* stream);

/* The decay part of one step of an alpha or dual-exponential synapse over
 * plan.lanes x plan.n (plan.syn_kind, psc, syn_h1, syn_h2, tau_syn, k_syn,
 * tau_syn2, k_syn2 as in btx_rsnn_plan): ALPHA psc and h decay,
 * ALPHA_BILLEH psc = c*psc + c*h and h = c*h, DUAL_EXP both conductances
 * decay; the caller then adds the push (AlphaPSC / AlphaPSCBilleh /
 * DualExponentialPSC.conductance_charge and the decay of adaptation_charge,
 * btorch/models/synapse.py:239-253, 304-317, 390-406). */
int btx_syn_decay(const btx_rsnn_plan* plan, void* stream);

/* btx_rsnn_run with a HIP event pair around every kernel launch of the
 * step sequence, recorded on `stream`.  On return (after a stream
 * synchronisation) ms_out[k] holds the summed duration of kernel kind k and
 * calls_out[k] its launch count, k = 0 fused step (with compaction),
 * 1 ordered reduction, 2 push, 3 final psc.  If per_launch_ms is not NULL, per_launch_ms[k*per_launch_cap + i]
 * receives the duration of the i-th launch of kind k (i < per_launch_cap).
 * For measurement only: the events serialise the launches. */
int btx_rsnn_run_timed(const btx_rsnn_plan* plan, int64_t n_steps, void* stream,
                       double* ms_out, int64_t* calls_out, double* per_launch_ms,
                       int64_t per_launch_cap);

/* ------------------------------------------------------------------------ */
/* Surrogate-gradient adjoint (BPTT), one step backward per call             */
/* ------------------------------------------------------------------------ */
/* All arrays of one step are neuron-major (n, lanes): element (j, b) at
 * j*lanes + b.  Replaces the autograd chain of the reference's eager step
 * (lif.py:144-187, surrogate/base.py:21-25, synapse.py:171-178). */
typedef struct btx_adj_lif_args {
    int64_t n;
    int64_t lanes;
    float dt;
    int hard_reset;
    int detach_reset;
    int surrogate;        /* BTX_SURR_* */
    float alpha;
    float damping;
    const float* tau;      int64_t tau_stride;
    const float* v_th;     int64_t v_th_stride;
    const float* v_reset;  int64_t v_reset_stride;
    const float* c_m;      int64_t c_m_stride;
    const float* tau_syn;  int64_t tau_syn_stride;
    const float* k_syn;    int64_t k_syn_stride;
    const float* v_prev;   /* membrane entering the step */
    const float* v_pre;    /* membrane before reset */
    const float* z;        /* spikes of the step */
    const float* ref_prev; /* refractory entering the step, NULL = none */
    const float* gz;       /* dL/dz of the step */
    const float* gsyn;     /* (W dL/dpsc')[j]: from btx_adj_transpose */
    float* gv;             /* in dL/dv after the step, out dL/dv before it */
    float* gpsc;           /* in dL/dpsc after the step, out before it */
    float* gx;             /* out dL/dx of the step */
    float* gtau;           /* accumulated dL/dtau ((n,) or 1 when tau_stride = 0), or NULL */
} btx_adj_lif_args;

int btx_adj_lif_step(const btx_adj_lif_args* args, void* stream);

/* The reverse sweep of a K-step chunk of an undelayed LIF + exponential
 * synapse in one call (the backward of the reference's checkpointed chunk,
 * rnn.py:177-197, through the eager step's autograd chain, lif.py:144-187,
 * synapse.py:171-178): for t = K-1 .. 0, btx_adj_transpose of args->gpsc into
 * args->gsyn, btx_adj_wgrad of masks[t] (n words) into gw, then
 * btx_adj_lif_step.  v_prev, v_pre, z, ref_prev, gz and gx of args are the
 * bases of (K, n, lanes) arrays; gv, gpsc and gtau carry across the steps. */
int btx_adj_lif_chunk(const btx_adj_lif_args* args, int64_t K, const int64_t* row_ptr,
                      const int32_t* col, const float* w, const uint64_t* masks, float* gw,
                      void* stream);

/* g_src[j, :] = sum over CSR row j of w[e] * g_dst[col[e], :] (every row;
 * (n, lanes) layout).  The transposed product of the backward pass. */
int btx_adj_transpose(const int64_t* row_ptr, const int32_t* col, const float* w,
                      int64_t n_rows, int64_t lanes, const float* g_dst, float* g_src,
                      void* stream);

/* gw[e] += sum over lanes b with bit b of masks[j] set of g_dst[col[e], b] for
 * every edge e of row j (lanes <= 64; masks[j] = spike bits of neuron j over
 * lanes).  Event-driven weight gradient: rows with an empty mask cost one
 * load.  gw is in CSR order. */
int btx_adj_wgrad(const int64_t* row_ptr, const int32_t* col, int64_t n_rows, int64_t lanes,
                  const uint64_t* masks, const float* g_dst, float* gw, void* stream);

/* Reverse delay ring of delayed synapses (HeterSynapsePSC,
 * btorch/models/synapse.py:626-761 with SpikeHistory.get_flattened,
 * history.py:299-345; DelayedPSC, synapse.py:537-623): virtual row
 * j*csr_d + sub carries the spike of neuron j age(sub) steps late, so
 *   g_src[j, :] (+)= sum_sub sum_{e in row j*csr_d+sub} w[e] * G[tau + age(sub)][col[e], :]
 * over the sub with 0 <= tau + age(sub) < K; G is dL/dpsc of the K steps of a
 * window, [K][n][lanes]; ages[csr_d] device ints (-1 = row offset unused). */
int btx_adj_transpose_delayed(const int64_t* row_ptr, const int32_t* col, const float* w,
                              int64_t n, int csr_d, int64_t lanes, const float* G, int64_t K,
                              int64_t tau, const int* ages, float* g_src, int accumulate,
                              void* stream);

/* Event-driven weight gradient of the delayed product: for every neuron j
 * with a nonzero lane mask (masks[j], bit b = lane b spiked at window step
 * tau; lanes <= 64), gw[e] += sum_b G[tau + age(sub)][col[e], b] over the
 * edges e of its virtual rows whose step lies in [0, K). */
int btx_adj_wgrad_delayed(const int64_t* row_ptr, const int32_t* col, int64_t n, int csr_d,
                          int64_t lanes, const uint64_t* masks, const float* G, int64_t K,
                          int64_t tau, const int* ages, float* gw, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* BTORCH_AMD_H */
