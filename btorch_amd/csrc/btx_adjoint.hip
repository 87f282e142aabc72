// Surrogate-gradient adjoint (BPTT) of the fused LIF + exponential-synapse
// + sparse-recurrent step, run backward in time one step per call.
//
// Layout: neuron-major (n, lanes) -- element (j, b) at j*lanes + b -- so the
// lanes of one neuron are contiguous and a wavefront reads one destination's
// 64 lanes with a single 256-byte access.
//
// Forward step (btorch/models/rnn.py:517-528, lif.py:144-187, synapse.py:171-178):
//   v1  = v + dt*((-(v - vr))/tau + (psc + x)/cm)
//   s   = H((v1 - vth)/(vth - vr)) * (ref == 0)
//   v'  = v1 - (vth - vr)*s            (hard: v1 - (v1 - vr)*s)
//   psc' = psc*(1 - k/tau_s) + W^T s
// Reverse step, given dL/dv', dL/dpsc', dL/dz:
//   gs   = gz + (W gpsc')[j] + gv' * d(v')/ds        (reset term unless detached)
//   gv1  = gv' * d(v')/dv1 + gs * mask * g(u) / (vth - vr)
//   gv   = gv1 * (1 - dt/tau),  gx = gv1 * dt/cm,  gpsc = gx + gpsc' * (1 - k/tau_s)
//   gtau += gv1 * dt * (v - vr)/tau^2,  gW[e] += s[src] * gpsc'[dst]
// with g the surrogate derivative (surrogate/sigmoid.py:22-29,
// superspike.py:25-31, erf.py:22-28) and the refractory mask detached as in
// lif.py:165-167.

#include <hip/hip_runtime.h>

#include "../../include/btorch_amd.h"
#include "btx_common.h"

namespace btx {

struct AdjArgs {
    int64_t n;
    int64_t lanes;
    float dt;
    int hard_reset;
    int detach_reset;
    int surrogate;
    float alpha;
    float damping;
    const float* tau;   int64_t tau_s;
    const float* vth;   int64_t vth_s;
    const float* vr;    int64_t vr_s;
    const float* cm;    int64_t cm_s;
    const float* tsyn;  int64_t tsyn_s;
    const float* ksyn;  int64_t ksyn_s;
    const float* v_prev;
    const float* v_pre;
    const float* z;
    const float* ref_prev;
    const float* gz;
    const float* gsyn;
    float* gv;
    float* gpsc;
    float* gx;
    float* gtau;
};

__device__ __forceinline__ float surrogate_grad(int kind, float u, float alpha, float damping) {
    if (kind == BTX_SURR_SIGMOID) {
        const float k = 1.7627471740390860f;  // 2*ln(sqrt(2)+1)
        const float sg = 1.f / (1.f + expf(-k * alpha * u));
        return damping * 4.f * sg * (1.f - sg);
    }
    if (kind == BTX_SURR_SUPERSPIKE) {
        const float k = 0.41421356237309515f;  // sqrt(2)-1
        const float d = 1.f + k * alpha * fabsf(u);
        return damping / (d * d);
    }
    const float au = alpha * u;  // BTX_SURR_ERF
    return damping * expf(-0.6931471805599453f * au * au);
}

__global__ __launch_bounds__(256) void k_adj_lif(AdjArgs a) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t m = a.n * a.lanes;
    if (i >= m) return;
    const int64_t j = i / a.lanes;
    const float tau = a.tau[a.tau_s * j];
    const float vth = a.vth[a.vth_s * j];
    const float vr = a.vr[a.vr_s * j];
    const float cm = a.cm[a.cm_s * j];
    const float v = a.v_prev[i];
    const float v1 = a.v_pre[i];
    const float s = a.z[i];
    const float mask = a.ref_prev ? (a.ref_prev[i] == 0.f ? 1.f : 0.f) : 1.f;
    const float gvn = a.gv[i];
    const float gpn = a.gpsc[i];
    float gs = a.gz[i] + a.gsyn[i];
    float dvdv1 = 1.f;
    if (a.hard_reset) {
        dvdv1 = 1.f - s;
        if (!a.detach_reset) gs += gvn * (-(v1 - vr));
    } else if (!a.detach_reset) {
        gs += gvn * (-(vth - vr));
    }
    const float u = (v1 - vth) / (vth - vr);
    const float gu = gs * mask * surrogate_grad(a.surrogate, u, a.alpha, a.damping);
    const float gv1 = gvn * dvdv1 + gu / (vth - vr);
    a.gv[i] = gv1 * (1.f - a.dt / tau);
    const float gI = gv1 * (a.dt / cm);
    a.gx[i] = gI;
    const float k = a.ksyn[a.ksyn_s * j];
    const float ts = a.tsyn[a.tsyn_s * j];
    a.gpsc[i] = gI + gpn * (1.f - k / ts);
    if (a.gtau) {
        const float gt = gv1 * a.dt * (v - vr) / (tau * tau);
        atomicAdd(&a.gtau[a.tau_s * j], gt);
    }
}

// g_src[j, b] = sum over edges e of row j of w[e] * g_dst[col[e], b]; one
// wavefront per row, lane = batch lane (rows of more than 64 lanes loop).
// The row's edges are read 64 at a time, lane k holding edge c + k, and
// handed to every lane by readlane, so the g_dst loads of a chunk do not wait
// on an index load each.  Sums: even edge offsets of the row into acc0, odd
// ones into acc1, each in edge order, then acc0 + acc1.
__global__ __launch_bounds__(256) void k_adj_transpose(const int64_t* __restrict__ rp,
                                                       const int32_t* __restrict__ col,
                                                       const float* __restrict__ w,
                                                       int64_t n_rows, int64_t lanes,
                                                       const float* __restrict__ g_dst,
                                                       float* __restrict__ g_src) {
    // uniform across the wavefront (four rows per workgroup)
    const int64_t row =
        __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4u + (threadIdx.x >> 6)));
    const int ln = threadIdx.x & 63;
    if (row >= n_rows) return;
    const int64_t e0 = rp[row], e1 = rp[row + 1];
    for (int64_t b0 = 0; b0 < lanes; b0 += 64) {
        const int64_t b = b0 + ln;
        const bool on = b < lanes;
        // lanes past `lanes` read column 0 of the row (never stored)
        const float* gb = g_dst + (on ? b : 0);
        float acc0 = 0.f, acc1 = 0.f;
        for (int64_t c = e0; c < e1; c += 64) {
            const int m = (int)(e1 - c < 64 ? e1 - c : 64);
            const int my_d = ln < m ? col[c + ln] : 0;
            const int my_w = ln < m ? __float_as_int(w[c + ln]) : 0;
            int k = 0;
            for (; k + 8 <= m; k += 8) {
                // eight loads in flight, then the sums in edge order
                float g8[8], w8[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int64_t d = __builtin_amdgcn_readlane(my_d, k + j);
                    w8[j] = __int_as_float(__builtin_amdgcn_readlane(my_w, k + j));
                    g8[j] = gb[d * lanes];
                }
#pragma unroll
                for (int j = 0; j < 8; j += 2) {
                    acc0 += w8[j] * g8[j];
                    acc1 += w8[j + 1] * g8[j + 1];
                }
            }
            for (; k + 2 <= m; k += 2) {
                const int64_t d0 = __builtin_amdgcn_readlane(my_d, k);
                const int64_t d1 = __builtin_amdgcn_readlane(my_d, k + 1);
                const float w0 = __int_as_float(__builtin_amdgcn_readlane(my_w, k));
                const float w1 = __int_as_float(__builtin_amdgcn_readlane(my_w, k + 1));
                const float g0 = gb[d0 * lanes], g1 = gb[d1 * lanes];
                acc0 += w0 * g0;
                acc1 += w1 * g1;
            }
            if (k < m) {
                const int64_t d0 = __builtin_amdgcn_readlane(my_d, k);
                const float w0 = __int_as_float(__builtin_amdgcn_readlane(my_w, k));
                acc0 += w0 * gb[d0 * lanes];
            }
        }
        if (on) g_src[row * lanes + b] = acc0 + acc1;
    }
}

// gw[e] += sum over lanes b spiking at neuron j of g_dst[col[e], b], for
// every edge of every neuron with a nonzero spike mask (lanes <= 64, mask
// bit b = lane b), the lanes taken in ascending order.  One wavefront per
// neuron.  A neuron that spiked in a few lanes: lane k of the wavefront
// takes edge c + k and gathers its destinations' values lane by lane.  In
// more lanes: the g_dst rows of 32 edges are staged in the wavefront's LDS
// tile with 256-byte loads (lane = batch lane), and lane k then sums row k
// of the tile over the set bits.
constexpr int kWgradTile = 32;
constexpr int kWgradFew = 2;  // set bits up to which a lane gathers directly

__global__ __launch_bounds__(256) void k_adj_wgrad(const int64_t* __restrict__ rp,
                                                   const int32_t* __restrict__ col,
                                                   int64_t n_rows, int64_t lanes,
                                                   const uint64_t* __restrict__ masks,
                                                   const float* __restrict__ g_dst,
                                                   float* __restrict__ gw) {
    __shared__ float tile[4][kWgradTile * 65];
    const int wv = threadIdx.x >> 6;
    // uniform across the wavefront (four rows per workgroup)
    const int64_t row = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4u + (unsigned)wv));
    const int ln = threadIdx.x & 63;
    if (row >= n_rows) return;
    const uint64_t mk = masks[row];
    if (mk == 0) return;
    const int64_t e0 = rp[row], e1 = rp[row + 1];
    if (__popcll((unsigned long long)mk) <= kWgradFew) {
        for (int64_t e = e0 + ln; e < e1; e += 64) {
            const int64_t d = col[e];
            float acc = 0.f;
            uint64_t q = mk;
            while (q) {
                const int b = __ffsll((unsigned long long)q) - 1;
                acc += g_dst[d * lanes + b];
                q &= q - 1;
            }
            gw[e] += acc;
        }
        return;
    }
    float* t = tile[wv];
    const bool on = ln < lanes;
    for (int64_t c = e0; c < e1; c += kWgradTile) {
        const int m = (int)(e1 - c < kWgradTile ? e1 - c : kWgradTile);
        const int my_d = ln < m ? col[c + ln] : 0;
        int k = 0;
        for (; k + 8 <= m; k += 8) {
            float g8[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int64_t d = __builtin_amdgcn_readlane(my_d, k + j);
                g8[j] = on ? g_dst[d * lanes + ln] : 0.f;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) t[(k + j) * 65 + ln] = g8[j];
        }
        for (; k < m; ++k) {
            const int64_t d = __builtin_amdgcn_readlane(my_d, k);
            t[k * 65 + ln] = on ? g_dst[d * lanes + ln] : 0.f;
        }
        // the tile is the wavefront's own: its LDS writes land before its
        // reads, the fences keep the compiler from moving them across
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (ln < m) {
            float acc = 0.f;
            uint64_t q = mk;
            while (q) {
                const int b = __ffsll((unsigned long long)q) - 1;
                acc += t[ln * 65 + b];
                q &= q - 1;
            }
            gw[c + ln] += acc;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// k_adj_transpose and k_adj_wgrad of one reverse step in one pass (lanes <=
// 64): each destination row of g_dst is loaded once per edge, feeds the
// transposed sum of its lane and, when the neuron spiked in some lane, the
// wavefront's LDS tile for the weight gradient.  Same sums in the same order
// as the two kernels.
__global__ __launch_bounds__(256) void k_adj_transpose_wgrad(
    const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
    const float* __restrict__ w, int64_t n_rows, int64_t lanes,
    const uint64_t* __restrict__ masks, const float* __restrict__ g_dst,
    float* __restrict__ g_src, float* __restrict__ gw) {
    __shared__ float tile[4][kWgradTile * 65];
    const int wv = threadIdx.x >> 6;
    const int64_t row = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4u + (unsigned)wv));
    const int ln = threadIdx.x & 63;
    if (row >= n_rows) return;
    const uint64_t mk = masks[row];
    const int64_t e0 = rp[row], e1 = rp[row + 1];
    const bool on = ln < lanes;
    // lanes past `lanes` read column 0 (never stored, never summed)
    const float* gb = g_dst + (on ? ln : 0);
    float* t = tile[wv];
    float acc0 = 0.f, acc1 = 0.f;
    for (int64_t c = e0; c < e1; c += kWgradTile) {
        const int m = (int)(e1 - c < kWgradTile ? e1 - c : kWgradTile);
        const int my_d = ln < m ? col[c + ln] : 0;
        const int my_w = ln < m ? __float_as_int(w[c + ln]) : 0;
        int k = 0;
        for (; k + 8 <= m; k += 8) {
            float g8[8], w8[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int64_t d = __builtin_amdgcn_readlane(my_d, k + j);
                w8[j] = __int_as_float(__builtin_amdgcn_readlane(my_w, k + j));
                g8[j] = gb[d * lanes];
            }
#pragma unroll
            for (int j = 0; j < 8; j += 2) {
                acc0 += w8[j] * g8[j];
                acc1 += w8[j + 1] * g8[j + 1];
            }
            if (mk) {
#pragma unroll
                for (int j = 0; j < 8; ++j) t[(k + j) * 65 + ln] = g8[j];
            }
        }
        for (; k + 2 <= m; k += 2) {
            const int64_t d0 = __builtin_amdgcn_readlane(my_d, k);
            const int64_t d1 = __builtin_amdgcn_readlane(my_d, k + 1);
            const float w0 = __int_as_float(__builtin_amdgcn_readlane(my_w, k));
            const float w1 = __int_as_float(__builtin_amdgcn_readlane(my_w, k + 1));
            const float g0 = gb[d0 * lanes], g1 = gb[d1 * lanes];
            acc0 += w0 * g0;
            acc1 += w1 * g1;
            if (mk) {
                t[k * 65 + ln] = g0;
                t[(k + 1) * 65 + ln] = g1;
            }
        }
        if (k < m) {
            const int64_t d0 = __builtin_amdgcn_readlane(my_d, k);
            const float w0 = __int_as_float(__builtin_amdgcn_readlane(my_w, k));
            const float g0 = gb[d0 * lanes];
            acc0 += w0 * g0;
            if (mk) t[k * 65 + ln] = g0;
        }
        if (mk) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (ln < m) {
                float acc = 0.f;
                uint64_t q = mk;
                while (q) {
                    const int b = __ffsll((unsigned long long)q) - 1;
                    acc += t[ln * 65 + b];
                    q &= q - 1;
                }
                gw[c + ln] += acc;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    if (on) g_src[row * lanes + ln] = acc0 + acc1;
}

// Delayed synapses (HeterSynapsePSC / DelayedPSC): psc_s receives
// W_sub^T z_{s - age(sub)} through the virtual rows j*csr_d + sub, so the
// spike z_t of neuron j collects sum over sub of W_sub dL/dpsc_{t + age(sub)}
// -- the reverse delay ring.  G holds dL/dpsc of the K steps of a window,
// [K][n][lanes]; step tau of the window reads G[tau + age] when that index
// lies in [0, K) (later steps belong to another window).
__global__ __launch_bounds__(256) void k_adj_transpose_delayed(
    const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
    const float* __restrict__ w, int64_t n, int csr_d, int64_t lanes,
    const float* __restrict__ G, int64_t K, int64_t tau, const int* __restrict__ ages,
    float* __restrict__ g_src, int accumulate) {
    const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int ln = threadIdx.x & 63;
    if (j >= n) return;
    for (int64_t b0 = 0; b0 < lanes; b0 += 64) {
        const int64_t b = b0 + ln;
        const bool on = b < lanes;
        float acc = 0.f;
        for (int sub = 0; sub < csr_d; ++sub) {
            const int age = ages[sub];
            const int64_t st = tau + age;
            if (age < 0 || st < 0 || st >= K) continue;
            const float* g = G + st * n * lanes;
            const int64_t r = j * csr_d + sub;
            for (int64_t e = rp[r]; e < rp[r + 1]; ++e) {
                const int64_t d = col[e];
                const float wv = w[e];
                if (on) acc += wv * g[d * lanes + b];
            }
        }
        if (on) g_src[j * lanes + b] = accumulate ? g_src[j * lanes + b] + acc : acc;
    }
}

// gw[e] += sum over lanes b spiking at neuron j (mask bit b) of
// G[tau + age(sub)][col[e], b], for every virtual row j*csr_d + sub whose
// step lies in the window.  One wavefront per neuron; lanes stride the edges.
__global__ __launch_bounds__(256) void k_adj_wgrad_delayed(
    const int64_t* __restrict__ rp, const int32_t* __restrict__ col, int64_t n, int csr_d,
    int64_t lanes, const uint64_t* __restrict__ masks, const float* __restrict__ G, int64_t K,
    int64_t tau, const int* __restrict__ ages, float* __restrict__ gw) {
    const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int ln = threadIdx.x & 63;
    if (j >= n) return;
    const uint64_t mk = masks[j];
    if (mk == 0) return;
    for (int sub = 0; sub < csr_d; ++sub) {
        const int age = ages[sub];
        const int64_t st = tau + age;
        if (age < 0 || st < 0 || st >= K) continue;
        const float* g = G + st * n * lanes;
        const int64_t r = j * csr_d + sub;
        for (int64_t e = rp[r] + ln; e < rp[r + 1]; e += 64) {
            const int64_t d = col[e];
            float acc = 0.f;
            uint64_t q = mk;
            while (q) {
                const int b = __ffsll((unsigned long long)q) - 1;
                acc += g[d * lanes + b];
                q &= q - 1;
            }
            gw[e] += acc;
        }
    }
}

}  // namespace btx

using namespace btx;

BTX_API int btx_adj_lif_step(const btx_adj_lif_args* p, void* stream) {
    if (p == nullptr) return fail("btx_adj_lif_step: null args");
    const int64_t m = p->n * p->lanes;
    if (m <= 0) return 0;
    if (p->surrogate < BTX_SURR_SIGMOID || p->surrogate > BTX_SURR_ERF)
        return fail("btx_adj_lif_step: unknown surrogate");
    AdjArgs a{};
    a.n = p->n;
    a.lanes = p->lanes;
    a.dt = p->dt;
    a.hard_reset = p->hard_reset;
    a.detach_reset = p->detach_reset;
    a.surrogate = p->surrogate;
    a.alpha = p->alpha;
    a.damping = p->damping;
    a.tau = p->tau; a.tau_s = p->tau_stride;
    a.vth = p->v_th; a.vth_s = p->v_th_stride;
    a.vr = p->v_reset; a.vr_s = p->v_reset_stride;
    a.cm = p->c_m; a.cm_s = p->c_m_stride;
    a.tsyn = p->tau_syn; a.tsyn_s = p->tau_syn_stride;
    a.ksyn = p->k_syn; a.ksyn_s = p->k_syn_stride;
    a.v_prev = p->v_prev;
    a.v_pre = p->v_pre;
    a.z = p->z;
    a.ref_prev = p->ref_prev;
    a.gz = p->gz;
    a.gsyn = p->gsyn;
    a.gv = p->gv;
    a.gpsc = p->gpsc;
    a.gx = p->gx;
    a.gtau = p->gtau;
    hipLaunchKernelGGL(k_adj_lif, dim3((unsigned)((m + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, a);
    return check_launch("btx_adj_lif_step");
}

BTX_API int btx_adj_transpose(const int64_t* row_ptr, const int32_t* col, const float* w,
                              int64_t n_rows, int64_t lanes, const float* g_dst, float* g_src,
                              void* stream) {
    if (n_rows <= 0 || lanes <= 0) return 0;
    hipLaunchKernelGGL(k_adj_transpose, dim3((unsigned)((n_rows + 3) / 4)), dim3(256), 0,
                       (hipStream_t)stream, row_ptr, col, w, n_rows, lanes, g_dst, g_src);
    return check_launch("btx_adj_transpose");
}

BTX_API int btx_adj_wgrad(const int64_t* row_ptr, const int32_t* col, int64_t n_rows,
                          int64_t lanes, const uint64_t* masks, const float* g_dst, float* gw,
                          void* stream) {
    if (n_rows <= 0 || lanes <= 0) return 0;
    if (lanes > 64) return fail("btx_adj_wgrad: at most 64 lanes per mask word");
    hipLaunchKernelGGL(k_adj_wgrad, dim3((unsigned)((n_rows + 3) / 4)), dim3(256), 0,
                       (hipStream_t)stream, row_ptr, col, n_rows, lanes, masks, g_dst, gw);
    return check_launch("btx_adj_wgrad");
}

BTX_API int btx_adj_lif_chunk(const btx_adj_lif_args* p, int64_t K, const int64_t* row_ptr,
                              const int32_t* col, const float* w, const uint64_t* masks,
                              float* gw, void* stream) {
    if (p == nullptr) return fail("btx_adj_lif_chunk: null args");
    if (K <= 0 || p->n <= 0 || p->lanes <= 0) return 0;
    if (p->lanes > 64) return fail("btx_adj_lif_chunk: at most 64 lanes per mask word");
    if (p->gsyn == nullptr || p->gpsc == nullptr || masks == nullptr || gw == nullptr)
        return fail("btx_adj_lif_chunk: gsyn, gpsc, masks and gw are required");
    const int64_t m = p->n * p->lanes;
    btx_adj_lif_args s = *p;
    for (int64_t t = K - 1; t >= 0; --t) {
        // dL/dpsc after step t through the transposed product and the weight
        // gradient of step t's spikes (one pass), then the elementwise
        // reverse step
        hipLaunchKernelGGL(k_adj_transpose_wgrad, dim3((unsigned)((p->n + 3) / 4)), dim3(256), 0,
                           (hipStream_t)stream, row_ptr, col, w, p->n, p->lanes,
                           masks + t * p->n, p->gpsc, (float*)p->gsyn, gw);
        if (int rc = check_launch("btx_adj_lif_chunk")) return rc;
        s.v_prev = p->v_prev + t * m;
        s.v_pre = p->v_pre + t * m;
        s.z = p->z + t * m;
        s.ref_prev = p->ref_prev ? p->ref_prev + t * m : nullptr;
        s.gz = p->gz + t * m;
        s.gx = p->gx + t * m;
        if (int rc = btx_adj_lif_step(&s, stream)) return rc;
    }
    return 0;
}

BTX_API int btx_adj_transpose_delayed(const int64_t* row_ptr, const int32_t* col, const float* w,
                                      int64_t n, int csr_d, int64_t lanes, const float* G,
                                      int64_t K, int64_t tau, const int* ages, float* g_src,
                                      int accumulate, void* stream) {
    if (n <= 0 || lanes <= 0) return 0;
    if (csr_d < 1 || csr_d > BTX_MAX_BINS) return fail("btx_adj_transpose_delayed: bad csr_d");
    hipLaunchKernelGGL(k_adj_transpose_delayed, dim3((unsigned)((n + 3) / 4)), dim3(256), 0,
                       (hipStream_t)stream, row_ptr, col, w, n, csr_d, lanes, G, K, tau, ages,
                       g_src, accumulate);
    return check_launch("btx_adj_transpose_delayed");
}

BTX_API int btx_adj_wgrad_delayed(const int64_t* row_ptr, const int32_t* col, int64_t n,
                                  int csr_d, int64_t lanes, const uint64_t* masks,
                                  const float* G, int64_t K, int64_t tau, const int* ages,
                                  float* gw, void* stream) {
    if (n <= 0 || lanes <= 0) return 0;
    if (lanes > 64) return fail("btx_adj_wgrad_delayed: at most 64 lanes per mask word");
    if (csr_d < 1 || csr_d > BTX_MAX_BINS) return fail("btx_adj_wgrad_delayed: bad csr_d");
    hipLaunchKernelGGL(k_adj_wgrad_delayed, dim3((unsigned)((n + 3) / 4)), dim3(256), 0,
                       (hipStream_t)stream, row_ptr, col, n, csr_d, lanes, masks, G, K, tau,
                       ages, gw);
    return check_launch("btx_adj_wgrad_delayed");
}
