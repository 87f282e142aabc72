"""Torch-CPU restatement of btorch's recurrent SNN step -- TEST INFRASTRUCTURE ONLY.

Every function performs the reference's eager torch ops in the reference's
order, on CPU tensors, so in float32 it reproduces the reference's native
path bit for bit; with ``dtype=torch.float64`` the same code is the float64
error oracle.  Citations are to the reference (jiaxinou2022/btorch).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import scipy.sparse
import torch

_HERE = Path(__file__).resolve().parent


# ---------------------------------------------------------------------------
# Connectivity
# ---------------------------------------------------------------------------

def coalesce(conn) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(dst, src, w) of ``conn`` (n_src, n_dst) transposed and coalesced,
    sorted by (dst, src): BaseSparseConn.__init__, linear.py:274-291."""
    coo = scipy.sparse.coo_array(conn)
    t = coo.T.tocoo()
    t.sum_duplicates()
    return t.row.astype(np.int64), t.col.astype(np.int64), t.data.astype(np.float32)


def numpy_pairwise_sum(a) -> np.float32:
    """NumPy's float32 pairwise sum (numpy/_core/src/umath/loops_utils.h.src,
    pairwise_sum): sequential from -0.0 below 8 terms, eight accumulators up
    to 128, halves split at a multiple of 8 above."""
    f32 = np.float32
    n = len(a)
    if n < 8:
        res = f32(-0.0)
        for x in a:
            res = f32(res + x)
        return res
    if n <= 128:
        r = [f32(x) for x in a[:8]]
        i = 8
        while i < n - n % 8:
            for j in range(8):
                r[j] = f32(r[j] + a[i + j])
            i += 8
        res = f32(f32(f32(r[0] + r[1]) + f32(r[2] + r[3])) + f32(f32(r[4] + r[5]) + f32(r[6] + r[7])))
        for x in a[i:]:
            res = f32(res + x)
        return res
    n2 = n // 2
    n2 -= n2 % 8
    return f32(numpy_pairwise_sum(a[:n2]) + numpy_pairwise_sum(a[n2:]))


def duplicate_sum(values) -> np.float32:
    """Weight of duplicated entries after scipy's sum_duplicates
    (np.add.reduceat over the stably sorted entries, linear.py:283-284): the
    first value plus the pairwise sum of the rest, in input order."""
    values = [np.float32(v) for v in values]
    if len(values) == 1:
        return values[0]
    return np.float32(values[0] + numpy_pairwise_sum(values[1:]))


def expand_conn_for_delays(conn, delays, n_delay_bins: int):
    """connection.py:973-1027: row pre*D + clip(delay, 0, D-1)."""
    delays = np.asarray(delays)
    if np.any(delays < 0):
        raise ValueError("delays must be non-negative")
    coo = scipy.sparse.coo_array(conn)
    bins = np.clip(delays, 0, n_delay_bins - 1).astype(int)
    rows = coo.row * n_delay_bins + bins
    return scipy.sparse.coo_array((coo.data, (rows, coo.col)),
                                  shape=(coo.shape[0] * n_delay_bins, coo.shape[1]))


class NativeSparse:
    """linear.py:474-483: out = (A_{dst,src} @ x^T)^T with torch.sparse.mm."""

    def __init__(self, conn, dtype=torch.float32):
        dst, src, w = coalesce(conn)
        self.n_src, self.n_dst = scipy.sparse.coo_array(conn).shape
        idx = torch.stack([torch.as_tensor(dst), torch.as_tensor(src)])
        self.dst, self.src, self.w = dst, src, w
        self.A = torch.sparse_coo_tensor(idx, torch.as_tensor(w).to(dtype),
                                         size=(self.n_dst, self.n_src), is_coalesced=True)

    @classmethod
    def from_coalesced(cls, dst, src, w, n_src, n_dst, dtype=torch.float32):
        """Same operator from (dst, src)-sorted coalesced arrays."""
        obj = cls.__new__(cls)
        obj.n_src, obj.n_dst = n_src, n_dst
        obj.dst, obj.src, obj.w = dst, src, w
        idx = torch.stack([torch.as_tensor(dst), torch.as_tensor(src)])
        obj.A = torch.sparse_coo_tensor(idx, torch.as_tensor(w).to(dtype), size=(n_dst, n_src),
                                        is_coalesced=True)
        return obj

    def __call__(self, x2d: torch.Tensor) -> torch.Tensor:
        return torch.sparse.mm(self.A, x2d.T).T


def seq_spmm_c(dst, src, w, x2d: np.ndarray, n_dst: int) -> np.ndarray:
    """The C restatement of the per-destination sequential fmaf order."""
    lib = ctypes.CDLL(str(_HERE / "liboracle.so"))
    x2d = np.ascontiguousarray(x2d, dtype=np.float32)
    lanes, n_src = x2d.shape
    out = np.empty((lanes, n_dst), np.float32)
    dst = np.ascontiguousarray(dst, np.int64)
    src = np.ascontiguousarray(src, np.int64)
    w = np.ascontiguousarray(w, np.float32)
    P = ctypes.c_void_p
    lib.seq_spmm.argtypes = [P, P, P, ctypes.c_int64, P, ctypes.c_int64, ctypes.c_int64,
                             ctypes.c_int64, P]
    lib.seq_spmm(dst.ctypes.data, src.ctypes.data, w.ctypes.data, len(w), x2d.ctypes.data,
                 lanes, n_src, n_dst, out.ctypes.data)
    return out


# ---------------------------------------------------------------------------
# Neuron and synapse
# ---------------------------------------------------------------------------

@dataclass
class LIFParams:
    """LIF parameters as 0-d or (N,) tensors (lif.py:94-128, base.py:1032-1045)."""

    tau: torch.Tensor
    v_threshold: torch.Tensor
    v_reset: torch.Tensor
    c_m: torch.Tensor
    tau_ref: torch.Tensor | None
    hard_reset: bool = False

    @staticmethod
    def make(n, tau=20.0, v_threshold=1.0, v_reset=0.0, c_m=1.0, tau_ref=None,
             hard_reset=False, dtype=torch.float32):
        def t(v):
            return None if v is None else torch.as_tensor(v, dtype=dtype)
        return LIFParams(t(tau), t(v_threshold), t(v_reset), t(c_m), t(tau_ref), hard_reset)


def heaviside(x: torch.Tensor) -> torch.Tensor:
    """surrogate/base.py:6-8."""
    return (x >= 0).to(x)


def lif_step(v, ref, x, p: LIFParams, dt: float):
    """One LIF step (base.py:1108-1116 with lif.py:144-187).

    Returns (spike, v_new, ref_new, v_pre)."""
    derivative = -(v - p.v_reset) / p.tau + x / p.c_m
    v = v + dt * derivative
    v_pre = v
    spike = heaviside((v - p.v_threshold) / (p.v_threshold - p.v_reset))
    if p.tau_ref is not None:
        spike = spike * (ref == 0).to(v.dtype)
    if p.hard_reset:
        v = v - (v - p.v_reset) * spike
    else:
        v = v - (p.v_threshold - p.v_reset) * spike
    if p.tau_ref is not None:
        ref = torch.relu(ref + spike * p.tau_ref - dt)
    return spike, v, ref, v_pre


@dataclass
class GLIF3Params:
    """GLIF3 parameters (glif.py:128-237): 0-d or (N,) tensors; k and
    asc_amps (n_asc,) or (N, n_asc); v_rest defaults to v_reset."""

    tau: torch.Tensor
    v_threshold: torch.Tensor
    v_reset: torch.Tensor
    v_rest: torch.Tensor
    c_m: torch.Tensor
    tau_ref: torch.Tensor | None
    k: torch.Tensor
    asc_amps: torch.Tensor
    hard_reset: bool = False

    @staticmethod
    def make(n, tau=20.0, v_threshold=-50.0, v_reset=-70.0, v_rest=None, c_m=0.05,
             tau_ref=0.0, k=(0.2,), asc_amps=(0.0,), hard_reset=False, dtype=torch.float32):
        def t(v):
            return None if v is None else torch.as_tensor(v, dtype=dtype)
        return GLIF3Params(t(tau), t(v_threshold), t(v_reset),
                           t(v_reset if v_rest is None else v_rest), t(c_m), t(tau_ref),
                           t(k), t(asc_amps), hard_reset)


def glif3_step(v, iasc, ref, x, p: GLIF3Params, dt: float):
    """One GLIF3 step, charge -> adaptation -> fire -> reset (base.py:1108-1116)
    with glif.py:264-342 and exp_euler_step (ode.py:20-50).

    Returns (spike, v_new, iasc_new, ref_new, v_pre)."""
    # charge: exp_euler_step(dV, v, Iasc, x), glif.py:264-297
    derivative = -(v - p.v_rest) / p.tau + (x + iasc.sum(-1, dtype=iasc.dtype)) / p.c_m
    linear = -1.0 / p.tau
    v = v + torch.expm1(dt * linear) / linear * derivative
    # adaptation: exp_euler_step(dIasc, Iasc), glif.py:260-262, 299-300
    d_i, lin_i = -p.k * iasc, -p.k
    iasc = iasc + torch.expm1(dt * lin_i) / lin_i * d_i
    # fire, glif.py:302-310
    v_pre = v
    spike = heaviside((v - p.v_threshold) / (p.v_threshold - p.v_reset))
    if p.tau_ref is not None:
        spike = spike * (ref == 0).to(v.dtype)
    # reset, glif.py:312-336
    if p.hard_reset:
        v = v - (v - p.v_reset) * spike
    else:
        v = v - (p.v_threshold - p.v_reset) * spike
    iasc = iasc + p.asc_amps * spike[..., None]
    if p.tau_ref is not None:
        ref = torch.relu(ref + spike * p.tau_ref - dt)
    return spike, v, iasc, ref, v_pre


@dataclass
class TwoCompParams:
    """TwoCompartmentGLIF parameters (two_compartment.py:127-182): 0-d or
    (N,) tensors, defaults as in the reference constructor."""

    v_threshold: torch.Tensor
    v_reset: torch.Tensor
    tau_s: torch.Tensor
    R_s: torch.Tensor
    E_L: torch.Tensor
    tau_a: torch.Tensor
    tau_th: torch.Tensor
    delta_th: torch.Tensor
    delta_T: torch.Tensor
    w_Ca: torch.Tensor
    theta_Ca: torch.Tensor
    w_sa: torch.Tensor
    w_as: torch.Tensor

    @staticmethod
    def make(v_threshold=-50.0, v_reset=-65.0, tau_s=20.0, R_s=1.0, E_L=-70.0, tau_a=100.0,
             tau_th=50.0, delta_th=0.0, delta_T=0.0, w_Ca=0.0, theta_Ca=1.0, w_sa=0.5,
             w_as=1.0, dtype=torch.float32):
        vals = (v_threshold, v_reset, tau_s, R_s, E_L, tau_a, tau_th, delta_th, delta_T, w_Ca,
                theta_Ca, w_sa, w_as)
        return TwoCompParams(*(torch.as_tensor(v, dtype=dtype) for v in vals))

    def to(self, dtype):
        return TwoCompParams(*(getattr(self, f).to(dtype) for f in self.__dataclass_fields__))


def tc_step(v, i_a, i_bap, theta_th, i_soma, i_apical, p: TwoCompParams, dt: float):
    """One TwoCompartmentGLIF.single_step_forward (two_compartment.py:235-292):
    semi-implicit apical current (:208-217), somatic voltage with the
    exponential initiation term (:219-241), surrogate Heaviside spike, hard
    reset (:243-249), bAP current and adaptive threshold (:251-259).

    Returns (spike, v, i_a, i_bap, theta_th, v_pre)."""
    # _update_apical_current, :212-217
    tau_a = torch.clamp(p.tau_a, min=1e-4)
    drive = i_apical + p.w_Ca * torch.sigmoid(i_a - p.theta_Ca) + i_bap
    decay = dt / tau_a
    i_a_next = (i_a + decay * drive) / (1.0 + decay)
    threshold_eff = p.v_threshold + theta_th
    # _update_somatic_voltage, :219-241
    tau_s = torch.clamp(p.tau_s, min=1e-4)
    R_s = torch.clamp(p.R_s, min=1e-4)
    total_current = i_soma + p.w_as * i_a_next
    delta_T = torch.clamp(p.delta_T, min=0.0)
    exp_arg = torch.clamp((v - threshold_eff) / torch.clamp(delta_T, min=1e-4), max=10.0)
    exp_drive = torch.where(delta_T > 0.0, delta_T * torch.exp(exp_arg), torch.zeros_like(v))
    decay = dt / tau_s
    v_pre = (v + decay * (p.E_L + R_s * total_current + exp_drive)) / (1.0 + decay)
    spike = heaviside(v_pre - threshold_eff)
    # state updates, :281-284
    v = v_pre - (v_pre - p.v_reset) * spike
    i_bap = p.w_sa * spike
    tau_th = torch.clamp(p.tau_th, min=1e-4)
    theta_th = theta_th / (1.0 + dt / tau_th) + p.delta_th * spike
    return spike, v, i_a_next, i_bap, theta_th, v_pre


@dataclass
class MixedParams:
    """MixedNeuronPopulation (mixed.py:16-193): contiguous groups
    (start, count, LIFParams | GLIF3Params | TwoCompParams) along the
    neuron axis, in order."""

    groups: list

    def to(self, dtype):
        out = []
        for s0, c, g in self.groups:
            if isinstance(g, TwoCompParams):
                g = g.to(dtype)
            elif isinstance(g, GLIF3Params):
                g = GLIF3Params(*(None if t is None else t.to(dtype) for t in
                                  (g.tau, g.v_threshold, g.v_reset, g.v_rest, g.c_m, g.tau_ref,
                                   g.k, g.asc_amps)), g.hard_reset)
            else:
                g = LIFParams(*(None if t is None else t.to(dtype) for t in
                                (g.tau, g.v_threshold, g.v_reset, g.c_m, g.tau_ref)),
                              g.hard_reset)
            out.append((s0, c, g))
        return MixedParams(out)


def mixed_step(st: dict, I, xa, p: MixedParams, dt: float):
    """MixedNeuronPopulation.single_step_forward (mixed.py:113-141): each
    group steps on its slice of the input (the two-compartment groups with
    their slice of the apical input), the spikes are concatenated.  ``st``
    holds v, refractory, i_a, i_bap, theta_th (B, N) and per-group iasc;
    returns (spike, v_pre) and updates ``st``."""
    z = torch.empty_like(I)
    v_pre = torch.empty_like(I)
    v, ref = st["v"].clone(), st["refractory"].clone()
    ia, ib, th = st["i_a"].clone(), st["i_bap"].clone(), st["theta_th"].clone()
    for gi, (s0, c, g) in enumerate(p.groups):
        sl = slice(s0, s0 + c)
        if isinstance(g, TwoCompParams):
            zg, vg, iag, ibg, thg, vpg = tc_step(st["v"][:, sl], st["i_a"][:, sl],
                                                 st["i_bap"][:, sl], st["theta_th"][:, sl],
                                                 I[:, sl], xa[:, sl], g, dt)
            ia[:, sl], ib[:, sl], th[:, sl] = iag, ibg, thg
        elif isinstance(g, GLIF3Params):
            zg, vg, iag, rg, vpg = glif3_step(st["v"][:, sl], st["iasc"][gi],
                                              st["refractory"][:, sl], I[:, sl], g, dt)
            st["iasc"][gi] = iag
            ref[:, sl] = rg
        else:
            zg, vg, rg, vpg = lif_step(st["v"][:, sl], st["refractory"][:, sl], I[:, sl], g, dt)
            ref[:, sl] = rg
        z[:, sl], v[:, sl], v_pre[:, sl] = zg, vg, vpg
    st.update(v=v, refractory=ref, i_a=ia, i_bap=ib, theta_th=th)
    return z, v_pre


def exp_psc_decay(psc, tau_syn, dt: float):
    """ExponentialPSC.conductance_charge (synapse.py:171-178) through
    exp_euler_step (ode.py:20-50): x + expm1(dt*A)/A * f(x)."""
    derivative = -psc / tau_syn
    linear = -1.0 / tau_syn
    return psc + torch.expm1(dt * linear) / linear * derivative


def _exp_euler(x, derivative, tau, dt: float):
    """exp_euler_step (ode.py:20-50) with the linear term -1/tau."""
    linear = -1.0 / tau
    return x + torch.expm1(dt * linear) / linear * derivative


@dataclass
class AlphaSyn:
    """AlphaPSC (synapse.py:274-334): conductance_charge then adaptation_charge."""
    tau_syn: torch.Tensor
    g_max: torch.Tensor
    states = ("h",)

    def to(self, dtype):
        return AlphaSyn(torch.as_tensor(self.tau_syn).to(dtype),
                        torch.as_tensor(self.g_max).to(dtype))

    def conductance(self, st, dt):
        psc, h, tau = st["psc"], st["h"], self.tau_syn
        st["psc"] = _exp_euler(psc, -psc / tau + h / tau, tau, dt)     # synapse.py:296-305

    def adaptation(self, st, mm_z, dt):
        wz = self.g_max * mm_z                                           # synapse.py:314
        st["h"] = _exp_euler(st["h"], -st["h"] / self.tau_syn, self.tau_syn, dt) + wz


@dataclass
class BillehSyn:
    """AlphaPSCBilleh (synapse.py:201-271); dt must be 1."""
    tau_syn: torch.Tensor
    states = ("h",)

    def to(self, dtype):
        return BillehSyn(torch.as_tensor(self.tau_syn).to(dtype))

    @property
    def syn_decay(self):
        return torch.exp(-1.0 / self.tau_syn)                            # synapse.py:235-237

    def conductance(self, st, dt):
        c = self.syn_decay
        st["psc"] = c * st["psc"] + c * st["h"]                          # synapse.py:239-241

    def adaptation(self, st, mm_z, dt):
        st["h"] = self.syn_decay * st["h"] + torch.e / self.tau_syn * mm_z  # synapse.py:253


@dataclass
class DualExpSyn:
    """DualExponentialPSC (synapse.py:337-418) with the default A (or a given one)."""
    tau_decay: torch.Tensor
    tau_rise: torch.Tensor
    A: torch.Tensor | None = None
    states = ("g_rise", "g_decay")

    def to(self, dtype):
        return DualExpSyn(torch.as_tensor(self.tau_decay).to(dtype),
                          torch.as_tensor(self.tau_rise).to(dtype),
                          None if self.A is None else torch.as_tensor(self.A).to(dtype))

    @property
    def a(self):
        td, tr = self.tau_decay, self.tau_rise
        A = self.A
        if A is None:                                                    # synapse.py:369-375
            A = td / (td - tr) * (tr / td) ** (tr / (tr - td))
        return (td - tr) / tr / td * A                                   # synapse.py:377

    def conductance(self, st, dt):                                       # synapse.py:390-392
        st["g_rise"] = _exp_euler(st["g_rise"], -st["g_rise"] / self.tau_rise, self.tau_rise, dt)
        st["g_decay"] = _exp_euler(st["g_decay"], -st["g_decay"] / self.tau_decay,
                                   self.tau_decay, dt)

    def adaptation(self, st, mm_z, dt):                                  # synapse.py:403-406
        st["g_rise"] = st["g_rise"] + mm_z
        st["g_decay"] = st["g_decay"] + mm_z
        st["psc"] = self.a * (st["g_decay"] - st["g_rise"])


def syn_impulse_response(syn, T: int, dt: float, dtype=torch.float64) -> torch.Tensor:
    """psc over T steps of one synapse given W z = 1 at step 0 (the state the
    neuron reads at the next step)."""
    syn = syn.to(dtype)
    st = {"psc": torch.zeros(1, dtype=dtype)}
    st.update({k: torch.zeros(1, dtype=dtype) for k in syn.states})
    out = []
    for t in range(T):
        syn.conductance(st, dt)
        syn.adaptation(st, torch.full((1,), 1.0 if t == 0 else 0.0, dtype=dtype), dt)
        out.append(st["psc"].clone())
    return torch.cat(out)


# ---------------------------------------------------------------------------
# Delay ring (history.py:28-358)
# ---------------------------------------------------------------------------

class Ring:
    """SpikeHistory buffer (B, D, N) in circular or cat mode."""

    def __init__(self, lanes, n, D, circular=True, dtype=torch.float32, init=None, cursor=0):
        self.D, self.n, self.circular = D, n, circular
        self.h = torch.zeros(lanes, D, n, dtype=dtype) if init is None else init.clone()
        self.cursor = cursor

    def update(self, z):
        if self.circular:
            self.h[:, self.cursor] = z
            self.cursor = (self.cursor + 1) % self.D
        else:
            self.h = torch.cat([z.unsqueeze(1), self.h[:, :-1]], dim=1)

    def get_delay(self, d):
        return self.h[:, (self.cursor - 1 - d) % self.D] if self.circular else self.h[:, d]

    def get_flattened(self, n_delays):
        stacked = torch.stack([self.get_delay(d) for d in range(n_delays)], dim=-1)
        return stacked.reshape(stacked.shape[0], -1)  # index n*D + d


# ---------------------------------------------------------------------------
# RecurrentNN(LIF, ExponentialPSC-like(SparseConn)) time loop (rnn.py:517-528)
# ---------------------------------------------------------------------------

@dataclass
class Net:
    n: int
    conn: object               # scipy (n_src, n_dst); n_src = n*csr_d
    lif: LIFParams
    tau_syn: torch.Tensor
    dt: float
    delay_mode: str = "none"   # "none" | "hetero" (HeterSynapsePSC) | "fixed" (DelayedPSC)
    max_delay_steps: int = 1
    circular: bool = False
    syn: object = None         # AlphaSyn | BillehSyn | DualExpSyn; None = ExponentialPSC(tau_syn)


def simulate(net: Net, x_seq: torch.Tensor, state=None, dtype=torch.float32, teacher=None,
             x_apical=None):
    """Run T steps; x_seq (T, B, N).  ``teacher`` (T, B, N) forces the spikes
    fed to the synapse (teacher forcing) while still reporting the produced
    raster.  With TwoCompParams the cell is ApicalRecurrentNN over a
    TwoCompartmentGLIF population (rnn.py:599-649) and ``x_apical`` (T, B, N)
    its apical input (None = zeros).  Returns dict of stacked z, v,
    refractory, psc, v_pre (T, B, N) and the final state."""
    T, B, n = x_seq.shape
    glif = isinstance(net.lif, GLIF3Params)
    tc = isinstance(net.lif, TwoCompParams)
    mixed = isinstance(net.lif, MixedParams)
    if mixed:
        p = net.lif.to(dtype)
        p_ref = next((g for _, _, g in p.groups if not isinstance(g, TwoCompParams)), None)
        v_reset0 = torch.cat([torch.as_tensor(g.v_reset, dtype=dtype).expand(c)
                              for _, c, g in p.groups])
    elif tc:
        p = net.lif.to(dtype)
    elif glif:
        g = net.lif
        p = GLIF3Params(*(None if t is None else t.to(dtype) for t in
                          (g.tau, g.v_threshold, g.v_reset, g.v_rest, g.c_m, g.tau_ref, g.k,
                           g.asc_amps)), g.hard_reset)
        n_asc = int(p.k.shape[-1]) if p.k.ndim else 1
    else:
        p = LIFParams(*(None if t is None else t.to(dtype) for t in
                        (net.lif.tau, net.lif.v_threshold, net.lif.v_reset, net.lif.c_m,
                         net.lif.tau_ref)), net.lif.hard_reset)
    tau_syn = torch.as_tensor(net.tau_syn).to(dtype)
    mm = net.conn if isinstance(net.conn, NativeSparse) else NativeSparse(net.conn, dtype)
    st = state or {}
    # memory v starts at v_reset (register_memory("v", v_reset, ...), base.py:1022)
    v0 = torch.zeros(B, n, dtype=dtype) + (v_reset0 if mixed else p.v_reset)
    v = st.get("v", v0).to(dtype).clone()
    ref = st.get("refractory", torch.zeros(B, n, dtype=dtype)).to(dtype).clone()
    psc = st.get("psc", torch.zeros(B, n, dtype=dtype)).to(dtype).clone()
    syn = net.syn.to(dtype) if net.syn is not None else None
    sst = {}
    if syn is not None:
        sst = {k: st.get(k, torch.zeros(B, n, dtype=dtype)).to(dtype).clone()
               for k in syn.states}
    iasc = (st.get("iasc", torch.zeros(B, n, n_asc, dtype=dtype)).to(dtype).clone()
            if glif else None)
    if tc or mixed:
        zero = torch.zeros(B, n, dtype=dtype)
        i_a, i_bap, th = (st.get(k, zero).to(dtype).clone() for k in ("i_a", "i_bap", "theta_th"))
    if mixed:
        mst = {"v": v, "refractory": ref, "i_a": i_a, "i_bap": i_bap, "theta_th": th,
               "iasc": {}}
        for gi, (s0, c, g) in enumerate(p.groups):
            if isinstance(g, GLIF3Params):
                m = int(g.k.shape[-1]) if g.k.ndim else 1
                init = st.get("iasc_groups", {}).get(gi)
                mst["iasc"][gi] = (init.to(dtype).clone() if init is not None
                                   else torch.zeros(B, c, m, dtype=dtype))
    ring = None
    if net.delay_mode == "hetero" and net.max_delay_steps > 1:
        ring = Ring(B, n, net.max_delay_steps, net.circular, dtype, st.get("history"),
                    st.get("cursor", 0))
    elif net.delay_mode == "fixed" and net.max_delay_steps > 1:
        ring = Ring(B, n, net.max_delay_steps + 1, net.circular, dtype, st.get("history"),
                    st.get("cursor", 0))
    rec = {k: [] for k in ("z", "v", "refractory", "psc", "v_pre", "i_a", "theta_th",
                           *(syn.states if syn is not None else ()))}
    x_seq = x_seq.to(dtype)
    for t in range(T):
        if mixed:
            xa = x_apical[t].to(dtype) if x_apical is not None else torch.zeros_like(v)
            z, v_pre = mixed_step(mst, psc + x_seq[t], xa, p, net.dt)
            v, ref, i_a, i_bap, th = (mst[k] for k in ("v", "refractory", "i_a", "i_bap",
                                                       "theta_th"))
        elif tc:
            xa = x_apical[t].to(dtype) if x_apical is not None else torch.zeros_like(v)
            z, v, i_a, i_bap, th, v_pre = tc_step(v, i_a, i_bap, th, psc + x_seq[t], xa, p,
                                                  net.dt)
        elif glif:
            z, v, iasc, ref, v_pre = glif3_step(v, iasc, ref, psc + x_seq[t], p, net.dt)
        else:
            z, v, ref, v_pre = lif_step(v, ref, psc + x_seq[t], p, net.dt)
        zin = teacher[t].to(dtype) if teacher is not None else z
        if syn is None:
            psc = exp_psc_decay(psc, tau_syn, net.dt)
        else:
            sst["psc"] = psc
            syn.conductance(sst, net.dt)
        if ring is None:
            zf = zin
        elif net.delay_mode == "hetero":
            ring.update(zin)
            zf = ring.get_flattened(net.max_delay_steps)
        else:
            ring.update(zin)
            zf = ring.get_delay(net.max_delay_steps)
        if syn is None:
            psc = psc + mm(zf)
        else:
            syn.adaptation(sst, mm(zf), net.dt)
            psc = sst["psc"]
            for k in syn.states:
                rec[k].append(sst[k])
        for k, val in (("z", z), ("v", v), ("refractory", ref), ("psc", psc), ("v_pre", v_pre)):
            rec[k].append(val)
        if tc or mixed:
            rec["i_a"].append(i_a)
            rec["theta_th"].append(th)
    out = {k: torch.stack(vals) for k, vals in rec.items() if vals}
    out["final"] = {"v": v, "refractory": ref, "psc": psc}
    if syn is not None:
        out["final"].update({k: sst[k] for k in syn.states})
    if glif:
        out["final"]["iasc"] = iasc
    if tc or mixed:
        out["final"].update(i_a=i_a, i_bap=i_bap, theta_th=th)
    if mixed:
        out["final"]["iasc_groups"] = mst["iasc"]
    if ring is not None:
        out["final"]["history"] = ring.h
        out["final"]["cursor"] = ring.cursor
    return out


def active_lists(spikes: np.ndarray) -> list[np.ndarray]:
    """Ascending active indices per lane (the compaction's closed form)."""
    return [np.flatnonzero(row) .astype(np.int32) for row in np.asarray(spikes)]


# ---------------------------------------------------------------------------
# Autograd restatement (reference gradients; surrogate/base.py:11-62)
# ---------------------------------------------------------------------------

SIGMOID_K = 2.0 * np.log(np.sqrt(2.0) + 1.0)
SUPERSPIKE_K = np.sqrt(2.0) - 1.0


def surrogate_derivative(kind: str, u, alpha: float, damping: float):
    """sigmoid.py:22-29, superspike.py:25-31, erf.py:22-28, atan.py:22-27."""
    if kind == "sigmoid":
        s = torch.sigmoid(SIGMOID_K * alpha * u)
        return damping * 4.0 * s * (1.0 - s)
    if kind == "superspike":
        return damping / (1.0 + SUPERSPIKE_K * alpha * u.abs()) ** 2
    if kind == "atan":
        return damping / (1.0 + (alpha * u) ** 2)
    return damping * torch.exp(-np.log(2.0) * (alpha * u) ** 2)


def surrogate_primitive(kind: str, u, alpha: float):
    if kind == "sigmoid":
        return torch.sigmoid(SIGMOID_K * alpha * u)
    if kind == "superspike":
        ka = SUPERSPIKE_K * alpha
        return torch.where(u >= 0.0, 1.0 - 0.5 / (1.0 + ka * u), 0.5 / (1.0 - ka * u))
    if kind == "atan":  # atan.py:10-12
        return 0.5 + torch.atan(alpha * u) / np.pi
    return torch.special.erfc(-np.sqrt(np.log(2.0)) * alpha * u) / 2.0


class _Spike(torch.autograd.Function):
    """Heaviside (or primitive) forward, surrogate-derivative backward."""

    @staticmethod
    def forward(ctx, u, kind, alpha, damping, spiking):
        ctx.save_for_backward(u)
        ctx.cfg = (kind, alpha, damping)
        return heaviside(u) if spiking else surrogate_primitive(kind, u, alpha)

    @staticmethod
    def backward(ctx, g):
        (u,) = ctx.saved_tensors
        kind, alpha, damping = ctx.cfg
        return g * surrogate_derivative(kind, u, alpha, damping), None, None, None, None


def simulate_grad(n, dst, src, magnitude, x_seq, tau, dt, tau_syn, v_th=1.0, v_reset=0.0,
                  c_m=1.0, tau_ref=None, hard_reset=False, detach_reset=False,
                  surrogate=("superspike", 2.0, 1.0, True), state=None, traj=None,
                  csr_d=1, ages=(0,)):
    """Differentiable RecurrentNN(LIF, ExponentialPSC-like(SparseConn)), op
    for op; ``magnitude`` in coalesced (dst, virtual src) order.  With
    ``csr_d`` > 1 or a nonzero age, the synapse reads the flattened spike
    history (HeterSynapsePSC / DelayedPSC, synapse.py:537-761 with
    history.py:252-345): virtual row j*csr_d + sub carries z_{t - ages[sub]}
    of neuron j (zero before the run).  Returns z (T,B,N) and the final
    (v, refractory, psc); every op records autograd."""
    T, B, _ = x_seq.shape
    dtype = x_seq.dtype
    kind, alpha, damping, spiking = surrogate
    idx = torch.stack([torch.as_tensor(dst), torch.as_tensor(src)])
    A = torch.sparse_coo_tensor(idx, magnitude, size=(n, n * csr_d), is_coalesced=True)
    delayed = csr_d != 1 or tuple(ages) != (0,)
    hist = []
    vth = torch.as_tensor(v_th, dtype=dtype)
    vr = torch.as_tensor(v_reset, dtype=dtype)
    cm = torch.as_tensor(c_m, dtype=dtype)
    ts = torch.as_tensor(tau_syn, dtype=dtype)
    st = state or {}
    v = st.get("v", torch.zeros(B, n, dtype=dtype) + vr)
    ref = st.get("refractory", torch.zeros(B, n, dtype=dtype))
    psc = st.get("psc", torch.zeros(B, n, dtype=dtype))
    zs = []
    for t in range(T):
        xin = psc + x_seq[t]
        if traj is not None:
            traj.setdefault("v_prev", []).append(v.detach().clone())
            traj.setdefault("ref_prev", []).append(ref.detach().clone())
        v = v + dt * (-(v - vr) / tau + xin / cm)
        if traj is not None:
            traj.setdefault("v_pre", []).append(v.detach().clone())
        s = _Spike.apply((v - vth) / (vth - vr), kind, alpha, damping, spiking)
        if tau_ref is not None:
            s = s * (ref == 0).detach().to(dtype)
        sd = s.detach() if detach_reset else s
        v = v - (v - vr) * sd if hard_reset else v - (vth - vr) * sd
        if tau_ref is not None:
            ref = torch.relu(ref + sd * tau_ref - dt)
        psc = psc + torch.expm1(dt * (-1.0 / ts)) / (-1.0 / ts) * (-psc / ts)
        if delayed:
            hist.append(s)
            zero = torch.zeros_like(s)
            cols = [hist[-1 - a] if 0 <= a < len(hist) else zero for a in ages]
            zf = torch.stack(cols, dim=-1).reshape(B, n * csr_d)
            psc = psc + torch.sparse.mm(A, zf.T).T
        else:
            psc = psc + torch.sparse.mm(A, s.T).T
        zs.append(s)
    return torch.stack(zs), v, ref, psc


# ---------------------------------------------------------------------------
# Counter-based Poisson drive (the serving input; no reference counterpart:
# the reference feeds a precomputed input tensor, which this reproduces)
# ---------------------------------------------------------------------------

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def mix64(z: np.ndarray) -> np.ndarray:
    """SplitMix64 output finaliser (Steele, Lea, Flood, OOPSLA 2014) on uint64."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def poisson_drive(seed, prob, t0, T, n, kick) -> torch.Tensor:
    """(T, B, n) float32 drive: ``kick`` where neuron j's uniform at the
    lane's local step s = t0 + t falls below the lane's probability, else 0.

    stream = mix64(seed ^ s*0x100000001b3); h_q = mix64(stream + (q+1)*gamma)
    (SplitMix64 gamma 0x9e3779b97f4a7c15); neuron 2q takes bits 40..63 of h_q,
    neuron 2q+1 bits 16..39, each divided by 2^24.
    """
    seed = np.asarray(seed, dtype=np.int64).astype(np.uint64)
    prob = np.asarray(prob, dtype=np.float32)
    t0 = np.asarray(t0, dtype=np.int64)
    B = seed.size
    out = np.zeros((T, B, n), dtype=np.float32)
    nq = (n + 1) // 2
    q1 = np.arange(1, nq + 1, dtype=np.uint64)
    gamma = np.uint64(0x9E3779B97F4A7C15)
    scale = np.float32(1.0 / 16777216.0)
    for t in range(T):
        for b in range(B):
            s = np.uint64(int(t0[b]) + t)
            with np.errstate(over="ignore"):
                stream = mix64(seed[b] ^ (s * np.uint64(0x100000001B3)))
                h = mix64(stream + q1 * gamma)
            u = np.empty(2 * nq, dtype=np.float32)
            u[0::2] = (h >> np.uint64(40)).astype(np.float32) * scale
            u[1::2] = ((h >> np.uint64(16)) & np.uint64(0xFFFFFF)).astype(np.float32) * scale
            out[t, b] = np.where(u[:n] < prob[b], np.float32(kick), np.float32(0.0))
    return torch.from_numpy(out)


# ---------------------------------------------------------------------------
# Recorded spikes in coordinate form (io/serialization.py:80-147)
# ---------------------------------------------------------------------------

def spikes_sparse_repr(z: np.ndarray, var_name: str = "spikes") -> dict:
    """to_sparse_repr of a dense (T, B, N) array: np.nonzero's index arrays
    (int32), the values, and the marker's count and shape."""
    idx = np.nonzero(z)
    out = {f"{var_name}_idx_{d}": i.astype(np.int32) for d, i in zip(("time", "batch", "neuron"), idx)}
    out[f"{var_name}_data"] = np.asarray(z[idx])
    out[var_name] = {"nnz": int(len(idx[0])), "original_shape": list(z.shape)}
    return out
