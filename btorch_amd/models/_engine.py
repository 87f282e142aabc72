"""Fused time-loop engine: binds ``RecurrentNN(LIF, <exp synapse>(SparseConn))``
to ``btx_rsnn_run`` so that a whole sequence runs as device work enqueued
from C, with no per-step Python and no (B, N*D) history matrix.

Supported cells (everything else runs the generic per-step loop):
  neuron   LIF / IF (soft or hard reset, optional refractory, scalar or
           per-neuron parameters); GLIF3 with 1..4 after-spike currents;
           TwoCompartmentGLIF, with an apical input sequence, in a
           MixedNeuronPopulation of that one group (how the reference's
           ApicalRecurrentNN drives it); a one-group population of the
           other neurons too
  synapse  K(linear=SparseConn(N, N))
           DelayedPSC(K(SparseConn(N, N)), max_delay_steps=m)
           HeterSynapsePSC(base_psc=K, n_receptor=1,
                           linear=SparseConn(N*D, N), max_delay_steps=D)
           with kinetics K = ExponentialPSC, AlphaPSC, AlphaPSCBilleh or
           DualExponentialPSC (the last three with LIF or GLIF3 neurons)
  input    x (T, *batch, N), optionally through neuron_inp_module, applied to
           the whole sequence at once (a SparseConn becomes one batched
           event-driven product over T*B lanes)
Delay semantics follow rnn.py:517-528 + synapse.py:751-761: the spike of
step t enters the ring before the synapse reads it, so a synapse of delay d
delivers at the end of step t+d.
"""

from __future__ import annotations

import ctypes
import weakref
from dataclasses import dataclass, field

import torch

from .. import _lib
from . import _plan
from .history import SpikeHistory
from .linear import BaseSparseConn
from .neurons.glif import GLIF3
from .neurons.lif import LIF
from .neurons.mixed import MixedNeuronPopulation
from .neurons.two_compartment import TwoCompartmentGLIF
from .synapse import (AlphaPSC, AlphaPSCBilleh, DelayedPSC, DualExponentialPSC,
                      ExponentialPSC, HeterSynapsePSC)

KINETICS = (ExponentialPSC, AlphaPSC, AlphaPSCBilleh, DualExponentialPSC)


@dataclass
class CellSpec:
    neuron: LIF | GLIF3 | TwoCompartmentGLIF
    exp_psc: ExponentialPSC          # kinetics module: its psc decays and receives W z
    conn: BaseSparseConn
    history: SpikeHistory | None
    csr_d: int
    ring_d: int
    bins: list[tuple[int, int]]      # (ring age, virtual-row offset)
    psc_keys: list[str]              # dotted names whose state equals psc
    psc_modules: list = field(default_factory=list)
    neuron_key: str = "neuron."      # dotted prefix of the neuron's memories
    syn_keys: dict = field(default_factory=dict)  # dotted state name -> record key


def _bin_nonempty(conn: BaseSparseConn, csr_d: int) -> list[bool]:
    rp = conn._csr_row_ptr
    counts = (rp[1:] - rp[:-1]).reshape(-1, csr_d).sum(0)
    return [bool(c) for c in counts.tolist()]


class MixedState:
    """A multi-group MixedNeuronPopulation as one neuron of the fused step:
    population-wide state arrays (v, refractory, the two-compartment states,
    after-spike currents) that the engine gathers from the groups' memories
    before a run and writes back into them after it, and the groups' layout
    and parameters for the plan (``_plan.mixed_plan``)."""

    def __init__(self, pop: MixedNeuronPopulation):
        self.pop = pop
        self.size = pop.size
        self.n_neuron = (pop.size,)
        groups = pop.groups()
        self.has_tc = any(isinstance(g, TwoCompartmentGLIF) for g in groups)
        glifs = [g for g in groups if isinstance(g, GLIF3)]
        self.n_asc = glifs[0].n_Iasc if glifs else 0
        self._use_refractory = any(getattr(g, "_use_refractory", False) for g in groups
                                   if not isinstance(g, TwoCompartmentGLIF))
        self.v = self.refractory = self.i_a = self.i_bap = self.theta_th = self.Iasc = None
        self._lanes = None

    def modules(self) -> list:
        return self.pop.groups()

    def _buffers_for(self, lanes, dev):
        n = self.size
        if self._lanes != (lanes, str(dev)):
            z = lambda *shape: torch.zeros(shape, dtype=torch.float32, device=dev)  # noqa: E731
            self.v = z(lanes, n)
            self.refractory = z(lanes, n) if self._use_refractory else None
            if self.has_tc:
                self.i_a, self.i_bap, self.theta_th = z(lanes, n), z(lanes, n), z(lanes, n)
            self.Iasc = z(lanes, n, self.n_asc) if self.n_asc else None
            self._lanes = (lanes, str(dev))

    def _members(self):
        """(population array, group, memory name, trailing dims) to move."""
        out = []
        lay = self.pop.layout
        for g, mod in enumerate(self.pop.groups()):
            sl = lay.span(g)
            out.append((self.v, sl, mod, "v"))
            if isinstance(mod, TwoCompartmentGLIF):
                out += [(self.i_a, sl, mod, "i_a"), (self.i_bap, sl, mod, "i_bap"),
                        (self.theta_th, sl, mod, "theta_th")]
            elif getattr(mod, "_use_refractory", False):
                out.append((self.refractory, sl, mod, "refractory"))
            if isinstance(mod, GLIF3):
                out.append((self.Iasc, sl, mod, "Iasc"))
        return out

    def gather(self, lanes: int) -> None:
        """Population arrays from the groups' memories (v, refractory, ...)."""
        dev = self.pop.groups()[0].v.device
        self._buffers_for(lanes, dev)
        for arr, sl, mod, name in self._members():
            t = getattr(mod, name)
            c = sl.stop - sl.start
            want = (lanes, c) + tuple(arr.shape[2:])
            if t is None or t.numel() != lanes * c * (arr[0, 0].numel()):
                raise RuntimeError(f"{type(mod).__name__}.{name} is not initialised for "
                                   f"{lanes} lanes x {c} neurons (call init_net_state)")
            arr[:, sl] = t.reshape(want)

    def scatter(self) -> None:
        """The groups' memories from the population arrays (in place when a
        memory is already contiguous of the right shape)."""
        for arr, sl, mod, name in self._members():
            t = getattr(mod, name)
            src = arr[:, sl]
            if t.is_contiguous() and t.numel() == src.numel():
                t.view(src.shape).copy_(src)
            else:
                setattr(mod, name, src.contiguous().reshape(t.shape))


def match_cell(rnn) -> CellSpec | None:
    """CellSpec when ``rnn`` is a cell the engine runs, else None."""
    neuron, syn = getattr(rnn, "neuron", None), getattr(rnn, "synapse", None)
    neuron_key = "neuron."
    mixed = None
    if isinstance(neuron, MixedNeuronPopulation):
        if len(neuron.counts) == 1:
            # one group spans the whole neuron axis: the population is that neuron
            neuron_key = f"neuron.{neuron._group_names[0]}."
            neuron = neuron.groups()[0]
        else:
            mixed = neuron
    elif isinstance(neuron, TwoCompartmentGLIF):
        return None  # its (spike, v) output is not a synapse input
    if getattr(rnn, "synapse_apical", None) is not None:
        return None
    if mixed is not None:
        groups = mixed.groups()
        if not all(isinstance(g, (LIF, GLIF3, TwoCompartmentGLIF)) for g in groups):
            return None
        nasc = {g.n_Iasc for g in groups if isinstance(g, GLIF3)}
        if len(nasc) > 1 or (nasc and not 1 <= nasc.pop() <= 4):
            return None
        if any(tuple(g.n_neuron) != (g.size,) for g in groups):
            return None
        neuron = MixedState(mixed)
    elif not isinstance(neuron, (LIF, GLIF3, TwoCompartmentGLIF)):
        return None
    if isinstance(neuron, GLIF3) and not 1 <= neuron.n_Iasc <= 4:
        return None
    if getattr(rnn, "syn_inp_module", None) is not None:
        return None
    n = neuron.size
    if isinstance(syn, KINETICS):
        conn = syn.linear
        spec = CellSpec(neuron, syn, conn, None, 1, 1, [(0, 0)], ["synapse.psc"], [syn])
    elif isinstance(syn, DelayedPSC) and isinstance(syn.psc_module, KINETICS):
        conn = syn.psc_module.linear
        if syn.history is None:
            spec = CellSpec(neuron, syn.psc_module, conn, None, 1, 1, [(0, 0)],
                            ["synapse.psc_module.psc"], [syn.psc_module])
        else:
            m = syn.max_delay_steps
            spec = CellSpec(neuron, syn.psc_module, conn, syn.history, 1, m + 1, [(m, 0)],
                            ["synapse.psc_module.psc"], [syn.psc_module])
    elif (isinstance(syn, HeterSynapsePSC) and isinstance(syn.base_psc, KINETICS)
          and syn.n_receptor == 1):
        conn = syn.linear
        keys = ["synapse.psc", "synapse.base_psc.psc"]
        if syn.history is None:
            spec = CellSpec(neuron, syn.base_psc, conn, None, 1, 1, [(0, 0)], keys,
                            [syn, syn.base_psc])
        else:
            d = syn.max_delay_steps
            spec = CellSpec(neuron, syn.base_psc, conn, syn.history, d, d, [], keys,
                            [syn, syn.base_psc])
    else:
        return None
    if not isinstance(conn, BaseSparseConn) or conn.bias is not None:
        return None
    if conn.in_features != n * spec.csr_d or conn.out_features != n:
        return None
    if spec.exp_psc.size != n or tuple(spec.exp_psc.n_neuron) != (n,) and \
            spec.exp_psc.size != n:
        return None
    if spec.csr_d > 1:
        ne = _bin_nonempty(conn, spec.csr_d)
        spec.bins = [(d, d) for d in range(spec.csr_d) if ne[d]] or [(0, 0)]
    if len(spec.bins) > _lib.MAX_BINS:
        return None
    kin = spec.exp_psc
    if not isinstance(kin, ExponentialPSC):
        if isinstance(neuron, (TwoCompartmentGLIF, MixedState)):
            return None
        prefix = spec.psc_keys[-1][: -len("psc")]
        spec.syn_keys = {prefix + k: f"syn_h{j + 1}" for j, k in enumerate(kin._STATES)}
    spec.neuron_key = neuron_key
    return spec


def pack_bits(spikes: torch.Tensor) -> torch.Tensor:
    """(…, n) 0/1 -> (…, ceil(n/64)) int64 words (bit j of word w is 64w+j)."""
    n = spikes.shape[-1]
    nw = (n + 63) // 64
    b = torch.zeros(*spikes.shape[:-1], nw * 64, dtype=torch.int64, device=spikes.device)
    b[..., :n] = (spikes != 0).to(torch.int64)
    b = b.reshape(*spikes.shape[:-1], nw, 64)
    shifts = torch.arange(64, device=spikes.device, dtype=torch.int64)
    return (b << shifts).sum(-1)


def unpack_bits(words: torch.Tensor, n: int) -> torch.Tensor:
    """(…, W) int64 spike words -> (…, n) float32 0/1 (bit j of word w is 64w+j)."""
    shifts = torch.arange(64, device=words.device, dtype=torch.int64)
    bits = (words.unsqueeze(-1) >> shifts) & 1
    return bits.reshape(*words.shape[:-1], words.shape[-1] * 64)[..., :n].to(torch.float32)


_MISSING = object()


def _attr(m, name: str):
    """m.name, or None when m is None or has no such attribute (instance
    entries and nn.Module's parameter, buffer and submodule tables first;
    class attributes such as properties through getattr)."""
    if m is None:
        return None
    d = m.__dict__
    v = d.get(name, _MISSING)
    if v is not _MISSING:
        return v
    for table in ("_modules", "_parameters", "_buffers"):
        t = d.get(table)
        if t is not None and name in t:
            return t[name]
    return getattr(m, name, None) if hasattr(type(m), name) else None


def _tkey(t):
    """Identity and in-place version of a tensor (None passes through)."""
    if t is None or not torch.is_tensor(t):
        return t
    return (id(t), t._version, t.data_ptr())


class _RingState:
    """The device ring a sequence of runs shares: ring_d + 1 slots of spike
    lists, per-segment counts and words, the slot the next run starts at, and
    the (tensor, version, cursor) of the module history it last matched."""

    def __init__(self, ring_d, lanes, n, dev):
        self.idx, self.cnt, self.words, self.lane_bits = Engine._ring_buffers(ring_d, lanes, n,
                                                                              dev)
        self.R = ring_d + 1
        self.cursor = ring_d          # device slot of the next run's first step
        self.token = None             # module history the ring holds (Engine._hist_token), or None
        self.fresh = True             # lists not built yet


class _HistoryWriteback:
    """Writes a SpikeHistory memory from the device ring that holds it: module
    slot j holds delay d(j) (circular: d = (cursor - 1 - j) mod D; cat: d = j),
    and delay d sits in device slot (dev_cursor - 1 - d) mod R."""

    __slots__ = ("engine", "ring", "lanes", "n", "ring_d", "R", "cursor", "dev_cursor")

    def __init__(self, engine, ring, lanes, n, ring_d, R, cursor, dev_cursor):
        self.engine, self.ring, self.lanes, self.n = engine, ring, lanes, n
        self.ring_d, self.R, self.cursor, self.dev_cursor = ring_d, R, cursor, dev_cursor

    def __call__(self, h) -> None:
        lanes, n, ring_d = self.lanes, self.n, self.ring_d
        buf = h._buffers.get("history")
        if buf is None or buf.numel() != lanes * ring_d * n:
            raise RuntimeError("SpikeHistory changed shape while its delay history was held "
                               "on the device")
        if not (buf.is_contiguous() and buf.dtype == torch.float32):
            buf = torch.empty(buf.shape, dtype=torch.float32, device=buf.device)
            h._buffers["history"] = buf
        j = torch.arange(ring_d, dtype=torch.long, device=buf.device)
        d = (self.cursor - 1 - j) % ring_d if h.use_circular_buffer else j
        slots = (self.dev_cursor - 1 - d) % self.R
        words = self.ring.words.index_select(0, slots).transpose(0, 1).contiguous()
        from .._kernels import unpack_rows
        unpack_rows(words.reshape(lanes * ring_d, -1), n, out=buf.view(lanes * ring_d, n))
        if h.use_circular_buffer:
            h._buffers["_cursor"].fill_(self.cursor)
        self.ring.token = self.engine._hist_token(h)


class Engine:
    """Runs a matched cell for T steps; owns the ring and the workspace.

    Across runs it keeps, per lane count, a plan template (parameters,
    coefficients and connectivity, rebuilt when a parameter tensor, the
    weights, dt or the configuration changes), the workspace with its
    delay-bin tables, and the device ring, which the next run continues
    from instead of rebuilding it from the module's dense history."""

    def __init__(self, spec: CellSpec):
        self.spec = spec
        self._ws: dict = {}
        self._tmpl: dict = {}
        self._rings: dict = {}
        self._tables_ok: set = set()

    def _workspace(self, lanes, n, slots, ordered, device):
        key = (lanes, n, slots, ordered, str(device))
        ws = self._ws.get(key)
        if ws is None:
            nbytes = _lib.load().btx_rsnn_workspace_bytes(lanes, n, slots, int(ordered))
            ws = torch.zeros(nbytes, dtype=torch.uint8, device=device)
            self._ws = {key: ws}
            self._tables_ok = set()
        return ws

    # -- plan -----------------------------------------------------------------
    _SCALAR_ATTRS = ("hard_reset", "_use_refractory", "n_Iasc")

    def _static_key(self, lanes: int, dev) -> tuple:
        """What the plan template depends on besides the state tensors."""
        from . import environ
        sp = self.spec
        conn = sp.conn
        cfg = conn.sparse_config
        parts = [lanes, str(dev), float(environ.get("dt")), bool(torch.is_grad_enabled()),
                 int(cfg.ordered), int(cfg.slots), int(getattr(cfg, "bin_cap", 0)),
                 self.dense_pull(),
                 id(conn._csr_row_ptr), id(conn._csc_ptr), conn.weight_key(),
                 id(conn._prepared_sparse_weight)]
        mods = (sp.neuron.modules() if isinstance(sp.neuron, MixedState) else [sp.neuron])
        for m in (*mods, sp.exp_psc):
            d = m.__dict__
            mem = d.get("_memories_rv") or ()
            for name, t in d["_parameters"].items():
                parts.append((name, _tkey(t)))
            for name, t in d["_buffers"].items():
                if name not in mem:
                    parts.append((name, _tkey(t)))
            for name in self._SCALAR_ATTRS:
                parts.append(_attr(m, name))
        return tuple(parts)

    def dense_pull(self) -> bool:
        """Whether runs hand the kernel the ring's lane bits (and with them
        the dense pull of busy steps)."""
        return bool(getattr(self.spec.conn.sparse_config, "dense_pull", True))

    def _template(self, lanes: int):
        """(template plan, tensors it points into): parameters,
        coefficients, connectivity and bins, cached per lane count."""
        sp = self.spec
        neuron, syn, conn = sp.neuron, sp.exp_psc, sp.conn
        dev = neuron.v.device
        key = self._static_key(lanes, dev)
        hit = self._tmpl.get(lanes)
        if hit is not None and hit[0] == key:
            return hit[1], hit[2]
        keep: list = []
        cfg = conn.sparse_config
        if isinstance(neuron, MixedState):
            p = _plan.mixed_plan(neuron, lanes, dev, keep)
        else:
            p = _plan.neuron_plan(neuron, lanes, dev, keep)
        p.ordered = int(cfg.ordered)
        p.slots = int(cfg.slots)
        p.bin_cap = int(getattr(cfg, "bin_cap", 0))
        syn.synapse_plan(p, dev, keep)
        w = conn.packed_weight().detach()
        if w.dtype != torch.float32:
            raise TypeError("the fused engine supports float32 weights only")
        keep.append(w)
        p.row_ptr = conn._csr_row_ptr.data_ptr()
        p.col = conn._csr_col.data_ptr()
        p.w = w.data_ptr()
        p.nnz = conn.nnz
        p.csr_d = sp.csr_d
        w_dst = conn.effective_weight_f32()
        keep.append(w_dst)
        keep += [conn._csr_row_ptr, conn._csr_col, conn._csc_ptr, conn._csc_src]
        p.csc_ptr = conn._csc_ptr.data_ptr()
        p.csc_src = conn._csc_src.data_ptr()
        p.csc_w = w_dst.data_ptr()
        p.ring_d = sp.ring_d
        p.nbins = len(sp.bins)
        for j, (age, sub) in enumerate(sp.bins):
            p.bin_delay[j] = age
            p.bin_sub[j] = sub
        self._tmpl[lanes] = (key, p, keep)
        return p, keep

    def _base_plan(self, lanes: int, keep: list):
        """A plan for one run: the cached template plus the state pointers
        (checked every run: a module may be given new state tensors)."""
        sp = self.spec
        neuron, syn = sp.neuron, sp.exp_psc
        n = neuron.size
        dev = neuron.v.device
        _lib.require_hip(neuron.v)
        tmpl, tkeep = self._template(lanes)
        keep += tkeep
        p = _lib.RsnnPlan()
        ctypes.memmove(ctypes.addressof(p), ctypes.addressof(tmpl), ctypes.sizeof(p))
        states = [(neuron, "v"), (syn, "psc")] + [(syn, k) for k in getattr(syn, "_STATES", ())]
        tc_state = isinstance(neuron, TwoCompartmentGLIF) or getattr(neuron, "has_tc", False)
        if tc_state:
            states += [(neuron, k) for k in ("i_a", "i_bap", "theta_th")]
        for mod, name in states:
            t = getattr(mod, name)
            if t is None or t.numel() != lanes * n or t.dtype != torch.float32:
                raise RuntimeError(f"{type(mod).__name__}.{name} is not initialised for "
                                   f"{lanes} lanes x {n} neurons (call init_net_state)")
            if not t.is_contiguous():
                setattr(mod, name, t.contiguous())
        p.v = neuron.v.data_ptr()
        p.psc = syn.psc.data_ptr()
        keep += [neuron.v, syn.psc]
        for j, k in enumerate(getattr(syn, "_STATES", ())):
            t = getattr(syn, k)
            setattr(p, f"syn_h{j + 1}", t.data_ptr())
            keep.append(t)
        if tc_state:
            for k in ("i_a", "i_bap", "theta_th"):
                t = getattr(neuron, k)
                setattr(p, k, t.data_ptr())
                keep.append(t)
        if isinstance(neuron, MixedState) and neuron.Iasc is not None:
            p.iasc = neuron.Iasc.data_ptr()
            keep.append(neuron.Iasc)
        if isinstance(neuron, GLIF3):
            m = neuron.n_Iasc
            ia = neuron.Iasc
            if ia is None or ia.numel() != lanes * n * m or ia.dtype != torch.float32:
                raise RuntimeError(f"GLIF3.Iasc is not initialised for {lanes} lanes x {n} "
                                   "neurons")
            if not ia.is_contiguous():
                neuron.Iasc = ia = ia.contiguous()
            p.iasc = ia.data_ptr()
            keep.append(ia)
        if isinstance(neuron, MixedState):
            if neuron._use_refractory:
                p.refractory = neuron.refractory.data_ptr()
                keep.append(neuron.refractory)
        elif getattr(neuron, "_use_refractory", False):
            ref = _plan.lane_state(neuron, "refractory", tuple(neuron.v.shape))
            p.refractory = ref.data_ptr()
            keep.append(ref)
        return p, n, dev

    @staticmethod
    def _ring_buffers(ring_d, lanes, n, dev):
        """Device ring of ring_d + 1 slots: lists, counts per segment, words,
        and the lane bits of the dense pull ((R, n, ceil(lanes/64)) words,
        bit b of word (slot, j, b // 64) = lane b pushed neuron j)."""
        R = ring_d + 1
        nwords = (n + 63) // 64
        nchunks = (n + _lib.RING_CHUNK - 1) // _lib.RING_CHUNK
        return (torch.empty((R, lanes, n), dtype=torch.int32, device=dev),
                torch.zeros((R, lanes, nchunks), dtype=torch.int32, device=dev),
                torch.zeros((R, lanes, nwords), dtype=torch.int64, device=dev),
                torch.zeros((R, n, (lanes + 63) // 64), dtype=torch.int64, device=dev))

    @staticmethod
    def clear_lane_bits(lane_bits: torch.Tensor, sl: slice) -> None:
        """Clear lanes ``sl`` from a ring's lane bits (lanes emptied of a job)."""
        lanes = range(*sl.indices(lane_bits.shape[-1] * 64))
        for g in range(lane_bits.shape[-1]):
            mask = 0
            for b in lanes:
                if b // 64 == g:
                    mask |= 1 << (b % 64)
            if mask:
                keep = ~mask & 0xFFFFFFFFFFFFFFFF
                keep = keep - (1 << 64) if keep >= 1 << 63 else keep  # as int64
                lane_bits[..., g].bitwise_and_(keep)

    @staticmethod
    def _set_outputs(p, T, lanes, n, drive, counts, sample):
        if drive is not None:
            for key, dt_ in (("seed", torch.int64), ("prob", torch.float32), ("t0", torch.int64)):
                t_ = drive[key]
                if t_.shape != (lanes,) or t_.dtype != dt_ or not t_.is_contiguous():
                    raise ValueError(f"drive[{key!r}] must be contiguous {dt_} of shape ({lanes},)")
            p.in_seed = drive["seed"].data_ptr()
            p.in_prob = drive["prob"].data_ptr()
            p.in_t0 = drive["t0"].data_ptr()
            p.in_kick = float(drive["kick"])
            if drive.get("t_end") is not None:
                te = drive["t_end"]
                if te.shape != (lanes,) or te.dtype != torch.int64 or not te.is_contiguous():
                    raise ValueError("drive['t_end'] must be contiguous int64 (lanes,)")
                p.lane_t_end = te.data_ptr()
        if counts is not None:
            if counts.shape != (lanes, n) or counts.dtype != torch.int32 \
                    or not counts.is_contiguous():
                raise ValueError("counts must be contiguous int32 (lanes, n)")
            p.counts = counts.data_ptr()
        if sample is not None:
            if sample.dim() != 3 or sample.shape[:2] != (T, lanes) or not sample.is_contiguous() \
                    or sample.dtype != torch.float32 or not 1 <= sample.shape[2] <= n:
                raise ValueError("sample must be contiguous float32 (T, lanes, S), 1 <= S <= n")
            p.rec_sample = sample.data_ptr()
            p.n_sample = sample.shape[2]

    # -- delay ring -----------------------------------------------------------
    def _ring(self, lanes, n, dev) -> _RingState:
        key = (lanes, str(dev))
        ring = self._rings.get(key)
        if ring is None:
            ring = _RingState(self.spec.ring_d, lanes, n, dev)
            self._rings = {key: ring}
        return ring

    @staticmethod
    def _hist_token(hist):
        """Identifies the module's history memory as it is now: the buffer
        itself (by weak reference: an id can be reused once it is freed), its
        version counter and the cursor."""
        buf = hist._buffers.get("history")
        return None if buf is None else (weakref.ref(buf), buf._version, hist.cursor)

    @staticmethod
    def _holds(token, hist) -> bool:
        """Whether `token` still names the module's current history memory."""
        if token is None:
            return False
        buf = hist._buffers.get("history")
        ref, version, cursor = token
        return buf is not None and ref() is buf and buf._version == version and \
            hist.cursor == cursor

    def _prepare_ring(self, ring: _RingState, lanes, n, dev, keep) -> int:
        """Make the ring hold the module's history; returns ring_rebuild."""
        sp = self.spec
        hist = sp.history
        ring_d = sp.ring_d
        if hist is None:
            # no delay memory: only slots this run writes are read
            rebuild = 1 if ring.fresh else 0
            ring.fresh = False
            return rebuild
        fn = hist.deferred_writeback()
        if fn is not None and getattr(fn, "ring", None) is ring:
            # the module's history is the state this ring left: continue
            hist.discard_writeback()
            ring.token = None
            return 0
        if fn is not None:
            hist._sync_memories()  # another engine's deferred state
        held = not ring.fresh and self._holds(ring.token, hist)
        # the run moves the ring on: it no longer holds this memory
        ring.token = None
        if held:
            return 0
        # rebuild from the dense memory: device slot ring_d - 1 - d holds
        # delay d; the run starts at slot ring_d
        h = hist.history
        if h is None or h.numel() != lanes * ring_d * n:
            raise RuntimeError("SpikeHistory is not initialised for this batch")
        h = h.reshape(lanes, ring_d, n).to(torch.float32).contiguous()
        keep.append(h)
        cursor = hist.cursor
        j = torch.arange(ring_d, dtype=torch.long, device=dev)
        order = (j + cursor) % ring_d if hist.use_circular_buffer else ring_d - 1 - j
        from .._kernels import pack_rows
        words = pack_rows(h.reshape(lanes * ring_d, n)).reshape(lanes, ring_d, -1)
        ring.words[:ring_d] = words.index_select(1, order).transpose(0, 1)
        ring.words[ring_d:] = 0
        ring.cnt.zero_()
        ring.cursor = ring_d
        ring.fresh = False
        return 1

    def _defer_history(self, ring: _RingState, lanes, n, T) -> None:
        """After a run: the module's history is what the ring holds; write it
        back only when it is read."""
        hist = self.spec.history
        if hist is None:
            return
        cursor = (hist.cursor + T) % hist.max_delay if hist.use_circular_buffer else 0
        ring_d, R = self.spec.ring_d, ring.R
        fn = _HistoryWriteback(self, ring, lanes, n, ring_d, R, cursor, ring.cursor)
        hist.defer_writeback(fn, cursor)

    # -- run --------------------------------------------------------------------
    def run(self, x: torch.Tensor | None, T: int, lanes: int, record: set[str],
            teacher: torch.Tensor | None = None, timings: dict | None = None,
            drive: dict | None = None, counts: torch.Tensor | None = None,
            sample: torch.Tensor | None = None, x_apical: torch.Tensor | None = None):
        """x: (T, lanes, n) fp32 contiguous or None.  Returns (raster, records).
        x_apical: (T, lanes, n) apical input of a two-compartment neuron.

        drive: {"seed": (lanes,) int64, "prob": (lanes,) f32, "t0": (lanes,) int64,
        "kick": float, optional "t_end": (lanes,) int64} in-kernel Poisson input;
        counts: (lanes, n) int32 accumulated spike counts; sample: (T, lanes, S)
        f32 membrane of neurons 0..S-1 each step."""
        sp = self.spec
        keep: list = []
        mixed = isinstance(sp.neuron, MixedState)
        if mixed:
            sp.neuron.gather(lanes)
        p, n, dev = self._base_plan(lanes, keep)
        if x is not None:
            if x.shape != (T, lanes, n) or not x.is_contiguous() or x.dtype != torch.float32:
                raise ValueError("engine input must be contiguous float32 (T, lanes, n)")
            p.x = x.data_ptr()
            p.x_step_stride = lanes * n
        if x_apical is not None:
            if p.neuron_model != _lib.NEURON_TWO_COMPARTMENT and not (
                    mixed and sp.neuron.has_tc):
                raise ValueError("an apical input needs a two-compartment neuron")
            if x_apical.shape != (T, lanes, n) or not x_apical.is_contiguous() \
                    or x_apical.dtype != torch.float32:
                raise ValueError("apical input must be contiguous float32 (T, lanes, n)")
            p.x_apical = x_apical.data_ptr()
            p.x_apical_step_stride = lanes * n
        ring = self._ring(lanes, n, dev)
        p.ring_rebuild = self._prepare_ring(ring, lanes, n, dev, keep)
        p.ring_idx = ring.idx.data_ptr()
        p.ring_cnt = ring.cnt.data_ptr()
        p.ring_words = ring.words.data_ptr()
        p.ring_lane_bits = ring.lane_bits.data_ptr() if self.dense_pull() else None
        p.ring_cursor = ring.cursor
        self._set_outputs(p, T, lanes, n, drive, counts, sample)
        nwords = (n + 63) // 64
        raster = torch.empty((T, lanes, nwords), dtype=torch.int64, device=dev)
        p.raster = raster.data_ptr()
        if teacher is not None:
            if teacher.shape != (T, lanes, nwords) or teacher.dtype != torch.int64 \
                    or not teacher.is_contiguous():
                raise ValueError("teacher must be contiguous int64 (T, lanes, ceil(n/64))")
            p.teacher = teacher.data_ptr()
        recs = {}
        # the spike record is not written by the step: it is the raster's
        # bits, unpacked after the run by one streaming kernel, so the steps
        # move 4 B less per neuron-lane (and without other records run the
        # record-free variant)
        want_z = "z" in record
        for key in ("v", "refractory", "psc", "v_pre", "v_prev", "i_a", "i_bap",
                    "theta_th", "syn_h1", "syn_h2"):
            if key in record:
                recs[key] = torch.empty((T, lanes, n), dtype=torch.float32, device=dev)
        for key in ("i_a", "i_bap", "theta_th"):
            if key in recs:
                if p.neuron_model != _lib.NEURON_TWO_COMPARTMENT and not (
                        mixed and sp.neuron.has_tc):
                    raise ValueError(f"'{key}' is a two-compartment state")
                setattr(p, "rec_" + key, recs[key].data_ptr())
        p.rec_z = None
        p.rec_v = recs["v"].data_ptr() if "v" in recs else None
        p.rec_refractory = recs["refractory"].data_ptr() if "refractory" in recs else None
        p.rec_psc = recs["psc"].data_ptr() if "psc" in recs else None
        p.rec_v_pre = recs["v_pre"].data_ptr() if "v_pre" in recs else None
        p.rec_v_prev = recs["v_prev"].data_ptr() if "v_prev" in recs else None
        for key in ("syn_h1", "syn_h2"):
            if key in recs:
                if key not in sp.syn_keys.values():
                    raise ValueError(f"'{key}' is not a state of this synapse")
                setattr(p, "rec_" + key, recs[key].data_ptr())
        ws = self._workspace(lanes, n, p.slots, p.ordered, dev)
        p.workspace = ws.data_ptr()
        p.workspace_bytes = ws.numel()
        tables = (ws.data_ptr(), tuple(sp.bins))
        p.tables_ready = 1 if tables in self._tables_ok else 0
        stream = _lib.stream_ptr()
        if timings is None:
            _lib.check(_lib.load().btx_rsnn_run(p, T, stream), "btx_rsnn_run")
        else:
            ms = (ctypes.c_double * 4)()
            calls = (ctypes.c_int64 * 4)()
            cap = T + 1
            per = (ctypes.c_double * (4 * cap))()
            _lib.check(_lib.load().btx_rsnn_run_timed(p, T, stream, ms, calls, per, cap),
                       "btx_rsnn_run_timed")
            for k, name in enumerate(("step", "reduce", "push", "finish")):
                timings[name] = (float(ms[k]), int(calls[k]))
                timings[name + "_each"] = [per[k * cap + i] for i in range(min(calls[k], cap))]
        self._tables_ok.add(tables)
        if want_z:
            from .._kernels import unpack_rows
            z = torch.empty((T, lanes, n), dtype=torch.float32, device=dev)
            unpack_rows(raster.view(T * lanes, nwords), n, out=z.view(T * lanes, n))
            recs["z"] = z
        if mixed:
            sp.neuron.scatter()
        ring.cursor = (ring.cursor + T) % ring.R
        self._defer_history(ring, lanes, n, T)
        return raster, recs

    def resident(self, lanes: int, T: int, drive: dict | None = None,
                 counts: torch.Tensor | None = None,
                 sample: torch.Tensor | None = None) -> "ResidentRun":
        """A fixed launch of T steps over `lanes` lanes whose delay ring stays
        on the device between launches (the module's history memory is not
        read or written).  Every pointer is fixed, so a launch can be captured
        in a HIP graph.  T must be a multiple of the ring depth, so that every
        launch starts at the same ring slot."""
        return ResidentRun(self, lanes, T, drive, counts, sample)

    def history_sequence(self, raster, hist0, T, lanes, n, cursor):
        """(T, lanes, D, n) history memory after every step (for recording)."""
        D = self.spec.ring_d
        hist = self.spec.history
        z = unpack_bits(raster, n)                   # (T, lanes, n)
        init = hist0.reshape(lanes, D, n)
        out = torch.empty((T, lanes, D, n), dtype=torch.float32, device=z.device)
        for t in range(T):
            for j in range(D):
                if hist.use_circular_buffer:
                    back = (t - (j - cursor)) % D     # steps since slot j was written
                    ts = t - back
                    out[t, :, j] = z[ts] if ts >= 0 else init[:, j]
                else:
                    ts = t - j
                    out[t, :, j] = z[ts] if ts >= 0 else init[:, j - t - 1]
        return out


class ResidentRun:
    """T steps of a matched cell with a device-resident delay ring (see
    ``Engine.resident``).  ``launch()`` enqueues one run on the current
    stream; ``reset_lanes(sl)`` empties the ring of a lane range (a new job);
    ``raster`` holds the spike words of the last launch."""

    def __init__(self, engine: Engine, lanes: int, T: int, drive, counts, sample):
        sp = engine.spec
        if isinstance(sp.neuron, MixedState):
            raise NotImplementedError("a resident run holds single-model state; a multi-group "
                                      "MixedNeuronPopulation runs through Engine.run")
        if T % (sp.ring_d + 1):
            raise ValueError(f"T ({T}) must be a multiple of the device ring's "
                             f"{sp.ring_d + 1} slots")
        self.keep: list = []
        p, n, dev = engine._base_plan(lanes, self.keep)
        # the state this run owns (the module may later be given other tensors)
        neuron = sp.neuron
        self.v = neuron.v
        self.refractory = (neuron.refractory if getattr(neuron, "_use_refractory", False)
                           else None)
        self.tc_state = ([neuron.i_a, neuron.i_bap, neuron.theta_th]
                         if isinstance(neuron, TwoCompartmentGLIF) else [])
        self.psc = sp.exp_psc.psc
        self.syn_state = [getattr(sp.exp_psc, k) for k in getattr(sp.exp_psc, "_STATES", ())]
        self.iasc = getattr(neuron, "Iasc", None)
        self.v_reset = neuron.v_reset
        self.idx, self.cnt, self.words, self.lane_bits = engine._ring_buffers(sp.ring_d, lanes,
                                                                              n, dev)
        self.keep += [self.idx, self.cnt, self.words, self.lane_bits]
        p.ring_idx = self.idx.data_ptr()
        p.ring_cnt = self.cnt.data_ptr()
        p.ring_words = self.words.data_ptr()
        p.ring_lane_bits = self.lane_bits.data_ptr() if engine.dense_pull() else None
        p.ring_cursor = 0
        p.ring_rebuild = 0  # lists stay current between launches (all empty at start)
        engine._set_outputs(p, T, lanes, n, drive, counts, sample)
        self.raster = torch.empty((T, lanes, (n + 63) // 64), dtype=torch.int64, device=dev)
        p.raster = self.raster.data_ptr()
        self.ws = torch.zeros(_lib.load().btx_rsnn_workspace_bytes(lanes, n, p.slots, p.ordered),
                              dtype=torch.uint8, device=dev)
        p.workspace = self.ws.data_ptr()
        p.workspace_bytes = self.ws.numel()
        self.plan, self.T, self.lanes = p, T, lanes

    def launch(self) -> None:
        _lib.check(_lib.load().btx_rsnn_run(self.plan, self.T, _lib.stream_ptr()),
                   "btx_rsnn_run")

    def reset_lanes(self, sl: slice) -> None:
        """Lanes `sl` at rest: v = v_reset, no refractory time, no current,
        empty delay ring."""
        self.v.view(self.lanes, -1)[sl] = self.v_reset.to(self.v.dtype)
        if self.refractory is not None:
            self.refractory.view(self.lanes, -1)[sl] = 0.0
        self.psc.view(self.lanes, -1)[sl] = 0.0
        for t in self.syn_state:
            t.view(self.lanes, -1)[sl] = 0.0
        if self.iasc is not None:
            self.iasc.view(self.lanes, -1)[sl] = 0.0
        for t in self.tc_state:
            t.view(self.lanes, -1)[sl] = 0.0
        self.cnt[:, sl] = 0
        self.words[:, sl] = 0
        Engine.clear_lane_bits(self.lane_bits, sl)
