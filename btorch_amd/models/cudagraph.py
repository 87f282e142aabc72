"""HIP-graph replay of the fused engine for ``RecurrentNN(cudagraph=True)``
(the reference's CudaGraphRunner, btorch/models/cudagraph.py:101-242, and
rnn.py:355-405).

A call of T steps is T + 2 kernel launches enqueued from C; what a graph
removes is the host work around them (plan assembly, checks, Python).  One
graph is captured per input signature (shape, dtype, device of x and the set
of recorded states) and replayed on later calls:

* The graph owns static buffers: the input sequence, the state (v, psc,
  refractory and the other memories the step updates in place), the delay
  ring, the workspace and the outputs (spike words and records).  Replay
  copies x and, when the module holds other tensors than the graph's, the
  state into them by value; afterwards the module's memories are the graph's
  state tensors, so a following call copies nothing.
* The delay ring starts every replay in one canonical order (delay d in
  slot ring_d - 1 - d, the run starting at slot ring_d): the graph ends by
  rotating its spike words back into that order, and begins by rebuilding the
  lists from the words.  The module's history memory is written back from
  the ring only when read, as in eager calls; a history the caller replaced
  is packed into the ring before the replay.
* Outputs are cloned out of the graph's memory, so a later replay does not
  rewrite tensors the caller holds.

Inference only, as in the reference: a call that would record an autograd
graph, ``grad_checkpoint`` and ``save_grad_history`` raise.  The input
projection (``neuron_inp_module``) runs eagerly before the replay.
"""

from __future__ import annotations

import ctypes

import torch

from .. import _lib
from .neurons.glif import GLIF3
from .neurons.two_compartment import TwoCompartmentGLIF


def _state_slots(spec) -> list[tuple[object, str]]:
    """(module, memory name) of every state the fused step updates in place."""
    neuron, syn = spec.neuron, spec.exp_psc
    slots = [(neuron, "v"), (syn, "psc")] + [(syn, k) for k in getattr(syn, "_STATES", ())]
    if getattr(neuron, "_use_refractory", False):
        slots.append((neuron, "refractory"))
    if isinstance(neuron, GLIF3):
        slots.append((neuron, "Iasc"))
    if isinstance(neuron, TwoCompartmentGLIF):
        slots += [(neuron, k) for k in ("i_a", "i_bap", "theta_th")]
    return slots


class _GraphEntry:
    """One captured run: static buffers, plan and graph."""

    def __init__(self, eng, T: int, lanes: int, record: frozenset, x: torch.Tensor):
        from ._engine import _RingState

        sp = eng.spec
        n = sp.neuron.size
        dev = x.device
        self.T, self.lanes, self.record = T, lanes, record
        self.keep: list = []
        self.x = torch.empty_like(x)
        self.slots = _state_slots(sp)
        # the graph's own state tensors, seeded with the module's
        self.state = [getattr(m, k).detach().clone().contiguous() for m, k in self.slots]
        for (m, k), t in zip(self.slots, self.state):
            setattr(m, k, t)
        p, _, _ = eng._base_plan(lanes, self.keep)
        # the plan points at the module's tensors after its checks
        self.state = [getattr(m, k) for m, k in self.slots]
        self.ring = _RingState(sp.ring_d, lanes, n, dev)
        self.ring.fresh = False
        R = self.ring.R
        ring_d = sp.ring_d
        # after a run the step t = T - 1 wrote slot (ring_d + T - 1) mod R:
        # delay d sits in (ring_d + T - 1 - d) mod R and goes to ring_d - 1 - d
        d = torch.arange(ring_d, dtype=torch.long, device=dev)
        self.rot_src = (ring_d + T - 1 - d) % R
        self.rot_dst = ring_d - 1 - d
        self.tmp_words = torch.zeros_like(self.ring.words)
        p.x = self.x.data_ptr()
        p.x_step_stride = lanes * n
        p.ring_idx = self.ring.idx.data_ptr()
        p.ring_cnt = self.ring.cnt.data_ptr()
        p.ring_words = self.ring.words.data_ptr()
        p.ring_lane_bits = self.ring.lane_bits.data_ptr() if eng.dense_pull() else None
        p.ring_cursor = ring_d
        p.ring_rebuild = 1
        p.tables_ready = 0
        nwords = (n + 63) // 64
        self.raster = torch.empty((T, lanes, nwords), dtype=torch.int64, device=dev)
        p.raster = self.raster.data_ptr()
        self.recs = {}
        for key in record:
            self.recs[key] = torch.empty((T, lanes, n), dtype=torch.float32, device=dev)
        # the spike record is the raster's bits, unpacked in the graph after
        # the run (as Engine.run does)
        for key, field in (("v", "rec_v"), ("refractory", "rec_refractory"),
                           ("psc", "rec_psc"), ("v_pre", "rec_v_pre"), ("v_prev", "rec_v_prev"),
                           ("i_a", "rec_i_a"), ("i_bap", "rec_i_bap"),
                           ("theta_th", "rec_theta_th"), ("syn_h1", "rec_syn_h1"),
                           ("syn_h2", "rec_syn_h2")):
            if key in self.recs:
                setattr(p, field, self.recs[key].data_ptr())
        self.ws = torch.zeros(_lib.load().btx_rsnn_workspace_bytes(lanes, n, p.slots, p.ordered),
                              dtype=torch.uint8, device=dev)
        p.workspace = self.ws.data_ptr()
        p.workspace_bytes = self.ws.numel()
        self.plan = p
        self.graph = None

    def _device_work(self) -> None:
        _lib.check(_lib.load().btx_rsnn_run(self.plan, self.T, _lib.stream_ptr()),
                   "btx_rsnn_run")
        if "z" in self.recs:
            from .._kernels import unpack_rows
            n = self.raster.shape[0] * self.raster.shape[1]
            unpack_rows(self.raster.view(n, -1), self.recs["z"].shape[-1],
                        out=self.recs["z"].view(n, -1))
        # slot ring_d of tmp_words stays zero: the next run writes it first
        w = self.ring.words
        self.tmp_words.index_copy_(0, self.rot_dst, w.index_select(0, self.rot_src))
        w.copy_(self.tmp_words)

    def capture(self) -> None:
        g = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            with torch.cuda.graph(g, stream=side):
                self._device_work()
        torch.cuda.current_stream().wait_stream(side)
        self.graph = g


class GraphRunner:
    """Captured runs of one RecurrentNN's fused engine, keyed by signature."""

    def __init__(self, warmup: int = 3):
        self.warmup = warmup
        self.entries: dict[tuple, _GraphEntry] = {}

    def reset(self) -> None:
        """Drop captured graphs and their buffers."""
        self.entries.clear()

    @staticmethod
    def check_supported(rnn, x: torch.Tensor) -> None:
        if x.device.type != "cuda":
            raise RuntimeError(f"cudagraph=True requires inputs on a HIP device, got {x.device}.")
        if torch.is_grad_enabled() and (
                x.requires_grad or any(p.requires_grad for p in rnn.parameters())):
            raise RuntimeError(
                "cudagraph=True is inference-only: each call rewrites the graph's static "
                "buffers in place, which would invalidate an autograd graph recorded over "
                "them. Run under torch.no_grad() / torch.inference_mode().")
        for attr, why in (("grad_checkpoint", "it only matters for a backward pass"),
                          ("save_grad_history", "its hooks only fire in the backward pass")):
            if getattr(rnn, attr, False):
                raise RuntimeError(f"cudagraph=True is incompatible with {attr}=True ({why}).")

    def run(self, eng, xs: torch.Tensor, record: set[str]):
        """Replay (capturing first when needed) T = xs.shape[0] steps of
        ``eng`` over xs (T, lanes, n); returns (raster, records) cloned."""
        sp = eng.spec
        from ._engine import MixedState
        if isinstance(sp.neuron, MixedState):
            raise NotImplementedError("cudagraph=True replays single-model populations; a "
                                      "multi-group MixedNeuronPopulation runs eagerly")
        T, lanes = xs.shape[0], xs.shape[1]
        key = (tuple(xs.shape), xs.dtype, str(xs.device), frozenset(record),
               eng._static_key(lanes, xs.device))
        entry = self.entries.get(key)
        if entry is None:
            entry = self._build(eng, xs, T, lanes, frozenset(record))
            self.entries[key] = entry
        entry.x.copy_(xs)
        for (m, k), t in zip(entry.slots, entry.state):
            cur = getattr(m, k)
            if cur is not t:
                if cur is None or cur.numel() != t.numel():
                    raise RuntimeError(f"{type(m).__name__}.{k} changed shape since capture; "
                                       "call reset() on the runner")
                t.copy_(cur.reshape(t.shape))
                setattr(m, k, t)
        self._load_history(eng, entry)
        entry.graph.replay()
        entry.ring.cursor = sp.ring_d  # the graph leaves the ring in canonical order
        eng._defer_history(entry.ring, lanes, sp.neuron.size, T)
        return entry.raster.clone(), {k: v.clone() for k, v in entry.recs.items()}

    def _build(self, eng, xs, T, lanes, record) -> _GraphEntry:
        sp = eng.spec
        # snapshot by value: warm-up and capture advance the module
        slots = _state_slots(sp)
        snap = [getattr(m, k).detach().clone() for m, k in slots]
        hist = sp.history
        hsnap = None
        if hist is not None:
            hsnap = (hist.history.detach().clone(), hist.cursor)
        # warm-up runs settle the allocator and the plan cache
        for _ in range(max(0, self.warmup)):
            eng.run(xs, T, lanes, set(record))
        for (m, k), t in zip(slots, snap):
            setattr(m, k, t.clone())
        if hsnap is not None:
            hist.history = hsnap[0].clone()
            hist.set_cursor(hsnap[1])
        entry = _GraphEntry(eng, T, lanes, record, xs)
        torch.cuda.synchronize()
        entry.capture()
        torch.cuda.synchronize()
        # capture records without running: the graph's state is the seeded
        # entry state; the ring is loaded from the module before each replay
        for (m, k), t, s in zip(entry.slots, entry.state, snap):
            t.copy_(s.reshape(t.shape))
            setattr(m, k, t)
        entry.ring.token = None
        return entry

    @staticmethod
    def _load_history(eng, entry: _GraphEntry) -> None:
        """Put the module's history into the entry's ring (canonical order)
        unless the ring already holds it."""
        sp = eng.spec
        hist = sp.history
        ring = entry.ring
        if hist is None:
            return
        fn = hist.deferred_writeback()
        if fn is not None and getattr(fn, "ring", None) is ring:
            hist.discard_writeback()
            return
        if fn is not None:
            hist._sync_memories()
        held = eng._holds(ring.token, hist)
        # the replay moves the ring on: it no longer holds this memory
        ring.token = None
        if held:
            return
        lanes, n, ring_d = entry.lanes, sp.neuron.size, sp.ring_d
        h = hist.history
        if h is None or h.numel() != lanes * ring_d * n:
            raise RuntimeError("SpikeHistory is not initialised for this batch")
        h = h.reshape(lanes, ring_d, n).to(torch.float32).contiguous()
        j = torch.arange(ring_d, dtype=torch.long, device=h.device)
        cursor = hist.cursor
        order = (j + cursor) % ring_d if hist.use_circular_buffer else ring_d - 1 - j
        from .._kernels import pack_rows
        words = pack_rows(h.reshape(lanes * ring_d, n)).reshape(lanes, ring_d, -1)
        ring.words[:ring_d] = words.index_select(1, order).transpose(0, 1)
        ring.words[ring_d:] = 0
        ring.token = None


__all__ = ["GraphRunner"]
