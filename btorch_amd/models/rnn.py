"""Time-unrolled recurrent loop (btorch/models/rnn.py:27-528).

``RecurrentNN(neuron, synapse, syn_inp_module=None, neuron_inp_module=None,
*, update_state_names=None, unroll=8, chunk_size=None, cpu_offload=False,
grad_checkpoint=False, allow_buffer=False)``.  ``model(x_seq)`` in step mode
"m" returns ``(z_seq, states)`` with ``z_seq`` (T, *batch, N) and ``states``
a dict of stacked per-step hidden states (all of them when
``update_state_names`` is None), exactly as the reference.

Cells the fused engine supports (see ``_engine``) run their whole sequence
in the HIP library; with gradients enabled they run chunk by chunk through
the hand-written adjoint with recompute checkpointing every ``chunk_size``
steps.  Other cells run the per-step loop of ``single_step_forward``.
"""

from __future__ import annotations

from collections.abc import Sequence

import torch
import torch.nn as nn

from . import base
from ._engine import Engine, MixedState, _attr, match_cell
from .functional import filter_hidden_states, hidden_state_names, prepare_sparse_modules
from .linear import BaseSparseConn


class RecurrentNNAbstract(base.MemoryModule):
    """Loop driver; subclasses define ``single_step_forward -> (z, states)``."""

    def __init__(self, update_state_names: Sequence[str] | None = None, step_mode="m",
                 unroll: int | bool = 8, chunk_size: int | None = None,
                 cpu_offload: bool = False, grad_checkpoint: bool = False,
                 save_grad_history: bool = False, grad_state_names=None,
                 cudagraph: bool = False, cudagraph_warmup: int = 3):
        super().__init__()
        self.step_mode = step_mode
        self.update_state_names = update_state_names
        self.unroll = unroll
        self.chunk_size = chunk_size
        self.cpu_offload = cpu_offload
        self.grad_checkpoint = grad_checkpoint
        self.save_grad_history = save_grad_history
        self.grad_state_names = grad_state_names
        self._grad_history: dict[str, list] = {}
        self.cudagraph = cudagraph
        self.cudagraph_warmup = cudagraph_warmup
        self._graph_runner = None

    def single_step_forward(self, *args, **kwargs):
        raise NotImplementedError

    # -- gradient history (rnn.py:86-104, 319-326, 407-413) ------------------
    def _should_save_grad(self, state_name: str) -> bool:
        if not self.save_grad_history:
            return False
        return self.grad_state_names is None or state_name in self.grad_state_names

    def get_grad_history(self) -> dict[str, list]:
        """dL/d(state after step t) of every saved state, filled by backward."""
        return self._grad_history

    def clear_grad_history(self) -> None:
        self._grad_history = {}

    def _hook_grad_history(self, states: dict[str, list], T: int) -> None:
        """Hooks on the per-step state tensors that record their gradient."""
        for name, tensors in states.items():
            if not self._should_save_grad(name):
                continue
            slots = self._grad_history.setdefault(name, [None] * T)
            for t, tensor in enumerate(tensors):
                if torch.is_tensor(tensor) and tensor.requires_grad:
                    def hook(g, slots=slots, t=t):
                        slots[t] = g.detach().clone()
                        return g
                    tensor.register_hook(hook)

    def _detect_loop_args(self, *args):
        if len(args) == 1:
            return args[0].shape[0], (0,)
        shapes = [a.shape[0] if torch.is_tensor(a) and a.ndim > 0 else None for a in args]
        T = shapes[0]
        assert T is not None
        return T, tuple(i for i, s in enumerate(shapes) if s == T)

    def _chunk_plan(self, *args, loop_args=None):
        if loop_args is None:
            T, loop_args = self._detect_loop_args(*args)
        else:
            T = args[loop_args[0]].shape[0]
        unroll = T if self.unroll is False else int(self.unroll)
        if self.chunk_size is None:
            large = unroll if self.grad_checkpoint else T
        else:
            large = self.chunk_size
            if self.unroll is not False and large % unroll != 0:
                raise ValueError(f"chunk_size ({large}) must be a multiple of unroll ({unroll})")
        return T, loop_args, unroll, large, (T + large - 1) // large

    def multi_step_forward(self, *args, loop_args=None, **kwargs):
        first = args[0]
        if self._runs_fused_without_grad(*args, **kwargs):
            # the fused engine keeps its own packed weights (rebuilt when a
            # weight changes); the per-run packing is for the step loop
            return self._multi_step_loop(*args, loop_args=loop_args, **kwargs)
        batch = (first[0].numel() // first.shape[-1]
                 if torch.is_tensor(first) and first.ndim >= 2 else None)
        with prepare_sparse_modules(self, batch_size=batch):
            return self._multi_step_loop(*args, loop_args=loop_args, **kwargs)

    def _runs_fused_without_grad(self, *args, **kwargs) -> bool:
        return False

    def _multi_step_loop(self, *args, loop_args=None, **kwargs):
        T, loop_args, _, large, nchunks = self._chunk_plan(*args, loop_args=loop_args)
        if self.save_grad_history:
            self._grad_history = {}
            if self.grad_state_names:
                self._grad_history = {k: [None] * T for k in self.grad_state_names}
        zs, states = [], {}
        for c in range(nchunks):
            lo, hi = c * large, min(T, (c + 1) * large)
            static = list(args)
            for t in range(lo, hi):
                for i in loop_args:
                    static[i] = args[i][t]
                z, st = self.single_step_forward(*static, **kwargs)
                if self.cpu_offload:
                    z = z.cpu()
                    st = {k: v.cpu() for k, v in st.items()}
                zs.append(z)
                for k, v in st.items():
                    states.setdefault(k, []).append(v)
        if self.save_grad_history and torch.is_grad_enabled():
            self._hook_grad_history(states, T)
        return torch.stack(zs, 0), {k: torch.stack(v, 0) for k, v in states.items()}


def make_rnn(obj=None, allow_buffer=False, **rnn_kwargs):
    """Wrap a MemoryModule (class, instance or decorator use) in the loop."""

    def build(cell):
        class RNNWrapped(RecurrentNNAbstract):
            def __init__(self, *args, **kwargs):
                super().__init__(**rnn_kwargs)
                self.rnn_cell = cell(*args, **kwargs) if isinstance(cell, type) else cell

            def single_step_forward(self, *args, **kwargs):
                out = self.rnn_cell(*args, **kwargs)
                return out, filter_hidden_states(self.rnn_cell, self.update_state_names,
                                                 allow_buffer=allow_buffer)

        RNNWrapped.__name__ = f"RNN_{getattr(cell, '__name__', type(cell).__name__)}"
        return RNNWrapped

    if isinstance(obj, type):
        return build(obj)
    if isinstance(obj, base.MemoryModule):
        return build(obj)()
    if obj is None:
        return build
    raise TypeError("`make_rnn` expects a MemoryModule class, a MemoryModule instance, or None.")


class RecurrentNN(RecurrentNNAbstract):
    """z_t = neuron(synapse.psc + x_t); synapse(z_t) -- btorch/models/rnn.py:517-528."""

    def __init__(self, neuron: nn.Module, synapse: nn.Module, syn_inp_module=None,
                 neuron_inp_module=None, *, update_state_names=None, unroll=8,
                 chunk_size=None, cpu_offload=False, grad_checkpoint=False, allow_buffer=False,
                 **kwargs):
        super().__init__(update_state_names=update_state_names, unroll=unroll,
                         chunk_size=chunk_size, cpu_offload=cpu_offload,
                         grad_checkpoint=grad_checkpoint, **kwargs)
        self.neuron = neuron
        self.synapse = synapse
        self.neuron_inp_module = neuron_inp_module
        self.syn_inp_module = syn_inp_module
        self.allow_buffer = allow_buffer
        self._engine: Engine | None = None
        self._engine_key = None

    def single_step_forward(self, x, x_syn=None):
        if self.neuron_inp_module is not None:
            x = self.neuron_inp_module(x)
        if self.syn_inp_module is not None:
            x_syn = self.syn_inp_module(x_syn)
        z = self.neuron(self.synapse.psc + x)
        self.synapse(z if x_syn is None else z + x_syn)
        return z, filter_hidden_states(self, self.update_state_names,
                                       allow_buffer=self.allow_buffer)

    # -- fused path ----------------------------------------------------------
    def _engine_identity(self) -> tuple:
        """Modules and layout tensors a matched CellSpec froze: a change of
        any of them (a new connection, a reloaded layout, a new history
        module) matches the cell again.  Read on every call, so attributes
        are looked up without the exception a missing nn.Module attribute
        raises."""
        syn = _attr(self, "synapse")
        inner = _attr(syn, "psc_module") or _attr(syn, "base_psc")
        conn = _attr(syn, "linear") or _attr(inner, "linear")
        hist = _attr(syn, "history")
        nrn = _attr(self, "neuron")
        groups = tuple(id(g) for g in nrn.groups()) if hasattr(type(nrn), "groups") else ()
        return (id(nrn), groups, id(syn), id(inner), id(_attr(self, "syn_inp_module")), id(conn),
                id(_attr(conn, "_csr_row_ptr")), id(hist), id(_attr(self, "synapse_apical")))

    def fused_engine(self) -> Engine | None:
        key = self._engine_identity()
        if self._engine is None or self._engine_key != key:
            spec = match_cell(self)
            self._engine = Engine(spec) if spec is not None else None
            self._engine_key = key
        return self._engine

    def _record_names(self) -> list[str]:
        key = (self._engine_identity(), None if self.update_state_names is None
               else tuple(self.update_state_names), self.allow_buffer)
        hit = getattr(self, "_names_cache", None)
        if hit is None or hit[0] != key:
            names = hidden_state_names(self, self.update_state_names,
                                       allow_buffer=self.allow_buffer)
            self._names_cache = hit = (key, names)
        return list(hit[1])

    def _runs_fused_without_grad(self, *args, **kwargs) -> bool:
        """A call the no-grad fused path serves (one input sequence on a HIP
        device, nothing requiring a gradient, every requested state
        recordable)."""
        if torch.is_grad_enabled() or kwargs or len(args) != 1 or self.cudagraph:
            return False
        x = args[0]
        if not torch.is_tensor(x) or x.device.type != "cuda":
            return False
        eng = self.fused_engine()
        return eng is not None and self._engine_records(eng)

    def _graph(self, eng, x: torch.Tensor):
        """The graph runner when ``cudagraph`` is set (checks first), else None."""
        if not self.cudagraph:
            return None
        from .cudagraph import GraphRunner
        GraphRunner.check_supported(self, x)
        if eng is None:
            raise NotImplementedError("cudagraph=True replays the fused engine; this cell is "
                                      "not one it runs")
        if self._graph_runner is None:
            self._graph_runner = GraphRunner(warmup=self.cudagraph_warmup)
        return self._graph_runner

    def _multi_step_loop(self, *args, loop_args=None, **kwargs):
        eng = self.fused_engine()
        x = args[0] if args else None
        if (eng is None or len(args) != 1 or kwargs or not torch.is_tensor(x)
                or x.device.type != "cuda"):
            if self.cudagraph:
                raise NotImplementedError("cudagraph=True replays the fused engine over one "
                                          "input sequence on a HIP device")
            return super()._multi_step_loop(*args, loop_args=loop_args, **kwargs)
        from ._adjoint import fused_adjoint_covers, needs_grad, run_with_grad

        if not self._engine_records(eng):
            # a requested state the engine does not record (e.g. GLIF3's
            # after-spike currents): the per-step loop returns it
            if self.cudagraph:
                raise NotImplementedError("cudagraph=True: the fused engine does not record "
                                          "every requested state")
            return super()._multi_step_loop(*args, loop_args=loop_args, **kwargs)
        graph = self._graph(eng, x)
        if needs_grad(self, x):
            lanes = x[0].numel() // max(1, x.shape[-1]) if x.ndim > 1 else 1
            if fused_adjoint_covers(eng.spec, lanes) and not self.save_grad_history:
                return run_with_grad(self, eng, x)
            # autograd through the standalone HIP steps (which raise for
            # what they cannot differentiate); per-step states carry the
            # gradient hooks of save_grad_history
            return super()._multi_step_loop(*args, loop_args=loop_args, **kwargs)
        with torch.no_grad():
            return self._fused_forward(eng, x, graph=graph)

    def _input_sequence(self, x: torch.Tensor, n: int):
        T = x.shape[0]
        if self.neuron_inp_module is not None:
            mod = self.neuron_inp_module
            flat = x.reshape(-1, x.shape[-1])
            x = mod(flat) if isinstance(mod, BaseSparseConn) else mod(flat)
        lead = x.shape[: -1] if x.ndim > 1 else ()
        x = x.reshape(T, -1, n) if x.numel() else x
        return x.to(torch.float32).contiguous(), lead

    def _fused_forward(self, eng: Engine, x: torch.Tensor, x_apical: torch.Tensor | None = None,
                       graph=None):
        sp = eng.spec
        n = sp.neuron.size
        T = x.shape[0]
        xs, _ = self._input_sequence(x, n)
        lanes = xs.shape[1]
        xa = None
        if x_apical is not None:
            xa = torch.broadcast_to(x_apical, x.shape).reshape(T, lanes, n)
            xa = xa.to(torch.float32).contiguous()
        names = self._record_names()
        want = self._wanted_records(names, sp)
        hist0 = sp.history.history.clone() if (sp.history is not None and
                                               "history" in want) else None
        cursor0 = sp.history.cursor if sp.history is not None else 0
        if graph is not None:
            if xa is not None:
                raise NotImplementedError("cudagraph=True replays runs without an apical input")
            raster, recs = graph.run(eng, xs, want)
        else:
            raster, recs = eng.run(xs, T, lanes, want, x_apical=xa)
        return self._assemble(eng, names, recs, raster, hist0, cursor0, T, lanes, n)

    # neuron memories the engine records, by name within the neuron module
    _NEURON_RECORDS = {"v": "v", "refractory": "refractory", "v_pre_spike": "v_pre",
                       "i_a": "i_a", "i_bap": "i_bap", "theta_th": "theta_th"}

    @classmethod
    def _record_of(cls, k: str, sp) -> str | None:
        if isinstance(sp.neuron, MixedState):
            # "neuron.<group>.<memory>": the population record, sliced to
            # the group's range ("<record>@<group index>")
            for g, name in enumerate(sp.neuron.pop.layout.names):
                pre = f"neuron.{name}."
                if k.startswith(pre):
                    rec = cls._NEURON_RECORDS.get(k[len(pre):])
                    return None if rec is None else f"{rec}@{g}"
        if k.startswith(sp.neuron_key):
            return cls._NEURON_RECORDS.get(k[len(sp.neuron_key):])
        if k in sp.psc_keys:
            return "psc"
        if k in sp.syn_keys:
            return sp.syn_keys[k]
        if k.endswith("history.history"):
            return "history"
        return None

    @classmethod
    def _wanted_records(cls, names, sp) -> set[str]:
        want = {"z"}
        for k in names:
            rec = cls._record_of(k, sp)
            if rec is None:
                raise NotImplementedError(f"recording '{k}' is not supported by the fused engine")
            want.add(rec.partition("@")[0])
        return want

    def _engine_records(self, eng) -> bool:
        """Whether the fused engine records every requested state."""
        sp = eng.spec
        return all(self._record_of(k, sp) is not None for k in self._record_names())

    def _assemble(self, eng, names, recs, raster, hist0, cursor0, T, lanes, n):
        sp = eng.spec
        mixed = isinstance(sp.neuron, MixedState)
        if mixed:
            lead = tuple(sp.neuron.pop.groups()[0].v.shape[:-1])
            shape = lead + (n,)
        else:
            shape = tuple(sp.neuron.v.shape)
        z = recs["z"].reshape(T, *shape)
        states = {}
        for k in names:
            rec = self._record_of(k, sp)
            if rec == "history":
                seq = eng.history_sequence(raster, hist0, T, lanes, n, cursor0)
                states[k] = seq.reshape(T, *sp.history.history.shape)
            elif "@" in rec:
                base, _, g = rec.partition("@")
                sl = sp.neuron.pop.layout.span(int(g))
                states[k] = recs[base][..., sl].reshape(T, *shape[:-1], sl.stop - sl.start)
            else:
                states[k] = recs[rec].reshape(T, *shape)
        if getattr(sp.neuron, "pre_spike_v", False) and "v_pre" in recs:
            sp.neuron.v_pre_spike = recs["v_pre"][-1].reshape(shape).clone()
        # HeterSynapsePSC keeps its own receptor-summed copy of psc
        for m in sp.psc_modules:
            if m is not sp.exp_psc:
                m.psc = sp.exp_psc.psc.reshape(shape).clone()
        if self.cpu_offload:
            z = z.cpu()
            states = {k: v.cpu() for k, v in states.items()}
        return z, states


class ApicalRecurrentNN(RecurrentNN):
    """RecurrentNN with an apical / top-down input for multi-port neurons
    (btorch/models/rnn.py:530-649): z_t = neuron(synapse.psc + x_t, A_t)
    with A_t = x_apical_t (+ synapse_apical.psc); both synapses receive z_t.
    Call it as ``model(x, None, x_apical)`` so the loop slices the apical
    sequence.  A one-group two-compartment population without
    ``synapse_apical`` runs in the fused engine."""

    def __init__(self, neuron: nn.Module, synapse: nn.Module, synapse_apical=None,
                 syn_inp_module=None, neuron_inp_module=None, *, update_state_names=None,
                 unroll=8, chunk_size=None, cpu_offload=False, grad_checkpoint=False,
                 allow_buffer=False, **kwargs):
        super().__init__(neuron=neuron, synapse=synapse, syn_inp_module=syn_inp_module,
                         neuron_inp_module=neuron_inp_module,
                         update_state_names=update_state_names, unroll=unroll,
                         chunk_size=chunk_size, cpu_offload=cpu_offload,
                         grad_checkpoint=grad_checkpoint, allow_buffer=allow_buffer, **kwargs)
        self.synapse_apical = synapse_apical

    def single_step_forward(self, x, x_syn=None, x_apical=None):
        if self.neuron_inp_module is not None:
            x = self.neuron_inp_module(x)
        if self.syn_inp_module is not None:
            x_syn = self.syn_inp_module(x_syn)
        total_input = self.synapse.psc + x
        apical_input = x_apical
        if self.synapse_apical is not None:
            apical_input = (self.synapse_apical.psc if apical_input is None
                            else self.synapse_apical.psc + apical_input)
        z = (self.neuron(total_input) if apical_input is None
             else self.neuron(total_input, apical_input))
        self.synapse(z if x_syn is None else z + x_syn)
        if self.synapse_apical is not None:
            self.synapse_apical(z if x_syn is None else z + x_syn)
        return z, filter_hidden_states(self, self.update_state_names,
                                       allow_buffer=self.allow_buffer)

    def fused_engine(self) -> Engine | None:
        if self.synapse_apical is not None:
            return None
        return super().fused_engine()

    def _multi_step_loop(self, *args, loop_args=None, **kwargs):
        eng = self.fused_engine()
        x = args[0] if args else None
        x_syn = args[1] if len(args) > 1 else None
        x_apical = args[2] if len(args) > 2 else None
        fusable = (eng is not None and not kwargs and len(args) <= 3 and torch.is_tensor(x)
                   and x.device.type == "cuda" and x_syn is None
                   and (x_apical is None or (torch.is_tensor(x_apical)
                                             and x_apical.shape[0] == x.shape[0])))
        if not fusable or not self._engine_records(eng):
            if self.cudagraph:
                raise NotImplementedError("cudagraph=True replays the fused engine; this call "
                                          "is not one it runs")
            return RecurrentNNAbstract._multi_step_loop(self, *args, loop_args=loop_args,
                                                        **kwargs)
        graph = self._graph(eng, x)
        from ._adjoint import fused_adjoint_covers, needs_grad, run_with_grad

        if needs_grad(self, x) or (x_apical is not None and x_apical.requires_grad
                                   and torch.is_grad_enabled()):
            lanes = x[0].numel() // max(1, x.shape[-1]) if x.ndim > 1 else 1
            if (x_apical is None and fused_adjoint_covers(eng.spec, lanes)
                    and not self.save_grad_history):
                return run_with_grad(self, eng, x)
            # autograd through the standalone HIP steps
            return RecurrentNNAbstract._multi_step_loop(self, *args, loop_args=loop_args,
                                                        **kwargs)
        with torch.no_grad():
            return self._fused_forward(eng, x, x_apical, graph=graph)


class SomaApicalRecurrentNN(ApicalRecurrentNN):
    """ApicalRecurrentNN with both a somatic and an apical synapse
    (btorch/models/rnn.py:652-712)."""

    def __init__(self, neuron: nn.Module, synapse_soma: nn.Module, synapse_apical: nn.Module,
                 syn_inp_module=None, neuron_inp_module=None, *, update_state_names=None,
                 **kwargs):
        super().__init__(neuron=neuron, synapse=synapse_soma, synapse_apical=synapse_apical,
                         syn_inp_module=syn_inp_module, neuron_inp_module=neuron_inp_module,
                         update_state_names=update_state_names, **kwargs)
        self.synapse_soma = synapse_soma


__all__ = ["RecurrentNNAbstract", "RecurrentNN", "ApicalRecurrentNN", "SomaApicalRecurrentNN",
           "make_rnn"]
