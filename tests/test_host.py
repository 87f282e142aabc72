"""CPU tests of the host side: config registry, module state management,
connection layout, delay expansion, and the C ABI library's exports."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse
import torch

import btorch_amd
from btorch_amd import _lib
from btorch_amd.config import sparse as sparse_registry
from btorch_amd.connectome import expand_conn_for_delays
from btorch_amd.models import (
    LIF, ExponentialPSC, HeterSynapsePSC, RecurrentNN, SparseConn, SpikeHistory, environ,
    functional,
)
from btorch_amd._sparse_config import HipSparseConfig, SparseConfigRegistry
from oracle import rsnn

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "btorch_amd.h").read_text()
    declared = set(re.findall(r"^\s*(?:int|int64_t|size_t|const char\*)\s+(btx_\w+)\(", header, re.M))
    assert declared, "no declarations parsed"
    lib = _lib.load()
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(_lib.declared_symbols())
    assert lib.btx_version() >= 100


@pytest.mark.parametrize("cname,pyname", [("btx_rsnn_plan", "RsnnPlan"),
                                           ("btx_adj_lif_args", "AdjLifArgs")])
def test_ctypes_structs_match_the_c_layout(tmp_path, cname, pyname):
    """Every field offset and the size of the ctypes mirrors equal what a C
    compiler makes of include/btorch_amd.h."""
    import shutil
    import subprocess

    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    py = getattr(_lib, pyname)
    lines = [f'printf("size %zu\\n", sizeof({cname}));']
    for f, _ in py._fields_:
        lines.append(f'printf("{f} %zu\\n", offsetof({cname}, {f}));')
    src = tmp_path / "layout.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"btorch_amd.h\"\n"
                   "int main(void) {\n" + "\n".join(lines) + "\nreturn 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)],
                   check=True)
    out = dict(line.split() for line in subprocess.run(
        [str(exe)], check=True, capture_output=True, text=True).stdout.splitlines())
    assert int(out["size"]) == ctypes.sizeof(py)
    for f, _ in py._fields_:
        assert int(out[f]) == getattr(py, f).offset, f


def test_sparse_registry_and_errors():
    reg = SparseConfigRegistry()
    with pytest.raises(RuntimeError):
        reg.backend()
    assert isinstance(reg.backend("HIP"), HipSparseConfig)
    assert reg.resolve("triton").ordered is False
    assert reg.resolve("native").ordered is True
    assert reg.resolve("hip", {"slots": 4}).slots == 4
    with pytest.raises(ValueError):
        reg.resolve("hip", {"hash": True})
    with pytest.raises(ValueError):
        reg.resolve("hip", {"slots": 1})
    with pytest.raises(ValueError):
        reg.backend("cusparse")


def test_sparse_conn_layout_matches_coalesced_matrix():
    rng = np.random.default_rng(0)
    dense = (rng.random((70, 50)) < 0.2) * rng.normal(size=(70, 50))
    conn = SparseConn(scipy.sparse.coo_array(dense.astype(np.float32)))
    dst, src, w = rsnn.coalesce(dense.astype(np.float32))
    assert torch.equal(conn.indices, torch.stack([torch.from_numpy(dst), torch.from_numpy(src)]))
    assert np.array_equal(conn.magnitude.detach().numpy(), w)
    rp, col, perm = conn._csr_row_ptr.numpy(), conn._csr_col.numpy(), conn._csr_perm.numpy()
    rebuilt = np.zeros_like(dense)
    for s in range(70):
        cols = col[rp[s]:rp[s + 1]]
        assert np.all(np.diff(cols) > 0)
        rebuilt[s, cols] = w[perm[rp[s]:rp[s + 1]]]
    assert np.array_equal(rebuilt.astype(np.float32), dense.astype(np.float32))
    sd = conn.state_dict()
    assert set(sd) == {"indices", "magnitude"}
    conn2 = SparseConn(scipy.sparse.coo_array(dense.astype(np.float32)))
    conn2.load_state_dict(sd)
    assert torch.equal(conn2._csr_perm, conn._csr_perm)


def test_sparse_conn_cpu_forward_fails_loudly():
    conn = SparseConn(scipy.sparse.random(10, 10, density=0.3, format="coo", random_state=1))
    with pytest.raises(RuntimeError, match="HIP device"):
        conn(torch.ones(2, 10))


def test_prepare_finish_refcount():
    conn = SparseConn(scipy.sparse.random(8, 8, density=0.3, format="coo", random_state=1))
    with pytest.raises(RuntimeError):
        conn.finish_sparse()
    conn.prepare_sparse()
    conn.prepare_sparse()
    conn.finish_sparse()
    assert conn._prepared_sparse_weight is not None
    conn.finish_sparse()
    assert conn._prepared_sparse_weight is None


def test_expand_conn_matches_oracle():
    rng = np.random.default_rng(4)
    m = scipy.sparse.random(30, 30, density=0.2, format="coo", random_state=4)
    d = rng.integers(0, 9, size=m.nnz)
    a = expand_conn_for_delays(m, d, 6).toarray()
    b = rsnn.expand_conn_for_delays(m, d, 6).toarray()
    assert np.array_equal(a, b)
    with pytest.raises(ValueError):
        expand_conn_for_delays(m, -np.ones(m.nnz), 6)


def test_memory_module_init_reset_inplace():
    lif = LIF(16, tau_ref=2.0, v_reset=-0.5)
    lif.init_state(batch_size=3)
    assert lif.v.shape == (3, 16) and torch.all(lif.v == -0.5)
    lif.v.fill_(7.0)
    ptr = lif.v.data_ptr()
    lif.reset(inplace=True)
    assert lif.v.data_ptr() == ptr and torch.all(lif.v == -0.5)
    with pytest.raises(ValueError):
        lif.reset(batch_size=4, inplace=True)
    lif.reset(batch_size=4)
    assert lif.v.shape == (4, 16)


def test_param_compaction_round_trip():
    tau = torch.full((16,), 20.0)
    lif = LIF(16, tau=tau, trainable_param={"tau"})
    sd = lif.state_dict()
    assert sd["tau"].shape == ()
    lif2 = LIF(16, tau=torch.linspace(10, 30, 16), trainable_param={"tau"})
    lif2.load_state_dict(lif.state_dict())
    assert torch.all(lif2.tau == 20.0)
    with pytest.raises(ValueError):
        LIF(16, tau=torch.ones(5))


def test_hidden_state_keys_and_set():
    n, D = 20, 4
    m = scipy.sparse.random(n, n, density=0.2, format="coo", random_state=0)
    conn = SparseConn(expand_conn_for_delays(m, np.ones(m.nnz, int), D))
    syn = HeterSynapsePSC(n, 1, None, conn, base_psc=ExponentialPSC, max_delay_steps=D,
                          tau_syn=5.0)
    rnn = RecurrentNN(LIF(n, tau_ref=2.0), syn)
    functional.init_net_state(rnn, batch_size=2)
    keys = set(functional.named_hidden_states(rnn))
    assert keys == {"neuron.v", "neuron.refractory", "synapse.psc", "synapse.history.history",
                    "synapse.base_psc.psc"}
    functional.set_hidden_states(rnn, {"neuron.v": torch.ones(2, n)})
    assert torch.all(rnn.neuron.v == 1)
    functional.set_hidden_states(rnn, {"neuron": {"v": torch.zeros(2, n)}})
    assert torch.all(rnn.neuron.v == 0)


def test_environ_context_stack():
    with pytest.raises(KeyError):
        environ.get("zz_missing")
    with environ.context(dt=0.5):
        with environ.context(dt=0.1):
            assert environ.get("dt") == 0.1
        assert environ.get("dt") == 0.5


def test_spike_history_modes_match_oracle():
    for circular in (True, False):
        h = SpikeHistory(5, max_delay_steps=3, use_circular_buffer=circular)
        h.init_state(batch_size=2)
        o = rsnn.Ring(2, 5, 3, circular)
        rng = np.random.default_rng(0)
        for _ in range(7):
            z = torch.from_numpy((rng.random((2, 5)) < 0.5).astype(np.float32))
            h.update(z)
            o.update(z)
            assert torch.equal(h.get_flattened(3), o.get_flattened(3))
            assert torch.equal(h.get_delay(2), o.get_delay(2))


def test_engine_matches_supported_cells():
    from btorch_amd.models._engine import match_cell
    n = 12
    m = scipy.sparse.random(n, n, density=0.3, format="coo", random_state=2)
    rnn = RecurrentNN(LIF(n), ExponentialPSC(n, 5.0, SparseConn(m)))
    spec = match_cell(rnn)
    assert spec is not None and spec.bins == [(0, 0)]
    delays = np.array([1, 3] * (m.nnz // 2) + [1] * (m.nnz % 2))
    conn = SparseConn(expand_conn_for_delays(m, delays, 5))
    syn = HeterSynapsePSC(n, 1, None, conn, base_psc=ExponentialPSC, max_delay_steps=5,
                          tau_syn=5.0)
    spec = match_cell(RecurrentNN(LIF(n), syn))
    assert spec.bins == [(1, 1), (3, 3)] and spec.csr_d == 5 and spec.ring_d == 5
    # an unsupported cell is not matched
    assert match_cell(RecurrentNN(LIF(n), ExponentialPSC(n, 5.0, torch.nn.Identity()))) is None


def test_engine_matches_two_compartment_cells():
    from btorch_amd.models import (ApicalRecurrentNN, MixedNeuronPopulation,
                                   TwoCompartmentGLIF)
    from btorch_amd.models._engine import match_cell
    n = 12
    m = scipy.sparse.random(n, n, density=0.3, format="coo", random_state=3)

    def syn():
        return ExponentialPSC(n, 5.0, SparseConn(m))

    one = MixedNeuronPopulation([(n, TwoCompartmentGLIF(n))])
    spec = match_cell(ApicalRecurrentNN(one, syn()))
    assert spec is not None and isinstance(spec.neuron, TwoCompartmentGLIF)
    assert spec.neuron_key == "neuron.group_0."
    assert match_cell(RecurrentNN(MixedNeuronPopulation([(n, LIF(n))]), syn())) is not None
    # a bare two-compartment neuron returns (spike, v), which no synapse takes
    assert match_cell(RecurrentNN(TwoCompartmentGLIF(n), syn())) is None
    # two groups run in the fused step through the population's state
    # arrays; an apical synapse runs the per-step loop
    from btorch_amd.models._engine import MixedState
    two = MixedNeuronPopulation([(6, TwoCompartmentGLIF(6)), (6, LIF(6))])
    spec2 = match_cell(ApicalRecurrentNN(two, syn()))
    assert spec2 is not None and isinstance(spec2.neuron, MixedState)
    assert spec2.neuron.has_tc and spec2.neuron.size == 12
    assert match_cell(ApicalRecurrentNN(one, syn(), synapse_apical=syn())) is None
    with pytest.raises(ValueError):
        MixedNeuronPopulation([])
    with pytest.raises(ValueError):
        MixedNeuronPopulation([(0, LIF(1))])


def _kinetics_modules(n, linear):
    from btorch_amd.models import AlphaPSC, AlphaPSCBilleh, DualExponentialPSC
    with environ.context(dt=1.0):
        return [AlphaPSC(n, 4.0, linear, g_max=1.5), AlphaPSCBilleh(n, 6.0, linear),
                DualExponentialPSC(n, 9.0, 2.0, linear)]


def test_psc_kernels_equal_oracle_impulse_responses():
    """get_kernel (the sequence form) of each kinetics equals the oracle's
    step loop driven by one impulse."""
    mods = _kinetics_modules(4, torch.nn.Identity())
    syns = [rsnn.AlphaSyn(torch.tensor(4.0), torch.tensor(1.5)), rsnn.BillehSyn(torch.tensor(6.0)),
            rsnn.DualExpSyn(torch.tensor(9.0), torch.tensor(2.0))]
    for m, s in zip(mods, syns):
        k = m.get_kernel(1.0, 64).double()
        r = rsnn.syn_impulse_response(s, 64, 1.0)
        np.testing.assert_allclose(k.numpy(), r.numpy(), rtol=1e-5, atol=1e-7)


def test_kinetics_modules_state_and_engine_match():
    from btorch_amd.models import AlphaPSC, DelayedPSC, TwoCompartmentGLIF
    from btorch_amd.models._engine import match_cell
    from btorch_amd.models.neurons.mixed import MixedNeuronPopulation
    n = 12
    m = scipy.sparse.random(n, n, density=0.3, format="coo", random_state=5)
    for syn in _kinetics_modules(n, SparseConn(m)):
        syn.init_state(batch_size=2)
        assert syn.psc.shape == (2, n)
        for k in syn._STATES:
            assert getattr(syn, k).shape == (2, n)
        spec = match_cell(RecurrentNN(LIF(n), syn))
        assert spec is not None and spec.exp_psc is syn
        assert spec.syn_keys == {f"synapse.{k}": f"syn_h{j + 1}"
                                 for j, k in enumerate(syn._STATES)}
        with pytest.raises(RuntimeError, match="HIP device"):
            syn.single_step_forward(torch.ones(2, n))
    # delayed and receptor-typed forms; the default base_psc is AlphaPSC
    delays = np.array([1, 2] * (m.nnz // 2) + [1] * (m.nnz % 2))
    conn = SparseConn(expand_conn_for_delays(m, delays, 3))
    het = HeterSynapsePSC(n, 1, None, conn, max_delay_steps=3, tau_syn=5.0)
    assert isinstance(het.base_psc, AlphaPSC)
    spec = match_cell(RecurrentNN(LIF(n), het))
    assert spec.syn_keys == {"synapse.base_psc.h": "syn_h1"}
    spec = match_cell(RecurrentNN(LIF(n), DelayedPSC(AlphaPSC(n, 5.0, SparseConn(m)), 2)))
    assert spec.syn_keys == {"synapse.psc_module.h": "syn_h1"}
    # the two-compartment neuron runs the exponential synapse only
    one = MixedNeuronPopulation([(n, TwoCompartmentGLIF(n))])
    from btorch_amd.models import ApicalRecurrentNN
    assert match_cell(ApicalRecurrentNN(one, AlphaPSC(n, 5.0, SparseConn(m)))) is None
    # Billeh's model is defined for dt == 1
    from btorch_amd.models import AlphaPSCBilleh
    with environ.context(dt=0.5), pytest.raises(ValueError):
        AlphaPSCBilleh(n, 5.0, SparseConn(m))


def _constrained(seed=0, n_src=30, n_dst=20, groups=3):
    from btorch_amd.models import SparseConstrainedConn
    rng = np.random.default_rng(seed)
    dense = (rng.random((n_src, n_dst)) < 0.3) * rng.normal(size=(n_src, n_dst))
    m = scipy.sparse.coo_array(dense.astype(np.float32))
    g = m.copy()
    g.data = rng.integers(1, groups + 1, size=m.nnz).astype(np.float64)
    return SparseConstrainedConn(m, g), dense.astype(np.float32), g.toarray().astype(int)


def test_constrained_conn_groups_and_effective_weight():
    conn, dense, gdense = _constrained()
    dst, src = conn.indices.numpy()
    # every connection maps to the group the constraint matrix gives it
    assert np.array_equal(conn._constraint_scatter_indices.numpy(), gdense[src, dst] - 1)
    assert torch.equal(conn.initial_weight, torch.from_numpy(dense[src, dst]))
    with torch.no_grad():
        conn.magnitude.copy_(torch.tensor([0.5, -2.0, 3.0]))
    w = conn._get_effective_weight()
    assert torch.equal(w, conn.initial_weight * conn.magnitude.detach()[gdense[src, dst] - 1])
    conn.constrain()
    assert torch.equal(conn.magnitude.detach(), torch.tensor([0.5, 0.0, 3.0]))
    info = conn.get_group_info(include_weights=True)
    assert info["num_connections"].sum() == conn.nnz
    conn.set_group_magnitude(group_id=1, value=4.0)
    assert float(conn.magnitude[1]) == 4.0
    groups = conn.get_weights_by_group()
    assert sum(v.numel() for v in groups.values()) == conn.nnz
    with pytest.raises(ValueError):
        conn.set_group_magnitude()
    # the magnitude gradient sums the per-connection gradient by group
    conn.reset_parameters()
    gw = torch.randn(conn.nnz, generator=torch.Generator().manual_seed(0))
    (conn._get_effective_weight() * gw).sum().backward()
    want = torch.zeros(3).index_add_(0, conn._constraint_scatter_indices,
                                     gw * conn.initial_weight)
    assert torch.allclose(conn.magnitude.grad, want, atol=1e-6)
    # a connection without a group is an error, as in the reference
    from btorch_amd.models import SparseConstrainedConn
    m = scipy.sparse.coo_array(np.eye(4, dtype=np.float32))
    g = scipy.sparse.coo_array(np.diag([1.0, 1.0, 0.0, 2.0]))
    with pytest.raises(AssertionError):
        SparseConstrainedConn(m, g)
    # state dict: magnitude only (initial weights on request)
    assert set(conn.state_dict()) == {"indices", "magnitude"}


def test_sheet_generator_statistics():
    """The 2-D sheet workload (largest single-GPU configuration), generated
    small on the CPU: degrees, displacement, E/I split and delays."""
    from btorch_amd import workloads
    side = 40
    pre, post, w, d = workloads.sheet_net_device(side=side, mean_degree=50.0, sigma=0.1,
                                                 device="cpu", chunk=4096)
    n = side * side
    assert pre.numel() == post.numel() == w.numel() == d.numel()
    assert abs(pre.numel() / n - 50.0) < 1.0
    assert int(post.min()) >= 0 and int(post.max()) < n
    assert int(d.min()) >= 1 and int(d.max()) <= 40
    # Gaussian displacement of 0.1 side = 4 cells per axis, wrapped
    for a, b in ((pre % side, post % side), (pre // side, post // side)):
        delta = torch.remainder(b - a + side // 2, side) - side // 2
        assert abs(float(delta.float().std()) - 4.0) < 0.2
    inh = (pre % 5) == 0
    assert bool((w[inh] < 0).all()) and bool((w[~inh] > 0).all())
    assert abs(float(inh.float().mean()) - 0.2) < 0.01


def test_fused_adjoint_coverage_decision():
    """Which cells the fused adjoint differentiates; the others train through
    the per-step loop of standalone steps."""
    from btorch_amd.models import GLIF3, AlphaPSC
    from btorch_amd.models._adjoint import fused_adjoint_covers
    from btorch_amd.models._engine import match_cell
    n = 12
    m = scipy.sparse.random(n, n, density=0.3, format="coo", random_state=6)
    lif_exp = match_cell(RecurrentNN(LIF(n), ExponentialPSC(n, 5.0, SparseConn(m))))
    assert fused_adjoint_covers(lif_exp, 8) and fused_adjoint_covers(lif_exp, 64)
    assert not fused_adjoint_covers(lif_exp, 65)
    with environ.context(dt=0.1):
        alpha = match_cell(RecurrentNN(LIF(n), AlphaPSC(n, 5.0, SparseConn(m))))
    assert not fused_adjoint_covers(alpha, 8)
    glif = match_cell(RecurrentNN(GLIF3(n), ExponentialPSC(n, 5.0, SparseConn(m))))
    assert not fused_adjoint_covers(glif, 8)
    trainable_vth = match_cell(RecurrentNN(LIF(n, trainable_param={"v_threshold"}),
                                           ExponentialPSC(n, 5.0, SparseConn(m))))
    assert not fused_adjoint_covers(trainable_vth, 8)


# ---------------------------------------------------------------------------
# state shapes, engine identity, deferred history write-back (host logic)
# ---------------------------------------------------------------------------

def test_lane_state_broadcasts_or_raises_before_any_launch():
    from btorch_amd.models import LIF, _plan
    n = 10
    lif = LIF(n, tau_ref=2.0)
    lif.init_state()
    assert lif.refractory.shape == (n,)
    ref = _plan.lane_state(lif, "refractory", (4, n))
    assert ref.shape == (4, n) and ref.is_contiguous() and lif.refractory is ref
    lif.refractory = torch.zeros(3, n)
    with pytest.raises(RuntimeError, match="does not match"):
        _plan.lane_state(lif, "refractory", (4, n))
    lif.refractory = None
    with pytest.raises(RuntimeError, match="not initialised"):
        _plan.lane_state(lif, "refractory", (4, n))


def test_engine_identity_follows_connection_and_layout():
    import scipy.sparse
    from btorch_amd.models import ExponentialPSC, LIF, RecurrentNN, SparseConn
    n = 12
    m = scipy.sparse.random(n, n, density=0.3, random_state=0, dtype=np.float32)
    model = RecurrentNN(LIF(n), ExponentialPSC(n, 5.0, SparseConn(m)), step_mode="m")
    k0 = model._engine_identity()
    model.synapse.linear = SparseConn(m)
    k1 = model._engine_identity()
    assert k1 != k0
    model.synapse.linear.load_state_dict(model.synapse.linear.state_dict())
    assert model._engine_identity() != k1  # the layout was rebuilt on load


def test_spike_history_deferred_writeback_semantics():
    from btorch_amd.models import SpikeHistory
    h = SpikeHistory(5, max_delay_steps=3)
    h.init_state(batch_size=2, persistent=True)
    calls = []

    def wb(mod):
        calls.append(1)
        mod._buffers["history"].fill_(1.0)

    # a read writes back once, at the read
    h.defer_writeback(wb, cursor=2)
    assert h.cursor == 2 and not calls
    assert float(h.history.sum()) == 2 * 3 * 5 and len(calls) == 1
    h.history  # noqa: B018
    assert len(calls) == 1
    # state_dict and memories() write back; assigning the memory discards it
    h.defer_writeback(wb, cursor=0)
    h._buffers["history"].zero_()
    assert float(h.state_dict()["history"].sum()) == 30 and len(calls) == 2
    h.defer_writeback(wb, cursor=0)
    h.history = torch.zeros(2, 3, 5)
    assert float(h.history.sum()) == 0 and len(calls) == 2
    h.defer_writeback(wb, cursor=1)
    h.set_cursor(2)
    assert len(calls) == 3 and h.cursor == 2
    h.defer_writeback(wb, cursor=1)
    h.reset()
    assert float(h.history.sum()) == 0 and len(calls) == 3


def test_ring_history_token_names_only_the_live_memory():
    # the engine reuses its device ring only while the module's history is
    # the very memory the ring was written back to: a replaced buffer (even
    # one that reuses a freed object's id), an in-place change or another
    # cursor all count as a different history
    import gc

    from btorch_amd.models import SpikeHistory
    from btorch_amd.models._engine import Engine
    h = SpikeHistory(5, max_delay_steps=3, use_circular_buffer=False)
    h.init_state(batch_size=2, persistent=True)
    tok = Engine._hist_token(h)
    assert Engine._holds(tok, h) and not Engine._holds(None, h)
    h.history.add_(0.0)
    assert not Engine._holds(tok, h)
    tok = Engine._hist_token(h)
    ids = set()
    for _ in range(20):  # fresh buffers, some of them at a freed buffer's id
        h.history = torch.zeros(2, 3, 5)
        ids.add(id(h._buffers["history"]))
        assert not Engine._holds(tok, h)
        gc.collect()
    c = SpikeHistory(5, max_delay_steps=3)
    c.init_state(batch_size=2, persistent=True)
    tok = Engine._hist_token(c)
    c.set_cursor(1)
    assert not Engine._holds(tok, c)


def test_mixed_plan_lays_parameters_out_per_neuron():
    import ctypes
    from btorch_amd import _lib
    from btorch_amd.models import GLIF3, MixedNeuronPopulation, TwoCompartmentGLIF
    from btorch_amd.models._engine import MixedState
    from btorch_amd.models._plan import mixed_plan
    pop = MixedNeuronPopulation({"a": (5, LIF(5, tau=12.0, tau_ref=2.0, hard_reset=True)),
                                 "b": (3, TwoCompartmentGLIF(3)),
                                 "c": (4, GLIF3(4, tau=torch.tensor([5.0, 6.0, 7.0, 8.0]),
                                                k=(0.1, 0.2), asc_amps=(1.0, -1.0)))})
    st = MixedState(pop)
    assert st.has_tc and st.n_asc == 2 and st._use_refractory
    keep: list = []
    with environ.context(dt=0.5):
        p = mixed_plan(st, 2, "cpu", keep)

    def read(ptr, count, ctype=ctypes.c_float):
        return list((ctype * count).from_address(ptr))

    assert p.neuron_model == _lib.NEURON_MIXED
    assert p.mixed_models == 0b111 and p.use_refractory == 1 and p.n_asc == 2
    kinds = read(p.neuron_kinds, 12, ctypes.c_uint8)
    assert kinds == [_lib.NEURON_LIF | _lib.MIXED_HARD_RESET] * 5 + \
        [_lib.NEURON_TWO_COMPARTMENT] * 3 + [_lib.NEURON_GLIF3] * 4
    assert read(p.tau, 12) == [12.0] * 5 + [1.0] * 3 + [5.0, 6.0, 7.0, 8.0]
    assert read(p.tau_ref, 12)[:5] == [2.0] * 5 and read(p.tau_ref, 12)[5:] == [0.0] * 7
    assert p.tau_stride == 1 and p.tc_stride[0] == 1
    asc_k = read(p.asc_k, 24)
    assert asc_k[-8:] == pytest.approx([0.1, 0.2] * 4) and asc_k[:16] == [0.0] * 16
