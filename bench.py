"""Benchmark: synaptic events/s of the 140k-neuron connectome time step on one GPU.

Workload (seeded, synthetic): N=140,000 LIF neurons (tau_m ~ U(10,30) ms per
neuron, v_th 1, v_reset 0, refractory 2 ms, soft reset), 70/30 E/I,
out-degree power law (exponent 2.1, k in [100, 5000]; E ~ 50M), lognormal
weights (I x4 negative), integer delays U{1..20} through
HeterSynapsePSC(ExponentialPSC tau_s=5 ms, max_delay_steps=21), dt=0.1 ms,
8 lanes, Poisson background 5 Hz per neuron-lane as a one-step current kick
of 12 (v jumps by dt*12 = 1.2, one spike per event from rest), drawn in the
step kernel from a counter-based hash of (lane seed, step, neuron).  Weights
lognormal(mu=-3.7, sigma=0.5) keep the net background-dominated (~0.04%
active per step, the configuration's ~0.05%); --w-mu/--kick change them.
A "step" is one simulation time step of all 8 lanes.

Prints one JSON line (rank 0).  value = synaptic events/s over the timed
steps (events = (lane, edge) deliveries, counted exactly from the raster).
``roofline`` prices the dominant kernel's algorithmic bytes per launch
against its HIP-event-timed average duration; ``cpu_baseline`` times the
CPU oracle (torch-CPU restatement of the reference's native path) on a
bounded sample of the same workload from the same state, and also checks
that sample's raster against the GPU's bit for bit.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--config", choices=("connectome", "sheet", "harness"), default="connectome",
                    help="connectome: the 140k-neuron headline (default); sheet: the largest "
                         "single-GPU configuration, 1M two-compartment neurons on a 2-D sheet "
                         "with ~1B synapses; harness: the reference benchmark's shape "
                         "(Hemibrain-sized net, T=128, batch 1, spike-rate sweep)")
    ap.add_argument("--sheet-steps", type=int, default=10_000,
                    help="timed steps of --config sheet")
    ap.add_argument("--n", type=int, default=140_000)
    ap.add_argument("--lanes", type=int, default=8)
    ap.add_argument("--kmin", type=int, default=100)
    ap.add_argument("--kmax", type=int, default=5000)
    ap.add_argument("--delay-max", type=int, default=20)
    ap.add_argument("--kick", type=float, default=12.0,
                    help="current of one background Poisson event (v jumps by dt*kick)")
    ap.add_argument("--w-mu", type=float, default=-3.7,
                    help="log-mean of the lognormal weights (sigma 0.5, inhibitory x4)")
    ap.add_argument("--mode", choices=["ordered", "fast"], default="ordered")
    ap.add_argument("--kernel-steps", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--serve-requests", type=int, default=48,
                    help="requests of the serving measurement (0 = skip it)")
    ap.add_argument("--serve-lanes", type=int, default=256)
    ap.add_argument("--serve-rate", type=float, default=200.0, help="arrivals per second")
    ap.add_argument("--serve-tmax", type=int, default=5000)
    ap.add_argument("--train-steps", type=int, default=2,
                    help="timed training steps of the differentiable 16k-neuron net (0 = skip)")
    ap.add_argument("--no-sweep", dest="sweep", action="store_false",
                    help="skip the activity sweep of the step kernel")
    ap.add_argument("--no-c0", dest="c0", action="store_false",
                    help="skip the 4096-neuron CPU-runnable configuration (GPU and CPU oracle)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 512 MiB device copy first (HBM counter calibration)")
    return ap.parse_args()


def build(args, dev):
    from btorch_amd import workloads
    from btorch_amd.models import (
        LIF, ExponentialPSC, HeterSynapsePSC, RecurrentNN, SparseConn, functional,
    )

    n, D = args.n, args.delay_max + 1
    c = workloads.powerlaw_net_fast(n=n, k_min=args.kmin, k_max=args.kmax, exc_frac=0.7,
                                    delay_lo=1, delay_hi=args.delay_max, seed=args.seed,
                                    w_mu=args.w_mu)
    # the delay-expanded, coalesced connectome and its CSR built on the device
    edges = [torch.from_numpy(a).to(dev) for a in (c.pre, c.post, c.w, c.delay)]
    backend = "native" if args.mode == "ordered" else "triton"
    SparseConn.from_edges(*(e[:1024] for e in edges[:3]), edges[3][:1024], D, shape=(n, n))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    conn = SparseConn.from_edges(*edges[:3], edges[3], D, shape=(n, n), sparse_backend=backend)
    torch.cuda.synchronize()
    build_ms = 1e3 * (time.perf_counter() - t0)
    rng = np.random.default_rng(args.seed + 7)
    tau = torch.from_numpy(rng.uniform(10.0, 30.0, n).astype(np.float32))
    syn = HeterSynapsePSC(n, 1, None, conn, base_psc=ExponentialPSC, max_delay_steps=D,
                          use_circular_buffer=True, tau_syn=5.0)
    model = RecurrentNN(LIF(n, tau=tau, tau_ref=2.0), syn, step_mode="m",
                        update_state_names=[]).to(dev)
    functional.init_net_state(model, batch_size=args.lanes, device=dev, dtype=torch.float32)
    # per-(neuron, delay bin) out-degree for exact event counting
    rp = conn._csr_row_ptr
    deg = (rp[1:] - rp[:-1]).reshape(n, D).to(torch.float64)
    return model, conn, deg, c.nnz, build_ms


def delivery_table(raster, n, deg):
    """A[t, d] = deliveries at delay d caused by the spikes of step t (summed over lanes)."""
    from btorch_amd.models._engine import unpack_bits
    T = raster.shape[0]
    out = torch.zeros((T, deg.shape[1]), dtype=torch.float64, device=deg.device)
    act = torch.zeros((T, deg.shape[1]), dtype=torch.float64, device=deg.device)
    nz = (deg > 0).to(torch.float64)
    # steps per chunk: the unpacked chunk stays below 2^30 elements
    step = max(1, min(64, (1 << 30) // max(1, raster.shape[1] * n)))
    for t0 in range(0, T, step):
        s = unpack_bits(raster[t0:t0 + step], n)                    # (t, lanes, n)
        t_idx, _, j = s.nonzero(as_tuple=True)
        # integer counts in float64: exact whatever the summation order
        out[t0:t0 + step].index_add_(0, t_idx, deg[j])
        act[t0:t0 + step].index_add_(0, t_idx, nz[j])
    return out, act


def unpack_count(raster):
    """Total spikes in a raster (popcount of its words)."""
    tot = 0
    for t0 in range(0, raster.shape[0], 16):
        w = raster[t0:t0 + 16]
        shifts = torch.arange(64, device=w.device, dtype=torch.int64)
        tot += int(((w.unsqueeze(-1) >> shifts) & 1).sum())
    return tot


def window_events(tables, lo, hi):
    """Deliveries (and active virtual rows) landing in steps [lo, hi) of the
    concatenated history of tables (each row t carries its delays d)."""
    out, act = tables
    T, D = out.shape
    ev = 0.0
    rows = 0.0
    for d in range(D):
        t0, t1 = max(0, lo - d), max(0, hi - d)
        ev += float(out[t0:t1, d].sum())
        rows += float(act[t0:t1, d].sum())
    return ev, rows


def _traffic(args):
    """HBM bytes per step launch measured by the rocprofv3 FETCH_SIZE and
    WRITE_SIZE counters (scaled by the same counters on a plain copy of known
    size), from the committed summary of this configuration; None if the
    configuration differs or no summary is present."""
    path = os.path.join(ROOT, "profiles", "k_step_140k_8lanes_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    cfg = d.get("config", {})
    if (cfg.get("n"), cfg.get("lanes"), cfg.get("kmin"), cfg.get("w_mu"), cfg.get("mode"),
            cfg.get("input")) != (args.n, args.lanes, args.kmin, args.w_mu, args.mode,
                                  "x tensor, record-free step (model(x))"):
        return None
    return {"bytes_per_launch": d.get("bytes_per_launch"),
            "source": "profiles/k_step_140k_8lanes_traffic.txt: rocprofv3 FETCH_SIZE and "
                      "WRITE_SIZE of a separate run, scaled by the same counters on a copy "
                      "of known size (L2-to-fabric requests, an upper bound on DRAM traffic)"}


def sheet_bench(args, dev):
    """The largest single-GPU configuration: N = 1000 x 1000 two-compartment
    GLIF neurons on a periodic unit square (tau_s ~ U(10,30) ms, threshold
    adaptation tau_th ~ U(100,500) ms, delta_th 1 mV), ~1e9 synapses from a
    Gaussian kernel (sigma 0.1 side, mean out-degree 1000), delays U{1..40}
    (D = 41), 80/20 E/I lognormal(0.0, 0.5) weights (I x4 negative),
    exponential synapse tau_s 5 ms, batch 1, dt 0.1 ms, 3 Hz Poisson drive as
    a one-step current kick of 8000 (one kick fires a neuron from rest).
    Edges are generated and the connectome built on the device.  Times
    --sheet-steps steps (default the configuration's 10,000) after 200
    warmup steps."""
    from btorch_amd import workloads
    from btorch_amd.models import (
        ExponentialPSC, HeterSynapsePSC, MixedNeuronPopulation, RecurrentNN, SparseConn,
        TwoCompartmentGLIF, environ, functional,
    )

    side, D, lanes = 1000, 41, 1
    n = side * side
    t0 = time.perf_counter()
    pre, post, w, delay = workloads.sheet_net_device(side=side, seed=args.seed + 3, device=dev)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    _progress(f"sheet: {pre.numel()} edges generated in {gen_s:.2f} s")
    t0 = time.perf_counter()
    conn = SparseConn.from_edges(pre, post, w, delay, D, shape=(n, n))
    torch.cuda.synchronize()
    build_ms = 1e3 * (time.perf_counter() - t0)
    edges = conn.nnz
    del pre, post, w, delay
    torch.cuda.empty_cache()
    _progress(f"sheet: connectome built in {build_ms:.1f} ms ({edges} unique edges)")
    rng = np.random.default_rng(args.seed + 11)
    tau_s = torch.from_numpy(rng.uniform(10.0, 30.0, n).astype(np.float32))
    tau_th = torch.from_numpy(rng.uniform(100.0, 500.0, n).astype(np.float32))
    cell = TwoCompartmentGLIF(n, tau_s=tau_s, tau_th=tau_th, delta_th=1.0)
    syn = HeterSynapsePSC(n, 1, None, conn, base_psc=ExponentialPSC, max_delay_steps=D,
                          use_circular_buffer=True, tau_syn=5.0)
    model = RecurrentNN(MixedNeuronPopulation([(n, cell)]), syn, step_mode="m",
                        update_state_names=[]).to(dev)
    functional.init_net_state(model, batch_size=lanes, device=dev, dtype=torch.float32)
    eng = model.fused_engine()
    if eng is None:
        raise RuntimeError("the sheet cell is not matched by the fused engine")
    rp = conn._csr_row_ptr
    deg = (rp[1:] - rp[:-1]).reshape(n, D).to(torch.float64)
    p_in = np.float32(3.0 * 0.1 * 1e-3)
    drive = {"seed": torch.tensor([1009 * args.seed + 17], dtype=torch.int64, device=dev),
             "prob": torch.full((lanes,), float(p_in), dtype=torch.float32, device=dev),
             "t0": torch.zeros(lanes, dtype=torch.int64, device=dev), "kick": 8000.0}
    W, K = 200, args.sheet_steps

    def run(T):
        raster, _ = eng.run(None, T, lanes, set(), drive=drive)
        drive["t0"] += T
        return raster

    from btorch_amd.models.functional import prepare_sparse_modules
    with environ.context(dt=0.1), torch.no_grad(), prepare_sparse_modules(model, lanes):
        r_w = run(W)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record()
        r_t = run(K)
        ev1.record()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        region_ms = ev0.elapsed_time(ev1)
    _progress(f"sheet: {K} steps in {elapsed:.3f} s")
    tables = delivery_table(torch.cat([r_w, r_t], 0), n, deg)
    ev_t, rows_t = window_events(tables, W, W + K)
    spikes = float(unpack_count(r_t))
    nwords = (n + 63) // 64
    # F: v, psc, i_a, theta_th read and written, i_bap read (36 B per
    # neuron-lane), per-neuron coefficients of tau_s and tau_th (12 B per
    # neuron), spike words to the raster and ring, ring indices; the push
    # and the reduction as for the headline (20 B per active row, 28 B per term)
    bytes_launch = (36.0 * n * lanes + 12.0 * n + 16.0 * nwords * lanes + 4.0 * spikes / K
                    + 20.0 * rows_t / K + 28.0 * ev_t / K)
    avg_ms = region_ms / K
    achieved = bytes_launch / (avg_ms * 1e-3)
    return {
        "metric": "synaptic events/s, 1M-neuron two-compartment sheet",
        "value": ev_t / elapsed, "unit": "events/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": 1e3 * elapsed / K, "steps_per_s": K / elapsed,
        "events_per_step": ev_t / K, "spikes_per_step": spikes / K,
        "rate_hz": spikes / K / n / 1e-4,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded 2-D Gaussian-kernel connectome generated on the device)",
        "config": {"workload": "largest single-GPU sim: 1M two-compartment GLIF on a 2-D "
                               "sheet, ~1B synapses (sigma 0.1, mean out-degree 1000), "
                               "delays 1-40 (D=41), 80/20 E/I, 3 Hz Poisson drive, batch 1, "
                               "dt 0.1 ms",
                   "n": n, "synapses": edges, "lanes": lanes, "delay_bins": D,
                   "edge_generation_s": round(gen_s, 2), "conn_build_ms": round(build_ms, 1)},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None,
                     "kernel": "k_step (fused step + push + reduction)",
                     "bytes_per_launch": bytes_launch, "avg_us": avg_ms * 1e3,
                     "avg_us_source": "HIP events around the timed region / steps"},
    }


def _progress(msg: str) -> None:
    """A progress line on stderr (stdout carries only the result line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported here "
              "(the path shards by postsynaptic neuron; see the design notes).", file=sys.stderr)
        sys.exit(2)
    from btorch_amd.models import environ

    dev = torch.device("cuda:0")
    torch.manual_seed(args.seed)
    if args.config == "sheet":
        print(json.dumps(sheet_bench(args, dev)), flush=True)
        return
    if args.config == "harness":
        print(json.dumps(harness_bench(args, dev)), flush=True)
        return
    if args.copy_probe:
        # a plain streaming kernel of known traffic (512 MiB read, 512 MiB
        # written) to calibrate the memory counters
        a = torch.ones(128 * 2**20, device=dev)
        b = torch.empty_like(a)
        torch.neg(a, out=b)
        torch.cuda.synchronize()
        del a, b
    t_build = time.perf_counter()
    model, conn, deg, nnz, conn_build_ms = build(args, dev)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t_build
    eng = model.fused_engine()
    n, lanes, D = args.n, args.lanes, args.delay_max + 1
    W, K, KS = args.warmup, args.steps, args.kernel_steps
    # Poisson background: lane b's neuron j gets the kick at step s when the
    # counter-based hash of (seed_b, s, j) falls below rate*dt; generated on
    # the device as the input tensor x (T, lanes, n) before the clock (the
    # CPU oracle draws the same numbers)
    seeds = (np.arange(lanes, dtype=np.int64) * 7919 + 1009 * args.seed) & 0x7FFFFFFFFFFFFFFF
    p_in = np.float32(5.0 * 0.1 * 1e-3)
    from btorch_amd._kernels import pack_rows, poisson_fill
    seed_t = torch.from_numpy(seeds).to(dev)
    prob_t = torch.full((lanes,), float(p_in), dtype=torch.float32, device=dev)
    x_all = poisson_fill(seed_t, prob_t, torch.zeros(lanes, dtype=torch.int64, device=dev),
                         args.kick, W + K + KS, n)
    x_w, x_t, x_k = x_all[:W], x_all[W:W + K], x_all[W + K:]
    torch.cuda.synchronize()

    def words(z):
        return pack_rows(z.reshape(-1, n)).reshape(z.shape[0], lanes, -1)

    # The timed region is one call of the module, model(x) -> (z, states),
    # with x resident in HBM: the drop-in API a btorch user calls.
    with environ.context(dt=0.1), torch.no_grad():
        # the W warm-up steps run as two calls (W - 1 steps, then 1): the
        # first builds the plan, workspace and delay ring, the second takes
        # the path every later call takes, so the timed call is not the
        # first run of that host code
        t_first = time.perf_counter()
        parts = [x_w[:W - 1], x_w[W - 1:]] if W >= 2 else [x_w] if W else []
        r_parts = []
        for i, xp in enumerate(parts):
            r_parts.append(words(model(xp)[0]))
            if i == 0:
                torch.cuda.synchronize()
                first_call_ms = 1e3 * (time.perf_counter() - t_first)
        r_w = torch.cat(r_parts, 0) if r_parts else None
        if not parts:
            first_call_ms = 0.0
        torch.cuda.synchronize()
        # snapshot of the state entering the timed region (for the CPU
        # sample); reading the history writes it back here, before the clock.
        # The copies stay on the device until the region has ended: host
        # copies here would leave the GPU idle and the runtime's staging
        # state changed just before the clock
        snap = {"v": model.neuron.v.detach().clone(),
                "refractory": model.neuron.refractory.detach().clone(),
                "psc": model.synapse.base_psc.psc.detach().clone(),
                "history": model.synapse.history.history.detach().clone(),
                "cursor": model.synapse.history.cursor}
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev0.record()
        z_t, _ = model(x_t)
        host_ms = 1e3 * (time.perf_counter() - t0)
        ev1.record()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        region_ms = ev0.elapsed_time(ev1)
        r_t = words(z_t)
        snap = {k: (v.cpu() if torch.is_tensor(v) else v) for k, v in snap.items()}
        z_cpu_sample = z_t[:args.cpu_steps].cpu() if not args.no_cpu else None
        del z_t
        # the same launches (input x, dense spike record) with a HIP event
        # pair around every launch
        from btorch_amd.models.functional import prepare_sparse_modules
        timings: dict = {}
        with prepare_sparse_modules(model, lanes):
            r_k, _ = eng.run(x_k, KS, lanes, {"z"}, timings=timings)
        torch.cuda.synchronize()

    rasters = [r for r in (r_w, r_t, r_k) if r is not None]
    full = torch.cat(rasters, 0)
    tables = delivery_table(full, n, deg)
    ev_t, rows_t = window_events(tables, W, W + K)
    spikes_t = float(unpack_count(r_t))
    ev_k, rows_k = window_events(tables, W + K, W + K + KS)
    value = ev_t / elapsed
    ms_step = 1e3 * elapsed / K

    per = {k: (v[0] / max(v[1], 1), v[1], v[0]) for k, v in timings.items()
           if not k.endswith("_each")}
    pct = {k[:-5]: [round(float(np.percentile(v, q)) * 1e3, 2) for q in (50, 90, 99)]
           for k, v in timings.items() if k.endswith("_each") and v}
    nwords = (n + 63) // 64
    # algorithmic bytes of one step launch: F reads v, psc and refractory
    # and the input x and writes v and psc (24 B per neuron-lane; the
    # refractory counter is written back only where it changes, a few per
    # cent of neuron-lanes, not counted),
    # per-neuron tau, spike words to the raster and the ring, ring indices of
    # the spikes; the previous push's term records read (10 B per term: row,
    # weight, destination).  P: per active row the index and the int64
    # row_ptr pair (20 B), per term col + weight read and the term record
    # written (18 B).
    def step_bytes(spikes, rows, events, steps):
        # (the spike record is unpacked from the raster after the run: 4 B per
        # neuron-lane more for the call)
        return (24.0 * n * lanes + 4.0 * n + 16.0 * nwords * lanes + 4.0 * spikes / steps
                + 20.0 * rows / steps + 28.0 * events / steps)

    # the kernel: algorithmic bytes of a k_step launch over the mean duration
    # of the launches (a HIP event pair around each, on the launch stream) of
    # a KS-step run of the same launches (input tensor, dense spike record)
    bytes_k = step_bytes(float(unpack_count(r_k)), rows_k, ev_k, KS)
    step_us = per["step"][0] * 1e3 if per.get("step") else float("nan")
    achieved = bytes_k / (step_us * 1e-6)
    # the call: the timed model(x) call's region over its steps
    bytes_launch = step_bytes(spikes_t, rows_t, ev_t, K) + 4.0 * n * lanes
    call_us = region_ms / K * 1e3
    traffic = _traffic(args)
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": achieved / HBM_PEAK,
                "traffic": None if traffic is None else traffic["bytes_per_launch"],
                "traffic_source": None if traffic is None else traffic["source"],
                "kernel": "k_step (fused step + push + reduction)",
                "bytes_per_launch": bytes_k, "avg_us": step_us,
                "avg_us_source": f"HIP event pair around each k_step launch of a {KS}-step "
                                 f"run of the timed call's launches, on the launch stream",
                "per_launch_event_us": {k: round(v[0] * 1e3, 3) for k, v in per.items()},
                "per_launch_event_p50_p90_p99_us": pct,
                "call": {"us_per_step": call_us, "bytes_per_step": bytes_launch,
                         "frac": bytes_launch / (call_us * 1e-6) / HBM_PEAK,
                         "what": "the timed model(x) call: HIP events around the call / "
                                 "steps (includes the host's work before the first launch "
                                 "and the final reduction launch)"}}
    calls = {"host_enqueue_ms": round(host_ms, 3), "first_call_ms": round(first_call_ms, 2),
             "what": "host_enqueue_ms: wall time of the timed model(x) call until it "
                     "returned (work enqueued); first_call_ms: the first warm-up call "
                     "(W - 1 steps), which builds the plan, workspace and delay ring; "
                     "the last warm-up step runs as a call of its own",
             "warmup_calls": [int(p.shape[0]) for p in parts]}
    _progress(f"kernel: {ms_step:.4f} ms per step")
    sweep = activity_sweep(args, model, eng, deg, dev) if args.sweep else None
    _progress("activity sweep done")
    c0 = c0_bench(args, dev) if args.c0 else None
    _progress("C0 done")
    serving = serve_bench(args, model, dev) if args.serve_requests > 0 else None
    _progress("serving done")
    training = train_bench(args, dev) if args.train_steps > 0 else None
    _progress("training done")

    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        from oracle import rsnn
        x_cpu = rsnn.poisson_drive(seeds, np.full(lanes, p_in, dtype=np.float32),
                                   np.full(lanes, W, dtype=np.int64), args.cpu_steps, n,
                                   args.kick)
        cpu = cpu_baseline(args, conn, snap, x_cpu, z_cpu_sample, tables, W)
        cpu["input_equal_to_gpu_x"] = bool(torch.equal(x_cpu, x_t[:args.cpu_steps].cpu()))

    line = {
        "metric": "synaptic events/s at 140k-neuron connectome",
        "value": value,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms_step,
        "steps_per_s": K / elapsed,
        "events_per_step": ev_t / K,
        "active_fraction_per_step": spikes_t / (K * lanes * n),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded power-law connectome, Poisson background)",
        "config": {"workload": "connectome forward sim, N=140k LIF, power-law out-degree "
                               "(2.1, 100..5000), delays 1-20 (D=21), 8 lanes, dt 0.1 ms, "
                               "5 Hz Poisson background; timed: one model(x) call of "
                               f"T={K} steps (x resident in HBM, dense spike record out)",
                   "steps_per_call": K,
                   "kick": args.kick,
                   "n": n, "synapses": nnz, "lanes": lanes, "delay_bins": D,
                   "weights": f"lognormal(mu={args.w_mu}, sigma=0.5), inhibitory x4",
                   "accumulation": args.mode, "build_s": round(t_build, 2)},
        "roofline": roofline,
        "calls": calls,
        "cpu_baseline": cpu,
        "conn_build": {"device_ms": round(conn_build_ms, 2), "edges": nnz,
                       "what": "delay expansion, coalescing and CSR/CSC layout "
                               "(btx_conn_build)"},
        "serving": serving,
        "training": training,
        "activity_sweep": sweep,
        "plain_4096": c0,
    }
    print(json.dumps(line), flush=True)


SWEEP_PROBS = (0.001, 0.002, 0.005, 0.01, 0.02, 0.05, 0.1)

# The reference benchmark's published rows (RTX 4060 Laptop, Hemibrain,
# T=128, batch 1, CUDA-graph replay, median of 30): microseconds per step,
# native torch.sparse.mm and Triton, for the product alone and for
# RecurrentNN(LIF(hard reset), ExponentialPSC(SparseConn))
# (benchmarks/sparse_rnn/results/hemibrain_rtx4060.md:26-32, 45-51).
REFERENCE_HARNESS = {
    "spmspv": {0.001: (396.93, 46.25), 0.002: (396.94, 38.19), 0.005: (396.90, 55.34),
               0.01: (396.98, 75.33), 0.02: (396.90, 118.08), 0.05: (396.91, 247.03),
               0.1: (396.88, 445.70)},
    "snn": {0.001: (434.86, 72.56), 0.002: (435.30, 75.02), 0.005: (434.67, 88.82),
            0.01: (434.91, 108.74), 0.02: (434.69, 152.05), 0.05: (434.99, 279.18),
            0.1: (434.46, 480.68)},
}


def harness_bench(args, dev):
    """The reference benchmark's shape on this GPU (benchmark_triton_spmspv.py:
    104-330): a Hemibrain-sized synthetic connectome (21,739 neurons,
    3,550,403 edges: power-law out-degree with mean 163.3, uniform targets,
    lognormal(0, 0.5) positive weights normalised to a mean absolute fan-out
    of 0.01 as the harness's normalize_weights does; the real Hemibrain
    dataset is not available), T = 128, batch 1, Bernoulli spike traces at
    0.1 .. 10 %.  "spmspv": 128 spike vectors through SparseConn, the
    sequence captured in a HIP graph; "snn": RecurrentNN(LIF(tau 20, hard
    reset), ExponentialPSC(tau 5, SparseConn), cudagraph=True) over the trace
    x 1.05 at dt = 1 ms, the state reset before each sample.  10 warm-up and
    30 timed replays, median per step, as the harness."""
    import statistics

    from btorch_amd import workloads
    from btorch_amd.models import (
        LIF, ExponentialPSC, RecurrentNN, SparseConn, environ, functional,
    )
    from btorch_amd.models.functional import prepare_sparse_modules, reset_net

    n, E, T = 21_739, 3_550_403, 128
    rng = np.random.default_rng(args.seed + 5)
    deg = workloads.powerlaw_degrees(rng, n, 2.1, 40, 4000).astype(np.float64)
    deg = np.maximum(1, np.round(deg * (E / deg.sum()))).astype(np.int64)
    pre = np.repeat(np.arange(n, dtype=np.int64), deg)
    post = rng.integers(0, n - 1, size=pre.size, dtype=np.int64)
    post[post >= pre] += 1
    key = np.unique(pre * n + post)
    pre, post = key // n, key % n
    w = rng.lognormal(0.0, 0.5, pre.size).astype(np.float32)
    w *= np.float32(0.01 / max(float(np.abs(w).sum()) / n, 1.0))
    import scipy.sparse
    mat = scipy.sparse.coo_array((w, (pre, post)), shape=(n, n))

    def trace(rate):
        g = torch.Generator(device=dev).manual_seed(args.seed)
        return (torch.rand((T, 1, n), device=dev, generator=g) < rate).to(torch.float32)

    def time_replays(run, before=None):
        for _ in range(10):
            if before:
                before()
            run()
        torch.cuda.synchronize()
        out = []
        for _ in range(30):
            if before:
                before()
                torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            run()
            e1.record()
            e1.synchronize()
            out.append(e0.elapsed_time(e1))
        return statistics.median(out) * 1e3 / T

    rows = []
    conn = SparseConn(mat, enforce_dale=False, device=dev)
    lif = LIF(n, tau=20.0, hard_reset=True)
    model = RecurrentNN(lif, ExponentialPSC(n, 5.0, SparseConn(mat, enforce_dale=False,
                                                               device=dev)),
                        unroll=8, cudagraph=True, update_state_names=[]).to(dev)
    functional.init_net_state(model, batch_size=1, device=dev, dtype=torch.float32)
    for rate in SWEEP_PROBS:
        tr = trace(rate)
        # the product alone: the 128-step sequence captured in one graph
        # no_grad, not inference_mode: the generator state that graph capture
        # registers must stay usable by the module's own capture below
        with torch.no_grad(), prepare_sparse_modules(conn, 1):
            outs = []

            def seq():
                outs.clear()
                for t in range(T):
                    outs.append(conn(tr[t]))

            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                for _ in range(3):
                    seq()
            torch.cuda.current_stream().wait_stream(side)
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                seq()
            us_prod = time_replays(g.replay)
        # the recurrent net, graph-replayed per call, reset before each sample
        drive = tr * 1.05
        with torch.no_grad(), environ.context(dt=1.0):
            us_snn = time_replays(lambda: model(drive),
                                  before=lambda: reset_net(model, batch_size=1, inplace=True))
            z, _ = model(drive)
        ref_p, ref_s = REFERENCE_HARNESS["spmspv"][rate], REFERENCE_HARNESS["snn"][rate]
        rows.append({"rate": rate, "snn_rate": float(z.mean()),
                     "spmspv_us_per_step": us_prod, "snn_us_per_step": us_snn,
                     "reference_rtx4060_spmspv_native_triton_us": ref_p,
                     "reference_rtx4060_snn_native_triton_us": ref_s,
                     "snn_events_per_s": float(z.sum()) * float(E / n) / (us_snn * T * 1e-6)})
        _progress(f"harness rate {rate}: product {us_prod:.2f} us/step, snn {us_snn:.2f} us/step")
    return {"metric": "us per step at the reference benchmark's shape (Hemibrain-sized net)",
            "value": rows[0]["snn_us_per_step"], "unit": "us/step", "n_gpus": 1, "steps": T,
            "higher_is_better": False, "dtype": "f32", "vs_baseline": None,
            "data": "synthetic Hemibrain-sized connectome (dataset unavailable offline)",
            "config": {"workload": "reference harness shape: N=21,739, E=3.55M, T=128, batch 1, "
                                   "rates 0.1-10%, graph replay", "n": n,
                       "synapses": int(mat.nnz)},
            "rows": rows}


def activity_sweep(args, model, eng, deg, dev):
    """The step kernel across activity levels, as the reference's harness
    sweeps the spike rate of its product's input (0.1 % .. 10 %): each lane's
    neurons receive a one-step kick with probability p per step (in-kernel
    drive, raster only), 20 warm-up and 50 timed steps per level from the
    state the headline run left.  The refractory period (2 ms = 20 steps)
    caps activity near 5 %; the realised active fraction is reported."""
    from btorch_amd.models import environ
    from btorch_amd.models.functional import prepare_sparse_modules

    n, lanes = args.n, args.lanes
    seeds = (np.arange(lanes, dtype=np.int64) * 104729 + 77 + args.seed) & 0x7FFFFFFFFFFFFFFF
    out = []
    t_clock = 0
    with environ.context(dt=0.1), torch.no_grad(), prepare_sparse_modules(model, lanes):
        for prob in SWEEP_PROBS:
            drive = {"seed": torch.from_numpy(seeds).to(dev),
                     "prob": torch.full((lanes,), prob, dtype=torch.float32, device=dev),
                     "t0": torch.full((lanes,), t_clock, dtype=torch.int64, device=dev),
                     "kick": args.kick}
            r_w, _ = eng.run(None, 20, lanes, set(), drive=drive)
            drive["t0"] += 20
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            r, _ = eng.run(None, 50, lanes, set(), drive=drive)
            e1.record()
            torch.cuda.synchronize()
            t_clock += 70
            ms = e0.elapsed_time(e1)
            tables = delivery_table(torch.cat([r_w, r], 0), n, deg)
            ev, _ = window_events(tables, 20, 70)
            spikes = float(unpack_count(r))
            out.append({"drive_prob": prob, "active_fraction": spikes / (50 * lanes * n),
                        "events_per_step": ev / 50, "us_per_step": 1e3 * ms / 50,
                        "events_per_s": ev / (ms * 1e-3)})
    return out


def c0_bench(args, dev):
    """The reference's CPU-runnable configuration at full size: N = 4096 LIF
    (80/20 E/I, ER p = 0.02, lognormal(-1, 0.5), inhibitory x4), exponential
    synapse 5 ms, 1000 Poisson channels at 10 Hz through a fixed 0.05-density
    projection (neuron_inp_module), B = 1, T = 1000, dt = 0.1 ms.  The GPU
    runs model(x) (projection and the fused engine); the CPU oracle runs the
    same T = 1000 steps, and the two rasters are compared."""
    from btorch_amd import workloads
    from btorch_amd.models import LIF, ExponentialPSC, RecurrentNN, SparseConn, environ, functional
    from oracle import rsnn

    n, n_in, B, T = 4096, 1000, 1, 1000
    c = workloads.plain_lif_net(n=n, p=0.02, seed=0)
    conn = c.matrix()
    win = workloads.input_weights(n_in, n, 0.05, 0)
    model = RecurrentNN(LIF(n, tau=20.0, tau_ref=2.0), ExponentialPSC(n, 5.0, SparseConn(conn)),
                        neuron_inp_module=SparseConn(win, enforce_dale=False), step_mode="m",
                        update_state_names=[]).to(dev)
    xin = torch.from_numpy(workloads.poisson_input(T, B, n_in, 10.0, 0.1, seed=0))
    xd = xin.to(dev)
    rp = model.synapse.linear._csr_row_ptr
    deg_n = (rp[1:] - rp[:-1]).to(torch.float64)
    res = {}
    with environ.context(dt=0.1), torch.no_grad():
        for rep in range(2):  # the second call is timed
            functional.init_net_state(model, batch_size=B, device=dev, dtype=torch.float32)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            z, _ = model(xd)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
    ev = float((z.reshape(T, n).to(torch.float64) @ deg_n).sum())
    res["gpu"] = {"s": el, "steps_per_s": T / el, "events_per_s": ev / el,
                  "spikes": int(z.sum())}
    threads = len(os.sched_getaffinity(0))
    torch.set_num_threads(threads)
    x_proj = rsnn.NativeSparse(win)(xin.reshape(T * B, n_in)).reshape(T, B, n)
    net = rsnn.Net(n, conn, rsnn.LIFParams.make(n, tau=20.0, tau_ref=2.0), torch.tensor(5.0), 0.1)
    t0 = time.perf_counter()
    ref = rsnn.simulate(net, x_proj)
    el_cpu = time.perf_counter() - t0
    res["cpu_baseline"] = {"value": ev / el_cpu, "unit": "events/s", "cores": threads,
                           "kind": "port", "s": el_cpu, "steps_per_s": T / el_cpu,
                           "sample": "the full T = 1000 run (torch-CPU oracle, input "
                                     "projection included)"}
    res["raster_bit_exact_vs_cpu"] = bool(torch.equal(z.cpu().reshape(T, B, n), ref["z"]))
    res["workload"] = ("N=4096 LIF, ER p=0.02 (~335k synapses), 1000 Poisson inputs at 10 Hz, "
                       "B=1, T=1000, dt=0.1 ms")
    res["events"] = ev
    return res


def _private_net(n, seed, dev):
    from btorch_amd import workloads
    from btorch_amd.models import LIF, ExponentialPSC, RecurrentNN, SparseConn

    # the shared connectome's 70/30 E/I split: with 80/20, inhibition x4
    # balances excitation on average and the 16k/32k nets run at ~2% active
    # per step, forty times the configured background
    c = workloads.er_net_fast(n=n, p=0.02, exc_frac=0.7, seed=seed, w_mu=-3.7)
    idx = torch.from_numpy(np.stack([c.pre, c.post])).to(dev)
    coo = torch.sparse_coo_tensor(idx, torch.from_numpy(c.w).to(dev), size=(n, n))
    return RecurrentNN(LIF(n, tau_ref=2.0), ExponentialPSC(n, 5.0, SparseConn(coo, device=dev)),
                       step_mode="m", update_state_names=[]).to(dev)


def serve_bench(args, model, dev):
    """Continuous batching vs one job at a time on the shared connectome plus
    private 4k-32k nets (every fifth request).  Returns the JSON object."""
    from btorch_amd.serving import LanePool, Server, SimRequest, run_plain

    rng = np.random.default_rng(args.seed + 3)
    R = args.serve_requests
    arrivals = np.cumsum(rng.exponential(1.0 / args.serve_rate, R))
    nets = {"a": 4096, "b": 16384, "c": 32768}
    reqs = []
    for i in range(R):
        private = rng.random() < 0.2
        reqs.append(SimRequest(i, T=int(rng.integers(200, args.serve_tmax + 1)),
                               batch=int(rng.integers(1, 33)),
                               rate_scale=float(rng.uniform(0.5, 2.0)),
                               seed=int(rng.integers(1 << 40)), arrival=float(arrivals[i]),
                               net=str(rng.choice(list(nets))) if private else "shared"))
    t_build = time.perf_counter()
    models = {k: _private_net(v, 100 + j, dev) for j, (k, v) in enumerate(nets.items())}
    models["shared"] = model
    D = args.delay_max + 1
    chunk = {"shared": 10 * (D + 1)}
    kw = dict(n_sample=16, base_rate_hz=5.0, kick=args.kick)
    shared = LanePool(model, lanes=args.serve_lanes, chunk=chunk["shared"], graph=True, **kw)

    # the private nets' pools are built up front, like the plain path's pools
    private_pools = {key: LanePool(models[key], lanes=64, chunk=200, stream=torch.cuda.Stream(),
                                   graph=True, **kw) for key in nets}

    def private(key):
        return private_pools[key]

    # plain path: one job at a time on a pool sized to the job (built up front)
    plain_pools = {}
    for r in reqs:
        key = (r.net, r.batch)
        if key not in plain_pools:
            plain_pools[key] = LanePool(models[r.net], lanes=r.batch,
                                        chunk=chunk.get(r.net, 200), graph=True, **kw)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t_build

    # admission by HBM budget: every shared lane plus a quarter of that for
    # the private nets' pools (bytes of resident state and delay ring)
    budget = int(1.25 * args.serve_lanes * shared.bytes_per_lane())
    mix = Server(shared, private, hbm_budget=budget).serve(reqs, realtime=True)
    plain = run_plain(lambda r: plain_pools[(r.net, r.batch)], reqs)
    lane_steps = sum(r.T * r.batch for r in reqs)

    def summ(st):
        return {"wall_s": round(st.wall, 3), "lane_steps_per_s": lane_steps / st.wall,
                "p50_latency_s": round(st.pct(50), 4), "p99_latency_s": round(st.pct(99), 4),
                "completed": len(st.results), "rejected": len(st.rejected)}

    m, pl = summ(mix), summ(plain)
    return {"requests": R, "arrival_rate_per_s": args.serve_rate,
            "trial_steps": f"U(200, {args.serve_tmax})", "batch": "U{1..32}",
            "private_nets": "20% of requests on ER p=0.02 nets of 4096/16384/32768 neurons",
            "shared_pool_lanes": args.serve_lanes, "lane_steps": lane_steps,
            "hbm_budget_bytes": budget,
            "continuous": m, "plain": pl,
            "throughput_vs_plain": m["lane_steps_per_s"] / pl["lane_steps_per_s"],
            "build_s": round(t_build, 2)}


def train_bench(args, dev):
    """One training step of the differentiable RSNN through the hand-written
    adjoint: N=16,384 LIF, Erdos-Renyi p=0.05 (13.4M synapses, weights
    N(0, 1/sqrt(pN))), 200 Poisson input channels with per-class rate
    patterns U(0, 50) Hz through a fixed 5%-density lognormal projection
    (scaled by 0.2), 10 classes,
    a linear readout of spike counts, cross-entropy, fast-sigmoid surrogate
    (slope 25), batch 64, T=200, dt=1 ms, recompute checkpointing every 50
    steps; gradients of the synaptic weights and per-neuron tau_m, one SGD
    update.  Returns the JSON object."""
    import math

    from btorch_amd import workloads
    from btorch_amd.models import (
        LIF, ExponentialPSC, RecurrentNN, SparseConn, SuperSpike, environ, functional,
    )

    n, p, B, T, C, n_in = 16384, 0.05, 64, 200, 10, 200
    rng = np.random.default_rng(args.seed + 2)
    c = workloads.er_net_fast(n=n, p=p, seed=args.seed + 2)
    w = rng.normal(0.0, 1.0 / math.sqrt(p * n), c.nnz).astype(np.float32)
    idx = torch.from_numpy(np.stack([c.pre, c.post])).to(dev)
    rec = SparseConn(torch.sparse_coo_tensor(idx, torch.from_numpy(w).to(dev), size=(n, n)),
                     enforce_dale=False, device=dev)
    win = workloads.input_weights(n_in, n, 0.05, args.seed)
    win.data *= np.float32(0.2)
    inp = SparseConn(win, enforce_dale=False, device=dev)
    inp.magnitude.requires_grad_(False)
    tau = torch.from_numpy(rng.uniform(10.0, 30.0, n).astype(np.float32))
    sur = SuperSpike(alpha=25.0 / (math.sqrt(2) - 1))
    model = RecurrentNN(LIF(n, tau=tau, tau_ref=2.0, trainable_param={"tau"},
                            surrogate_function=sur),
                        ExponentialPSC(n, 5.0, rec), neuron_inp_module=inp, step_mode="m",
                        grad_checkpoint=True, chunk_size=50, unroll=50,
                        update_state_names=[]).to(dev)
    readout = torch.nn.Linear(n, C).to(dev)
    params = [rec.magnitude, model.neuron.tau, *readout.parameters()]
    opt = torch.optim.SGD(params, lr=1e-3)
    rates = torch.from_numpy(rng.uniform(0.0, 50.0, (C, n_in)).astype(np.float32)).to(dev)
    g = torch.Generator(device=dev).manual_seed(args.seed)

    def batch():
        labels = torch.randint(0, C, (B,), device=dev, generator=g)
        prob = rates[labels] * 1e-3  # per 1 ms step
        x = (torch.rand((T, B, n_in), device=dev, generator=g) < prob).float()
        return x, labels

    split = {"forward": 0.0, "backward": 0.0}

    def step():
        x, labels = batch()
        functional.init_net_state(model, batch_size=B, device=dev, dtype=torch.float32)
        opt.zero_grad(set_to_none=True)
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record()
        with environ.context(dt=1.0):
            z, _ = model(x)
        logits = readout(z.mean(0))  # readout of spike counts / T
        loss = torch.nn.functional.cross_entropy(logits, labels)
        e1.record()
        loss.backward()
        e2.record()
        opt.step()
        e2.synchronize()
        split["forward"] += e0.elapsed_time(e1)
        split["backward"] += e1.elapsed_time(e2)
        return float(loss.detach()), float(z.detach().mean())

    step()  # warm-up
    torch.cuda.synchronize()
    split["forward"] = split["backward"] = 0.0
    t0 = time.perf_counter()
    out = [step() for _ in range(args.train_steps)]
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / args.train_steps
    fwd_ms = split["forward"] / args.train_steps
    bwd_ms = split["backward"] / args.train_steps
    # algorithmic bytes of the backward pass per step: the recompute of the
    # checkpointed forward (at this activity a dense pull: the CSC edge
    # stream, 8 B per edge; state and the three records, 44 B per
    # neuron-lane) and the reverse step (k_adj_transpose: col + w, 8 B per
    # edge; k_adj_wgrad: col read, gw read and written, 12 B per edge;
    # k_adj_lif: 11 arrays, 44 B per neuron-lane)
    bwd_bytes = T * (28.0 * c.nnz + 88.0 * n * B)
    bwd_achieved = bwd_bytes / (bwd_ms * 1e-3)
    cpu = None
    if not args.no_cpu:
        cpu = train_cpu_baseline(model, x_sample=batch()[0], T_s=10)
    return {"workload": "N=16384 LIF, ER p=0.05 (13.4M synapses), 200 Poisson inputs, "
                        "10 classes, batch 64, T=200, dt=1 ms, checkpoint every 50 steps",
            "synapses": int(c.nnz), "s_per_step": el, "samples_per_s": B / el,
            "neuron_steps_per_s": B * T * n / el, "loss": out[-1][0],
            "mean_rate_per_step": out[-1][1],
            "grads": "synaptic weights (13.4M), tau_m (16384), readout",
            "forward_ms": fwd_ms, "backward_ms": bwd_ms,
            "roofline": {"bound": "hbm", "kernel": "backward pass (chunk recomputes + reverse "
                                                  "sweep: k_step, k_adj_transpose, k_adj_wgrad, "
                                                  "k_adj_lif)",
                         "achieved": bwd_achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                         "frac": bwd_achieved / HBM_PEAK, "bytes_per_iteration": bwd_bytes,
                         "avg_ms": bwd_ms,
                         "avg_ms_source": "HIP events around loss.backward() of the timed "
                                          "iterations"},
            "cpu_baseline": cpu}


def train_cpu_baseline(model, x_sample, T_s: int):
    """The reference-style autograd (torch-CPU oracle) of the same net on a
    bounded sample: forward and backward over the first T_s steps of one
    batch, scaled to the T = 200 iteration."""
    from oracle import rsnn

    threads = len(os.sched_getaffinity(0))
    torch.set_num_threads(threads)
    conn = model.synapse.linear
    n = conn.out_features
    dst, src = conn.indices[0].cpu(), conn.indices[1].cpu()
    mag = conn.magnitude.detach().cpu().clone().requires_grad_(True)
    tau = model.neuron.tau.detach().cpu().clone().requires_grad_(True)
    x = model.neuron_inp_module(x_sample[:T_s].reshape(-1, x_sample.shape[-1])).detach()
    x = x.reshape(T_s, -1, n).cpu()
    sur = model.neuron.surrogate_function
    t0 = time.perf_counter()
    z, v, _, psc = rsnn.simulate_grad(n, dst, src, mag, x, tau, 1.0, 5.0, tau_ref=2.0,
                                      surrogate=("superspike", sur.alpha, 1.0, True))
    (z.mean(0).sum() + v.sum() * 1e-3).backward()
    el = time.perf_counter() - t0
    B = x.shape[1]
    per_iter = el * 200 / T_s
    return {"value": B / per_iter, "unit": "samples/s", "cores": threads, "kind": "port",
            "s_per_iteration": per_iter,
            "sample": f"forward + backward of the first {T_s} of the 200 steps of one batch "
                      f"(torch-CPU oracle autograd, {threads} threads), scaled to T = 200"}


def cpu_baseline(args, conn, snap, x_cpu, gpu_z, tables, W):
    """Time the CPU oracle on the first ``cpu_steps`` timed steps from the
    same state; check its raster against the GPU's."""
    from oracle import rsnn

    n, D = args.n, args.delay_max + 1
    threads = len(os.sched_getaffinity(0))
    torch.set_num_threads(threads)
    dst = conn.indices[0].cpu().numpy()
    src = conn.indices[1].cpu().numpy()
    w = conn.magnitude.detach().cpu().numpy()
    mm = rsnn.NativeSparse.from_coalesced(dst, src, w, n * D, n)
    rng = np.random.default_rng(args.seed + 7)
    tau = torch.from_numpy(rng.uniform(10.0, 30.0, n).astype(np.float32))
    net = rsnn.Net(n, mm, rsnn.LIFParams.make(n, tau=tau, tau_ref=2.0), torch.tensor(5.0), 0.1,
                   delay_mode="hetero", max_delay_steps=D, circular=True)
    S = x_cpu.shape[0]
    t0 = time.perf_counter()
    out = rsnn.simulate(net, x_cpu, state=snap)
    el = time.perf_counter() - t0
    ev, _ = window_events(tables, W, W + S)
    return {"value": ev / el, "unit": "events/s", "cores": threads, "kind": "port",
            "sample": f"{S} steps of the same workload from the timed region's entry state "
                      f"(torch-CPU oracle, {threads} threads)",
            "steps_per_s": S / el, "raster_bit_exact_vs_gpu": bool(torch.equal(out["z"], gpu_z))}


if __name__ == "__main__":
    main()
