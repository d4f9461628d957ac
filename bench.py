"""Benchmark of the hip jagged-array path on one MI355X.

Workload (one "step"): the jagged event pipeline on 50M events of records
{pt, eta, phi: float32, charge: int8}, multiplicity ~ Poisson(3) clipped to
[0, 20], seed 2:

    mask = (ev.pt > 25) & (abs(ev.eta) < 2.4)
    sel  = ev[mask][:, 1:]
    ak.num(sel), ak.sum(sel.pt, axis=-1), ak.argmax(sel.pt, axis=-1)

Inputs are resident in HBM before timing.  `value` is events/s of the whole
pipeline in its fused form (one kernel, replayed from a hipGraph); `forms`
also times the reference's per-kernel sequence ("plain") and the per-kernel
path with tuned reductions.  `roofline` is the HBM roofline of the fused
form's kernel (algorithmic bytes / its average duration, from HIP events
recorded around each replay on the stream it runs on).  `cpu_baseline` times the same
pipeline through the same host layer on the single-threaded CPU oracle over a
bounded sample of the events.  `segmented_reductions` reports the segmented reductions of
the 1M-list Poisson(10) float64 layout, HBM-cold (a read-only 1 GiB scrub
before each timed launch).

    python bench.py [--steps K] [--warmup W] [--events N] [--no-cpu]
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "segmented-op GB/s & % HBM roofline; events/s batched+learned vs plain"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_events(n_events, seed=2, lam=3, clip=20):
    import numpy as np

    rng = np.random.default_rng(seed)
    mult = np.clip(rng.poisson(lam, n_events), 0, clip).astype(np.int64)
    off = np.zeros(n_events + 1, dtype=np.int64)
    np.cumsum(mult, out=off[1:])
    N = int(off[-1])
    fields = {
        "pt": (5 + rng.exponential(20, N)).astype(np.float32),
        "eta": rng.uniform(-2.5, 2.5, N).astype(np.float32),
        "phi": rng.uniform(-np.pi, np.pi, N).astype(np.float32),
        "charge": (2 * rng.integers(0, 2, N) - 1).astype(np.int8),
    }
    return off, fields


def pipeline(ak, ev):
    mask = (ev.pt > 25) & (abs(ev.eta) < 2.4)
    sel = ev[mask][:, 1:]
    return ak.num(sel), ak.sum(sel.pt, axis=-1), ak.argmax(sel.pt, axis=-1)


def run_segmented_reductions(ak, backend, torch, reps=5):
    """Segmented reductions on 1M Poisson(10) float64 lists, HBM-cold."""
    import numpy as np

    rng = np.random.default_rng(1)
    counts = rng.poisson(10, 1_000_000)
    off = np.zeros(len(counts) + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    N = int(off[-1])
    x = rng.standard_normal(N)
    n = len(counts)
    dev_off = torch.from_numpy(off).cuda()
    dev_x = torch.from_numpy(x).cuda()
    scrub = torch.empty(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB > 256 MiB MALL
    out_f = torch.empty(n, dtype=torch.float64, device="cuda")
    out_i = torch.empty(n, dtype=torch.int64, device="cuda")
    starts = dev_off[:-1]
    ops = {
        "sum": (("awkward_reduce_sum", np.float64, np.float64, np.int64), (out_f, dev_x, dev_off, n),
                8 * (n + 1) + 8 * N + 8 * n),
        "max": (("awkward_reduce_max", np.float64, np.float64, np.int64), (out_f, dev_x, dev_off, n, -np.inf),
                8 * (n + 1) + 8 * N + 8 * n),
        "argmax": (("awkward_reduce_argmax", np.int64, np.float64, np.int64, np.int64),
                   (out_i, dev_x, dev_off, starts, n), 8 * (n + 1) + 8 * N + 8 * n),
        "count": (("awkward_reduce_count_64", np.int64, np.int64), (out_i, dev_off, n), 8 * (n + 1) + 8 * n),
    }
    from awkward_amd import dispatch

    v = dispatch.heuristic_variant(n, N, 8)
    mask = torch.empty(n, dtype=torch.int8, device="cuda")
    ops["sum_tuned"] = (lambda: dispatch.reduce_sum(out_f, dev_x, dev_off, n, v), None, 8 * (n + 1) + 8 * N + 8 * n)
    ops["argmax_tuned"] = (lambda: dispatch.reduce_arg("argmax", out_i, mask, dev_x, dev_off, n, v), None,
                           8 * (n + 1) + 8 * N + 9 * n)
    result = {"lists": n, "elements": N, "dtype": "float64", "variant": v,
              "method": "HBM-cold: a read-only 1 GiB scrub (sum) before each launch evicts L2 and MALL without "
                        "leaving dirty lines; *_tuned = fused entry point, heuristic tile; read_ceiling = a plain "
                        "16-byte streaming read of the same 80 MB content buffer, same scrub; *_dirty_scrub = after a read-modify-write "
                        "scrub, whose 256 MiB of dirty lines are written back during the timed kernel"}
    from awkward_amd.tuning import time_variants

    sweep = time_variants("sum", out_f, dev_x, dev_off, n, reps=reps, scrub=scrub)
    sweep_arg = time_variants("argmax", out_i, dev_x, dev_off, n, mask=mask, reps=reps, scrub=scrub)
    result["sweep_sum_ms"] = {str(k): round(t, 4) for k, t in sweep.items()}
    result["sweep_argmax_ms"] = {str(k): round(t, 4) for k, t in sweep_arg.items()}
    vb = min(sweep, key=sweep.get)
    result["best_variant_sum"] = int(vb)
    ops["sum_best"] = (lambda: dispatch.reduce_sum(out_f, dev_x, dev_off, n, vb), None, 8 * (n + 1) + 8 * N + 8 * n)
    from awkward_amd._kernels import KernelError
    from awkward_amd._lib import lib

    read = lib().awkward_hip_stream_read
    read.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    read.restype = KernelError
    sink = torch.zeros(1, dtype=torch.int64, device="cuda")
    ops["read_ceiling"] = (lambda: read(dev_x.data_ptr(), 8 * N, sink.data_ptr()), None, 8 * N)
    dispatch.reduce_sum(out_f, dev_x, dev_off, n, v)
    torch.cuda.synchronize()
    ACCURACY["segmented_reductions"] = (off, x, out_f.cpu().numpy())
    for name, (key, args, nbytes) in list(ops.items()) + [("sum_best_dirty_scrub", ops["sum_best"])]:
        if args is None:
            call = key
        else:
            kern = backend[key]
            call = (lambda k, a: (lambda: k(*a)))(kern, args)
        call()
        times = []
        for _ in range(reps):
            if name.endswith("_dirty_scrub"):
                scrub.add_(1.0)
            else:
                scrub.sum()
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record()
            call()
            e.record()
            torch.cuda.synchronize()
            times.append(s.elapsed_time(e))
        ms = sorted(times)[len(times) // 2]
        gbs = nbytes / (ms * 1e-3) / 1e9
        result[name] = {"ms": round(ms, 4), "GB/s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 3),
                        "bytes": nbytes}
    return result


def pairs_step(ak, np, ev):
    pairs = ak.combinations(ev, 2)
    p0, p1 = pairs["0"], pairs["1"]
    expr = 2 * p0.pt * p1.pt * (np.cosh(p0.eta - p1.eta) - np.cos(p0.phi - p1.phi))
    triples = ak.combinations(ev, 3)
    return ak.max(expr, axis=-1), ak.argmax(expr, axis=-1), pairs, triples


def pairs_plan(ev):
    """The same step recorded lazily: the reductions of the pair expression
    become one awkward_hip_pair_reduce launch, the n=2 and n=3 carries two
    decodes -- the same four outputs as pairs_step."""
    import numpy as np

    from awkward_amd.lazy import Plan, lazy

    lz = lazy(ev)
    pairs = lz.combinations(2)
    p0, p1 = pairs["0"], pairs["1"]
    expr = 2 * p0.pt * p1.pt * (np.cosh(p0.eta - p1.eta) - np.cos(p0.phi - p1.phi))
    return Plan([expr.max(), expr.argmax(), pairs, lz.combinations(3)])


def pairs_expr_device(ev):
    """The pair expression as a leg-aware device expression over ev's columns
    (what the lazy plan compiles for pairs_plan)."""
    from awkward_amd import lazy as L

    plan = pairs_plan(ev)
    return next(iter(plan.pair_groups.values()))["expr"]


def same_outputs(torch, a, b):
    """Bit equality of pairs_step's four outputs (masked max / argmax, two
    combinations layouts) between two forms."""
    for x, y in zip(a[:2], b[:2]):
        lx, ly = x.layout, y.layout
        if not (torch.equal(lx.mask.to(torch.int8), ly.mask.to(torch.int8)) and
                torch.equal(lx.content.data, ly.content.data)):
            return False
    for x, y in zip(a[2:], b[2:]):
        lx, ly = x.layout, y.layout
        if not torch.equal(lx.offsets, ly.offsets):
            return False
        if not all(torch.equal(cx.index, cy.index) for cx, cy in zip(lx.content.contents, ly.content.contents)):
            return False
    return True


def run_combinations(ak, backend, torch, n_events=20_000_000, steps=5):
    """Combinations n=2 and n=3 on 20M events (Poisson(4) clipped [0, 30],
    seed 3), the float32 pair expression, max/argmax over pairs per event;
    both forms write the same outputs (the masked max and argmax, the n=2 and
    the n=3 combinations)."""
    import numpy as np

    from awkward_amd.fused import Combinations, PairReduce

    off, fields = make_events(n_events, seed=3, lam=4, clip=30)
    fields = {k: fields[k] for k in ("pt", "eta", "phi")}
    ev = ak.from_offsets(off, fields, backend)
    torch.cuda.synchronize()
    out = {"events": n_events, "particles": int(off[-1])}
    out["plain"] = time_form(torch, lambda: pairs_step(ak, np, ev), steps, 1, n_events)
    plan = pairs_plan(ev)
    stats = plan.stats()
    assert stats["eager_outputs"] == 0 and stats["fused_launches"] == 1, stats
    fused_out = plan.run()
    plain_out = pairs_step(ak, np, ev)
    out["outputs_equal"] = same_outputs(torch, fused_out, plain_out)
    del plain_out, fused_out
    unit = plan._unit(next(iter(plan.units)))
    out["pairs"], out["triples"] = unit.tuples[2].total, unit.tuples[3].total
    graph = plan.graph()
    out["fused_graph"] = time_form(torch, graph.replay, steps, 1, n_events)
    out["speedup"] = round(out["fused_graph"]["events_per_s"] / out["plain"]["events_per_s"], 2)
    # the parts: the fused launch with its two length scans (writes the 2.56 GB
    # of pairs and 5.12 GB of triples while reducing the expression), the pair
    # reductions alone (ALU-bound) and the triples' decode alone (write-bound)
    pr = PairReduce(ev.layout.offsets, pairs_expr_device(ev), ("max", "argmax"))
    tri = Combinations(ev, 3)
    for name, fn, nbytes in (("combinations_pairs", unit.run, unit.algorithmic_bytes()), ("pair_reduce", pr.run, None),
                             ("combinations3", tri.run, tri.algorithmic_bytes())):
        fn()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        torch.cuda.synchronize()
        ms = s.elapsed_time(e)
        out[name] = {"ms": round(ms, 4)}
        if nbytes:
            out[name].update({"GB/s": round(nbytes / ms / 1e6, 1), "bytes": nbytes})
    out["what"] = ("plain: ak.combinations(ev,2) + the pair expression as one expression-VM pass over the carried "
                   "legs + ak.max/argmax + ak.combinations(ev,3); fused_graph: the same graph recorded with "
                   "lazy(ev) -- two combinations_length scans and one awkward_hip_combinations_pairs launch that "
                   "writes the n=2 and n=3 carries and reduces the expression over the pairs in place, replayed "
                   "from a hipGraph")
    return out


def make_batch(seed=4, narrays=256, min_exp=3.0, max_exp=5.0, zipf_cap=100_000):
    """The batched workload's generator (seed 4): per array a list count 10**U(min_exp,
    max_exp), lengths Poisson(2) / Poisson(50) / Zipf(1.5) capped, 5% of the
    arrays all-empty, dtype int32 / int64 / float32 / float64, values
    integers(-1000, 1000) or standard normal."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(narrays):
        n = int(10 ** rng.uniform(min_exp, max_exp))
        kind = int(rng.integers(0, 3))
        if kind == 0:
            counts = rng.poisson(2, n)
        elif kind == 1:
            counts = rng.poisson(50, n)
        else:
            counts = np.minimum(rng.zipf(1.5, n), zipf_cap)
        if rng.random() < 0.05:
            counts = np.zeros(n, dtype=np.int64)
        off = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(counts, out=off[1:])
        dt = [np.int32, np.int64, np.float32, np.float64][int(rng.integers(0, 4))]
        N = int(off[-1])
        x = rng.integers(-1000, 1000, N).astype(dt) if np.dtype(dt).kind == "i" else rng.standard_normal(N).astype(dt)
        out.append((off, x))
    return out


def run_batched(ak, torch, steps=3):
    """Batched uneven workload: per array sum, min, argmin and argsort.
    plain = the registry path array by array; batched = the persistent
    work-queue kernels (autotuned tiles) replayed from one hipGraph."""
    import numpy as np

    from awkward_amd.backend import HipBackend
    from awkward_amd.batch import BatchRunner
    from awkward_amd.fused import Graph

    batch = make_batch()
    n_lists = int(sum(len(o) - 1 for o, _ in batch))
    n_elem = int(sum(len(x) for _, x in batch))
    be = HipBackend(fused=False)
    arrays = [ak.from_offsets(o, x, be) for o, x in batch]
    torch.cuda.synchronize()

    def plain():
        for a in arrays:
            ak.sum(a, axis=-1)
            ak.min(a, axis=-1)
            ak.argmin(a, axis=-1)
            ak.argsort(a, axis=-1)

    out = {"arrays": len(batch), "lists": n_lists, "elements": n_elem,
           "sample": "list counts log-uniform in [1e3, 1e5] (the configuration's [1e3, 1e7] would hold ~2.6e10 "
                     "elements); every other parameter as configured, seed 4"}
    out["plain"] = time_form(torch, plain, steps, 1, n_lists)
    dev = [(a.layout.offsets, a.layout.content.data) for a in arrays]
    runner = BatchRunner(dev)
    out["autotune"] = runner.tune(reps=3)
    graph = Graph(runner.run)
    out["batched_graph"] = time_form(torch, graph.replay, steps * 3, 1, n_lists)
    # the same batch with the tile chosen on the device by the learned dispatcher
    learned_runner = BatchRunner(dev, variants="learned")
    lgraph = Graph(learned_runner.run)
    out["batched_learned_graph"] = time_form(torch, lgraph.replay, steps * 3, 1, n_lists)
    out["learned_tiles"] = {g.name: g.chosen() for g in learned_runner.groups}
    for form in ("plain", "batched_graph", "batched_learned_graph"):
        out[form]["lists_per_s"] = out[form].pop("events_per_s")
    out["speedup"] = round(out["batched_graph"]["lists_per_s"] / out["plain"]["lists_per_s"], 2)
    out["speedup_learned"] = round(out["batched_learned_graph"]["lists_per_s"] / out["plain"]["lists_per_s"], 2)
    ms = out["batched_graph"]["ms_per_step"]
    out["batched_GB/s"] = round(runner.algorithmic_bytes() / ms / 1e6, 1)
    return out


def run_dispatch_regret(torch, budget_s=40.0):
    """Learned dispatch: the trained cost model's regret against the
    exhaustive sweep (every tile timed) on held-out arrays -- seeds 9000+,
    drawn on the device by the training generators -- beside the autotuner
    table, the fixed heuristic and the default tile."""
    from awkward_amd import learned

    device = torch.device("cuda", torch.cuda.current_device())
    t0 = time.perf_counter()
    X, T, _ = learned.collect(list(range(9000, 9500)), 20, device, budget_s=budget_s)
    sweep_s = time.perf_counter() - t0
    model = learned.Model.load(learned.DEFAULT_MODEL)
    out = learned.regret(model, X, T, device)
    out["sweep_s"] = round(sweep_s, 1)
    out["model"] = {k: model.info.get(k) for k in ("rows", "epochs", "val_regret", "train_s")}
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "event_pipeline_traffic.json")
KERNEL_SOURCE = os.path.join(ROOT, "awkward_amd", "csrc", "fused.hip")


def source_digest():
    import hashlib

    with open(KERNEL_SOURCE, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def measured_traffic(events):
    """HBM bytes per launch of the fused pipeline kernel from the committed
    rocprofv3 counter runs (FETCH_SIZE and WRITE_SIZE, each scaled by the
    same counter's reading for a copy of known size), if they were taken on
    this kernel source and this many events; else None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("source_sha256") != source_digest() or t.get("events") != events:
        return None
    return t.get("traffic_bytes")


def traffic_probe(args):
    """The program the counter runs profile: an elementwise copy of known
    size (1 GiB read, 1 GiB written), then the fused pipeline launched
    eagerly, each three times.  Prints the algorithmic bytes."""
    import torch

    import awkward_amd as ak
    from awkward_amd.backend import HipBackend
    from awkward_amd.fused import EventPipeline

    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    off, fields = make_events(args.events)
    ev = ak.from_offsets(off, fields, HipBackend(fused=True))
    pipe = EventPipeline(ev, [("pt", ">", 25.0), ("eta", "abs<", 2.4)], slice(1, None), "pt",
                         outputs=("num", "sum", "argmax"))
    for _ in range(3):
        pipe.run()
    torch.cuda.synchronize()
    print(json.dumps({"copy_read_bytes": 4 << 28, "copy_write_bytes": 4 << 28, "events": args.events,
                      "pipeline_alg_bytes": pipe.algorithmic_bytes(int(off[-1])), "source_sha256": source_digest()}),
          flush=True)


def roofline_probe(args):
    """The program the per-kernel counter runs profile: the 1 GiB copy the
    counters are calibrated on, then every workload's kernels once, eagerly
    (graph replays are not profiled per dispatch): the 1M-list reductions
    (plain and tuned), the event pipeline (plain, tuned and fused), the
    20M-event combinations (plain and fused) and a 256-array batch (sum/min/argmin
    and argsort)."""
    import numpy as np
    import torch

    import awkward_amd as ak
    from awkward_amd.backend import HipBackend
    from awkward_amd.batch import BatchRunner
    from awkward_amd.fused import EventPipeline

    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    del a, b
    plain, tuned = HipBackend(fused=False), HipBackend(fused=True)
    rng = np.random.default_rng(1)
    counts = rng.poisson(10, 1_000_000)
    off = np.zeros(len(counts) + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    x = rng.standard_normal(int(off[-1]))
    for be in (plain, tuned):
        arr = ak.from_offsets(off, x, be)
        for fn in (ak.sum, ak.max, ak.argmax, ak.count):
            fn(arr, axis=-1)
    torch.cuda.synchronize()
    off2, fields = make_events(args.events)
    for be in (plain, tuned):
        pipeline(ak, ak.from_offsets(off2, fields, be))
    ev = ak.from_offsets(off2, fields, tuned)
    EventPipeline(ev, [("pt", ">", 25.0), ("eta", "abs<", 2.4)], slice(1, None), "pt",
                  outputs=("num", "sum", "argmax")).run()
    torch.cuda.synchronize()
    del ev
    off3, f3 = make_events(20_000_000, seed=3, lam=4, clip=30)
    f3 = {k: f3[k] for k in ("pt", "eta", "phi")}
    ev3 = ak.from_offsets(off3, f3, tuned)
    pairs_step(ak, np, ev3)
    pairs_plan(ev3).run()
    torch.cuda.synchronize()
    del ev3
    batch = make_batch(max_exp=4.5)
    dev = [(torch.from_numpy(o).cuda(), torch.from_numpy(v).cuda()) for o, v in batch]
    BatchRunner(dev, variants={n: 0 for n in ("int32", "int64", "float32", "float64")}).run()
    torch.cuda.synchronize()
    del dev
    # option types and non-local reductions on the 1M-list layout: 10% of the
    # values missing (IndexedOptionArray), and the same lists reduced across
    # (axis=0: by position)
    from awkward_amd import contents as C

    index = np.arange(len(x), dtype=np.int64)
    index[rng.random(len(x)) < 0.1] = -1
    opt = ak.Array(C.ListOffsetArray(tuned.nplike.asarray(off), C.IndexedOptionArray(
        tuned.nplike.asarray(index), C.NumpyArray(tuned.nplike.asarray(x), tuned), tuned), tuned))
    for fn in (ak.sum, ak.argmax):
        fn(opt, axis=-1)
    arr = ak.from_offsets(off, x, tuned)
    for fn in (ak.sum, ak.argmax):
        fn(arr, axis=0)
    torch.cuda.synchronize()
    print("roofline probe done", flush=True)


def cpu_baseline(ak, off, fields, sample_events):
    import numpy as np

    from oracle.backend import OracleBackend

    be = OracleBackend()
    n = min(sample_events, len(off) - 1)
    end = int(off[n])
    sub_off = off[: n + 1].copy()
    sub_fields = {k: np.ascontiguousarray(v[:end]) for k, v in fields.items()}
    ev = ak.from_offsets(sub_off, sub_fields, be)
    pipeline(ak, ev)  # warm-up
    t0 = time.perf_counter()
    reps = 0
    while True:
        pipeline(ak, ev)
        reps += 1
        if time.perf_counter() - t0 > 10.0 or reps >= 5:
            break
    dt = (time.perf_counter() - t0) / reps
    return {"value": round(n / dt, 1), "unit": "events/s", "cores": 1, "kind": "port",
            "sample": f"first {n} events of the same workload, {reps} passes of the full pipeline "
                      f"through the same host layer on the single-threaded C++ oracle"}


ACCURACY = {}  # host copies of inputs and GPU sums for the fsum report


def _fsum_errors(off, values, sums, exact):
    """Max and mean |sum - fsum| / sum(|x|) over the lists (0 for empty lists)."""
    import numpy as np

    vals = np.concatenate([np.abs(values.astype(np.float64)), [0.0]])  # reduceat needs start < len
    mags = np.add.reduceat(vals, off[:-1])
    mags[off[1:] == off[:-1]] = 0.0
    err = np.abs(sums.astype(np.float64) - exact)
    rel = np.where(mags > 0, err / np.where(mags > 0, mags, 1.0), 0.0)
    return {"max_rel_err": float(rel.max()) if len(rel) else 0.0, "mean_rel_err": float(rel.mean()) if len(rel) else 0.0}


def accuracy_report(ak, off, fields, gpu_sum, sample_events):
    """Float sums against math.fsum, for the GPU and the single-threaded CPU
    oracle: the 1M-list float64 sums and the event pipeline's
    float32 sum(pt) over the kept particles (the first `sample_events`
    events).  Error = |sum - fsum| / sum(|x|) per list."""
    import math

    import numpy as np

    from oracle.backend import OracleBackend

    be = OracleBackend()
    report = {"definition": "per list |sum - math.fsum| / sum(|x|); max and mean over lists"}
    if "segmented_reductions" in ACCURACY:
        off1, x1, g1 = ACCURACY["segmented_reductions"]
        exact = np.array([math.fsum(x1[a:b]) for a, b in zip(off1[:-1], off1[1:])])
        c1 = ak.sum(ak.from_offsets(off1, x1, be), axis=-1).layout
        cpu1 = np.asarray(c1.nplike.to_numpy(c1.data))
        report["segmented_sum_f64"] = {"lists": len(exact), "gpu": _fsum_errors(off1, x1, g1, exact),
                                     "cpu": _fsum_errors(off1, x1, cpu1, exact),
                                     "gpu_bits_equal_cpu": float(np.mean(g1.view(np.int64) == cpu1.view(np.int64)))}
    n = min(sample_events, len(off) - 1)
    pt = fields["pt"][: int(off[n])]
    eta = fields["eta"][: int(off[n])]
    keep = (pt > np.float32(25)) & (np.abs(eta) < np.float32(2.4))
    rank = np.cumsum(keep) - 1  # selected particles before each one, over all events
    sel_before = np.concatenate([[0], np.cumsum(keep)])[off[:n]]
    event = np.repeat(np.arange(n), np.diff(off[: n + 1]))
    kept = keep & ((rank - sel_before[event]) >= 1)  # [:, 1:] of the selected particles
    kept_off = np.concatenate([[0], np.cumsum(np.bincount(event[kept], minlength=n))])
    kept_pt = pt[kept]
    # sums of at most 20 float32 values of 5..~500 are exact in float64
    exact2 = np.zeros(n)
    np.add.at(exact2, event[kept], kept_pt.astype(np.float64))
    sub = ak.from_offsets(off[: n + 1].copy(), {k: np.ascontiguousarray(v[: int(off[n])]) for k, v in fields.items()}, be)
    c2 = pipeline(ak, sub)[1].layout
    cpu2 = np.asarray(c2.nplike.to_numpy(c2.data))
    g2 = gpu_sum[:n]
    report["pipeline_sum_pt_f32"] = {"events": n, "gpu": _fsum_errors(kept_off, kept_pt, g2, exact2),
                                    "cpu": _fsum_errors(kept_off, kept_pt, cpu2, exact2),
                                    "gpu_bits_equal_cpu": float(np.mean(g2.view(np.int32) == cpu2.view(np.int32)))}
    return report


def time_form(torch, fn, steps, warmup, events):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return {"ms_per_step": round(dt * 1e3, 4), "events_per_s": round(events / dt, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--events", type=int, default=50_000_000)
    ap.add_argument("--cpu-sample", type=int, default=2_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-reductions", action="store_true", help="skip the 1M-list segmented reductions")
    ap.add_argument("--no-combinations", action="store_true", help="skip the 20M-event combinations")
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-array batch")
    ap.add_argument("--no-dispatch", action="store_true", help="skip the dispatcher regret sweep")
    ap.add_argument("--traffic-probe", action="store_true", help="only the program profiled by the counter runs")
    ap.add_argument("--roofline-probe", action="store_true", help="every workload's kernels once, for counter runs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: there is no multi-GPU path (each op is independent per list; a "
              "sharded form would give each GPU whole lists and exchange only the k per-GPU totals that "
              "offsets need)", file=sys.stderr)
        sys.exit(2)

    if args.traffic_probe:
        traffic_probe(args)
        return
    if args.roofline_probe:
        roofline_probe(args)
        return

    import torch

    import awkward_amd as ak
    from awkward_amd.backend import HipBackend
    from awkward_amd.fused import EventPipeline, Graph

    log(f"generating {args.events} events")
    off, fields = make_events(args.events)
    plain_be = HipBackend(fused=False)
    tuned_be = HipBackend(fused=True)
    ev_plain = ak.from_offsets(off, fields, plain_be)
    ev = ak.from_offsets(off, fields, tuned_be)
    torch.cuda.synchronize()
    N = int(off[-1])
    log(f"resident: {args.events} events, {N} particles")

    # the three forms of the same pipeline
    forms = {}
    short = max(3, args.steps // 4)
    forms["plain"] = time_form(torch, lambda: pipeline(ak, ev_plain), short, 1, args.events)
    forms["plain"]["what"] = "the reference's kernel sequence, one registry kernel per step, host syncs for sizes"
    log(f"plain: {forms['plain']}")
    forms["tuned"] = time_form(torch, lambda: pipeline(ak, ev), short, 1, args.events)
    forms["tuned"]["what"] = "per-kernel path with fused/tuned reduction kernels (variant by mean list length)"
    log(f"tuned: {forms['tuned']}")

    terms = [("pt", ">", 25.0), ("eta", "abs<", 2.4)]
    pipe = EventPipeline(ev, terms, slice(1, None), "pt", outputs=("num", "sum", "argmax"))
    graph = Graph(pipe.run)
    for _ in range(args.warmup):
        graph.replay()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    stops = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record()
        graph.replay()
        stops[k].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = args.events / ((t1 - t0) / args.steps)
    kms = sum(a.elapsed_time(b) for a, b in zip(starts, stops)) / args.steps
    forms["fused_graph"] = {"ms_per_step": round(ms_per_step, 4), "events_per_s": round(value, 1),
                            "what": "awkward_hip_event_pipeline, one kernel, replayed from a hipGraph"}
    log(f"fused_graph: {forms['fused_graph']}")

    gpu_sum_pt = pipe.out["sum"][: args.cpu_sample].cpu().numpy()
    alg_bytes = pipe.algorithmic_bytes(N)
    achieved = alg_bytes / (kms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(args.events),
                "kernel": "event_pipeline_kernel (awkward_hip_event_pipeline)", "kernel_ms": round(kms, 4),
                "alg_bytes": alg_bytes}

    reductions = None if args.no_reductions else run_segmented_reductions(ak, tuned_be, torch)
    del ev, ev_plain, pipe, graph
    torch.cuda.empty_cache()
    combos = None if args.no_combinations else run_combinations(ak, tuned_be, torch)
    if combos:
        log(f"combinations: plain {combos['plain']['events_per_s']:.4g} ev/s, fused {combos['fused_graph']['events_per_s']:.4g} ev/s")
    torch.cuda.empty_cache()
    batched = None if args.no_batched else run_batched(ak, torch)
    if batched:
        log(f"batched: plain {batched['plain']['lists_per_s']:.4g} lists/s, batched {batched['batched_graph']['lists_per_s']:.4g} lists/s")
    regret = None if args.no_dispatch else run_dispatch_regret(torch)
    if regret:
        log(f"dispatch: learned regret {regret['learned']['aggregate']}, table {regret['table']['aggregate']}, "
            f"heuristic {regret['heuristic']['aggregate']} over {regret['rows']} held-out rows")
    cpu = None if args.no_cpu else cpu_baseline(ak, off, fields, args.cpu_sample)
    accuracy = None if args.no_cpu else accuracy_report(ak, off, fields, gpu_sum_pt, args.cpu_sample)
    if accuracy:
        log(f"accuracy: {json.dumps(accuracy)}")

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator of the event-pipeline workload)",
        "config": {"workload": "jagged event pipeline: 50M events {pt,eta,phi f32; charge i8}, "
                               "Poisson(3) clipped [0,20]; mask -> jagged getitem -> [:,1:] -> num/sum/argmax",
                   "events": args.events, "particles": N, "path": "fused_graph"},
        "forms": forms,
        "speedup_fused_vs_plain": round(forms["fused_graph"]["events_per_s"] / forms["plain"]["events_per_s"], 2),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "accuracy": accuracy,
        "segmented_reductions": reductions,
        "combinations": combos,
        "batched": batched,
        "dispatch_regret": regret,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
