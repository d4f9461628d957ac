"""Benchmark of the hip jagged-array path on one MI355X.

Workload (one "step"): the jagged event pipeline on 50M events of records
{pt, eta, phi: float32, charge: int8}, multiplicity ~ Poisson(3) clipped to
[0, 20], seed 2:

    mask = (ev.pt > 25) & (abs(ev.eta) < 2.4)
    sel  = ev[mask][:, 1:]
    ak.num(sel), ak.sum(sel.pt, axis=-1), ak.argmax(sel.pt, axis=-1)

Inputs are resident in HBM before timing.  `value` is events/s of the whole
pipeline.  `roofline` is the HBM roofline of the pipeline's dominant kernel
(algorithmic bytes / its average duration, from HIP events recorded on the
kernels' stream inside the timed steps).  `cpu_baseline` times the same
pipeline through the same host layer on the single-threaded CPU oracle over a
bounded sample of the events.  `config1` reports the segmented reductions of
the 1M-list Poisson(10) float64 layout, HBM-cold (a 1 GiB scrub before each
timed launch).

    python bench.py [--steps K] [--warmup W] [--events N] [--no-cpu]
"""

from __future__ import annotations

import argparse
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


class KernelTimer:
    """Wraps registry kernels of a backend to record HIP events around each
    launch on the current stream (the stream the library enqueues on)."""

    def __init__(self, backend, names):
        import torch

        self.torch = torch
        self.backend = backend
        self.names = set(names)
        self.events = {}
        self.enabled = False
        self._orig = backend.__class__.__getitem__

        timer = self

        def getitem(be, key):
            kernel = timer._orig(be, key)
            fam = key[0]
            if fam in timer.names and timer.enabled:
                return _Timed(kernel, fam, timer)
            return kernel

        backend.__class__.__getitem__ = getitem

    def restore(self):
        self.backend.__class__.__getitem__ = self._orig

    def record(self, fam):
        torch = self.torch
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        self.events.setdefault(fam, []).append((start, stop))
        return start, stop

    def mean_ms(self, fam):
        torch = self.torch
        torch.cuda.synchronize()
        ev = self.events.get(fam, [])
        if not ev:
            return None
        return sum(a.elapsed_time(b) for a, b in ev) / len(ev)


class _Timed:
    def __init__(self, kernel, fam, timer):
        self.kernel, self.fam, self.timer = kernel, fam, timer

    def __call__(self, *args):
        start, stop = self.timer.record(self.fam)
        start.record()
        out = self.kernel(*args)
        stop.record()
        return out


def run_config1(ak, backend, torch, reps=5):
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
    result = {"lists": n, "elements": N, "dtype": "float64", "method": "HBM-cold: 1 GiB scrub before each launch"}
    for name, (key, args, nbytes) in ops.items():
        kernel = backend[key]
        kernel(*args)
        times = []
        for _ in range(reps):
            scrub.add_(1.0)
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record()
            kernel(*args)
            e.record()
            torch.cuda.synchronize()
            times.append(s.elapsed_time(e))
        ms = sorted(times)[len(times) // 2]
        gbs = nbytes / (ms * 1e-3) / 1e9
        result[name] = {"ms": round(ms, 4), "GB/s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 3),
                        "bytes": nbytes}
    return result


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--events", type=int, default=50_000_000)
    ap.add_argument("--cpu-sample", type=int, default=2_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-config1", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each op is independent per "
              "list; sharding by list ranges is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch

    import awkward_amd as ak
    from awkward_amd.backend import hip_backend

    log(f"generating {args.events} events")
    off, fields = make_events(args.events)
    backend = hip_backend()
    ev = ak.from_offsets(off, fields, backend)
    torch.cuda.synchronize()
    N = int(off[-1])
    log(f"resident: {args.events} events, {N} particles")

    for _ in range(args.warmup):
        pipeline(ak, ev)
    torch.cuda.synchronize()

    dominant = "awkward_reduce_sum"
    timer = KernelTimer(backend, [dominant])
    timer.enabled = True
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        outs = pipeline(ak, ev)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    timer.enabled = False
    timer.restore()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = args.events / ((t1 - t0) / args.steps)
    log(f"{ms_per_step:.3f} ms/step, {value:.4g} events/s")

    # dominant kernel roofline: segmented sum over the sliced pt (float32)
    sel_n = args.events
    K = int(ak.num(ev[(ev.pt > 25) & (abs(ev.eta) < 2.4)][:, 1:]).layout.data.sum().item())
    alg_bytes = 8 * (sel_n + 1) + 4 * K + 4 * sel_n
    kms = timer.mean_ms(dominant)
    achieved = alg_bytes / (kms * 1e-3) / 1e9 if kms else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None, "traffic": None,
                "kernel": "segmented_reduce_kernel (awkward_reduce_sum_float32_float32_64)",
                "kernel_ms": round(kms, 4) if kms else None, "alg_bytes": alg_bytes}

    config1 = None if args.no_config1 else run_config1(ak, backend, torch)
    cpu = None if args.no_cpu else cpu_baseline(ak, off, fields, args.cpu_sample)

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
                   "events": args.events, "particles": N, "path": "plain (per-kernel launches)"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "config1": config1,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
