"""The learned tile dispatcher: an FP32 MLP cost model over 16 features of a
jagged array, trained and evaluated by the library's own HIP kernels
(``csrc/learned.hip``).

Which of the ten reduce tiles (``dispatch.VARIANTS``) is fastest depends on
the list-length distribution, the element size and the op.  The model
predicts each tile's log2 cost relative to the mean over tiles (plus the mean
itself) and the dispatcher takes the cheapest:

* features come from one device pass over the offsets (integer
  accumulators, so they are identical from run to run) and the network runs
  in the same launch, in the workgroup that finishes the statistics last
  (``awkward_hip_dispatch_choose``): the batch runner enqueues that launch
  and ten predicated tile launches without a host round trip, and a
  single-array call reads back one int;
* a feature outside the range seen in training selects the autotuner table
  (bucketed exhaustive timings), else the fixed heuristic;
* training data: the exhaustive sweep (``awkward_hip_dispatch_time_variants``)
  on synthetic arrays drawn on the device from the workloads' generators
  (Poisson event multiplicities, their pair counts, Poisson(50), Zipf(1.5)
  capped, and mixtures), seeds 1000-1999; evaluation on seeds 9000+.

Ops (feature 14): 0 = sum, 1 = min/max with mask, 2 = local argmin/argmax
with mask, 3 = the batch kernel's fused sum+min+argmin.

Command line (on a machine with the GPU)::

    python -m awkward_amd.learned collect --seeds 1000:2000 --per-seed 20 --out train.npz
    python -m awkward_amd.learned collect --seeds 9000:9100 --per-seed 20 --out heldout.npz
    python -m awkward_amd.learned train --data train.npz --out awkward_amd/data/dispatch_mlp.json
    python -m awkward_amd.learned evaluate --model awkward_amd/data/dispatch_mlp.json --data heldout.npz
"""

from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np
import torch

from awkward_amd._kernels import KernelError, set_stream
from awkward_amd._lib import lib

DEFAULT_MODEL = os.path.join(os.path.dirname(__file__), "data", "dispatch_mlp.json")
OP_CODE = {"sum": 0, "min": 1, "max": 1, "argmin": 2, "argmax": 2, "batch": 3}
DTYPE_CODE = {torch.int32: 0, torch.int64: 1, torch.float32: 2, torch.float64: 3}
N_VARIANTS = 10
OUTPUT_WEIGHTS = [1.0] * N_VARIANTS + [0.1]

P = ctypes.c_void_p
I32, I64, F32 = ctypes.c_int32, ctypes.c_int64, ctypes.c_float
_fns = {}


def _fn(name, argtypes, restype=KernelError):
    f = _fns.get(name)
    if f is None:
        f = getattr(lib(), name)
        f.argtypes = argtypes
        f.restype = restype
        _fns[name] = f
    return f


def _check(err):
    if err.str is not None:
        raise RuntimeError(err.str.decode())


def layout():
    f = _fn("awkward_hip_dispatch_layout", [I32], I32)
    keys = ["features", "hidden", "outputs", "params", "total", "acc_words", "stats_lists", "rows"]
    return {k: int(f(i)) for i, k in enumerate(keys)}


# ---------------------------------------------------------------------------
# device statistics + inference
# ---------------------------------------------------------------------------


class StatsPlan:
    """Device tables for one statistics pass: entries are (offsets tensor,
    number of lists, slot); every entry of a slot accumulates into it."""

    def __init__(self, entries, nslots, device):
        lay = layout()
        L = lay["stats_lists"]
        rows, tile = [], 0
        for offsets, n, slot in entries:
            if n <= 0:
                continue
            rows.append([offsets.data_ptr(), n, tile, slot])
            tile += -(-n // L)
        self.total_tiles = tile
        self.nitems = len(rows)
        self.nslots = nslots
        self.table = (torch.from_numpy(np.array(rows, dtype=np.uint64).view(np.int64)).to(device)
                      if rows else torch.zeros(4, dtype=torch.int64, device=device))
        self.acc = torch.zeros(nslots * lay["acc_words"], dtype=torch.int64, device=device)
        self._acc_clear = True  # the one-launch form leaves the accumulators at zero

    def run(self):
        """The statistics alone: the accumulators stay in ``acc``."""
        set_stream()
        _check(_fn("awkward_hip_dispatch_reset", [P, I64])(self.acc.data_ptr(), self.nslots))
        _check(_fn("awkward_hip_dispatch_stats", [P, I32, I64, P])(self.table.data_ptr(), self.nitems,
                                                                   self.total_tiles, self.acc.data_ptr()))
        self._acc_clear = False

    def choose(self, meta, params, features=None, costs=None, choice=None):
        """Statistics and inference in one launch (``awkward_hip_dispatch_choose``)."""
        set_stream()
        if not self._acc_clear:
            _check(_fn("awkward_hip_dispatch_reset", [P, I64])(self.acc.data_ptr(), self.nslots))
        _check(_fn("awkward_hip_dispatch_choose", [P, I32, I64, P, P, I32, P, P, P, P])(
            self.table.data_ptr(), self.nitems, self.total_tiles, self.acc.data_ptr(), meta.data_ptr(), self.nslots,
            params.data_ptr(), 0 if features is None else features.data_ptr(),
            0 if costs is None else costs.data_ptr(), choice.data_ptr()))
        self._acc_clear = True


def infer(acc, meta, nslots, params, features=None, costs=None, choice=None):
    set_stream()
    _check(_fn("awkward_hip_dispatch_infer", [P, P, I32, P, P, P, P])(
        acc.data_ptr(), meta.data_ptr(), nslots, params.data_ptr(),
        0 if features is None else features.data_ptr(), 0 if costs is None else costs.data_ptr(),
        choice.data_ptr()))


def device_features(offsets, n, itemsize, op, device=None):
    """The 16 features of one array as the device computes them (syncs)."""
    device = device or offsets.device
    lay = layout()
    plan = StatsPlan([(offsets, n, 0)], 1, device)
    plan.run()
    meta = torch.tensor([itemsize, op, 0, 0], dtype=torch.int32, device=device)
    params = torch.zeros(lay["total"], dtype=torch.float32, device=device)
    feats = torch.empty(lay["features"], dtype=torch.float32, device=device)
    choice = torch.empty(2, dtype=torch.int32, device=device)
    infer(plan.acc, meta, 1, params, features=feats, choice=choice)
    return feats.cpu().numpy()


# ---------------------------------------------------------------------------
# the model
# ---------------------------------------------------------------------------


class Model:
    """Parameters (flat float32, the device layout) + normalisation + the
    autotuner table used as fallback."""

    def __init__(self, params: np.ndarray, table: dict, info: dict | None = None):
        self.params = np.asarray(params, dtype=np.float32)
        self.table = table
        self.info = info or {}
        self._dev = {}

    @classmethod
    def load(cls, path=DEFAULT_MODEL):
        with open(path) as f:
            d = json.load(f)
        return cls(np.array(d["params"], dtype=np.float32), d.get("table", {}), d.get("info", {}))

    def save(self, path):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        with open(path, "w") as f:
            json.dump({"params": [float(x) for x in self.params], "table": self.table, "info": self.info}, f)

    def device_params(self, device):
        key = str(device)
        t = self._dev.get(key)
        if t is None:
            t = torch.from_numpy(self.params.copy()).to(device)
            self._dev[key] = t
        return t

    def table_choice(self, op: int, n: int, N: int, itemsize: int):
        from awkward_amd import dispatch

        v = self.table.get(bucket_key(op, n, N, itemsize))
        if v is None:
            return dispatch.heuristic_variant(n, N, itemsize)
        return int(v)


def bucket_key(op, n, N, itemsize):
    n, N = max(int(n), 0), max(int(N), 0)
    mean = N / max(n, 1)
    return f"{op}|{int(round(math.log2(1 + n)))}|{int(round(2 * math.log2(1 + mean)))}|{itemsize}"


_default = None


def default_model():
    global _default
    if _default is None:
        if not os.path.exists(DEFAULT_MODEL):
            raise FileNotFoundError(f"no trained dispatcher at {DEFAULT_MODEL}")
        _default = Model.load(DEFAULT_MODEL)
    return _default


class Dispatcher:
    """choose(op, offsets, n, N, itemsize) -> variant (one small sync, cached
    on the offsets tensor); choose_device(...) -> device int32[2] choice."""

    def __init__(self, model: Model | None = None):
        self.model = model or default_model()
        self.last_fallback = None

    def choose_device(self, op, entries, slot_sizes, itemsizes, device):
        """entries: (offsets, n, slot); slot_sizes[slot] = (lists, elements)
        and itemsizes[slot], known on the host from the shapes (they only
        pick the table's fallback).  Returns (plan, meta, choice) for
        run_device -- everything stays on the device."""
        nslots = len(slot_sizes)
        plan = StatsPlan(entries, nslots, device)
        meta = torch.tensor(
            [[itemsizes[s], op, self.model.table_choice(op, n, N, itemsizes[s]), 0]
             for s, (n, N) in enumerate(slot_sizes)], dtype=torch.int32).reshape(-1).to(device)
        choice = torch.zeros(2 * nslots, dtype=torch.int32, device=device)
        return plan, meta, choice

    def run_device(self, plan, meta, choice, costs=None):
        plan.choose(meta, self.model.device_params(plan.acc.device), costs=costs, choice=choice)

    def choose(self, op: str | int, offsets, n: int, N: int, itemsize: int) -> int:
        code = OP_CODE[op] if isinstance(op, str) else int(op)
        cache = getattr(offsets, "_akhip_choice", None)
        if cache is not None and code in cache:
            return cache[code]
        device = offsets.device
        plan = StatsPlan([(offsets, n, 0)], 1, device)
        meta = torch.tensor([itemsize, code, self.model.table_choice(code, n, N, itemsize), 0], dtype=torch.int32,
                            device=device)
        choice = torch.zeros(2, dtype=torch.int32, device=device)
        self.run_device(plan, meta, choice)
        c = choice.cpu().tolist()
        self.last_fallback = bool(c[1])
        try:
            if cache is None:
                cache = {}
                offsets._akhip_choice = cache
            cache[code] = int(c[0])
        except AttributeError:
            pass
        return int(c[0])

    def chooser(self, op, n_lists, n_elements, itemsize, offsets=None):
        """For ``dispatch.chooser``: the learned choice when the offsets are at
        hand, else the table."""
        if offsets is None:
            return self.model.table_choice(OP_CODE[op], n_lists, n_elements, itemsize)
        return self.choose(op, offsets, n_lists, n_elements, itemsize)


def enable(model: Model | None = None):
    """Route the tuned reduce path's tile choice through the dispatcher."""
    from awkward_amd import dispatch

    d = Dispatcher(model)
    dispatch.chooser = d.chooser
    return d


def disable():
    from awkward_amd import dispatch

    dispatch.chooser = None


# ---------------------------------------------------------------------------
# synthetic arrays on the device
# ---------------------------------------------------------------------------

KINDS = ("poisson", "events", "pairs", "poisson50", "zipf", "mixture")


def _counts(kind, n, rng, gen, device):
    def poisson(lam, m):
        return torch.poisson(torch.full((m,), float(lam), device=device), generator=gen).to(torch.int64)

    def zipf(s, cap, m):
        u = torch.rand(m, device=device, dtype=torch.float64, generator=gen).clamp_min(1e-300)
        return torch.clamp(torch.floor(u ** (-1.0 / (s - 1.0))), max=float(cap)).to(torch.int64)

    if kind == "poisson":
        return poisson(10 ** rng.uniform(-0.7, 2.5), n)
    if kind == "events":
        return torch.clamp(poisson(rng.uniform(1.5, 6.0), n), max=int(rng.choice([20, 30])))
    if kind == "pairs":
        m = torch.clamp(poisson(rng.uniform(2.0, 6.0), n), max=30)
        k = int(rng.choice([2, 3]))
        return m * (m - 1) // 2 if k == 2 else m * (m - 1) * (m - 2) // 6
    if kind == "poisson50":
        return poisson(50, n)
    if kind == "zipf":
        return zipf(rng.uniform(1.3, 2.5), int(10 ** rng.uniform(2, 5)), n)
    # mixture of two or three of the above, list by list
    parts = [_counts(k, n, rng, gen, device) for k in rng.choice(["poisson", "events", "poisson50", "zipf"],
                                                                 size=int(rng.integers(2, 4)), replace=False)]
    pick = torch.randint(0, len(parts), (n,), device=device, generator=gen)
    out = parts[0].clone()
    for i in range(1, len(parts)):
        out = torch.where(pick == i, parts[i], out)
    return out


def synthetic_array(rng, gen, device, max_elements=120_000_000, min_exp=2.5, max_exp=7.0):
    """One array: (offsets int64, data, spec dict).  The spec is drawn from
    `rng` (numpy), the lengths and values on the device from `gen`."""
    kind = KINDS[int(rng.integers(0, len(KINDS)))]
    n = int(10 ** rng.uniform(min_exp, max_exp))
    dtype = [torch.int32, torch.int64, torch.float32, torch.float64][int(rng.integers(0, 4))]
    empty = rng.random() < 0.03
    for _ in range(8):
        counts = torch.zeros(n, dtype=torch.int64, device=device) if empty else _counts(kind, n, rng, gen, device)
        N = int(counts.sum())
        if N <= max_elements:
            break
        n = max(1, n // 4)
    offsets = torch.zeros(n + 1, dtype=torch.int64, device=device)
    torch.cumsum(counts, 0, out=offsets[1:])
    if dtype in (torch.int32, torch.int64):
        data = torch.randint(-1000, 1000, (max(N, 1),), device=device, generator=gen, dtype=torch.int64).to(dtype)
    else:
        data = torch.randn(max(N, 1), device=device, generator=gen, dtype=torch.float64).to(dtype)
    return offsets, data, {"kind": kind, "n": n, "N": N, "dtype": str(dtype).split(".")[-1], "empty": bool(empty)}


def time_variants(op, offsets, data, n, reps):
    """ms per launch of each of the ten tiles (the exhaustive sweep)."""
    out = (ctypes.c_float * N_VARIANTS)()
    set_stream()
    _check(_fn("awkward_hip_dispatch_time_variants", [I32, I32, P, P, I64, I64, I32, P])(
        op, DTYPE_CODE[data.dtype], data.data_ptr(), offsets.data_ptr(), n, data.shape[0], reps,
        ctypes.addressof(out)))
    return np.array(out[:], dtype=np.float64)


def collect(seeds, per_seed, device, ops=(0, 1, 2, 3), log=None, max_elements=120_000_000, budget_s=None):
    """Rows of (features, ms per tile) for every op on synthetic arrays;
    stops after the first seed that ends past `budget_s` seconds."""
    feats, times, specs = [], [], []
    t0 = time.time()
    for si, seed in enumerate(seeds):
        if budget_s is not None and time.time() - t0 > budget_s:
            if log:
                log(f"collect: time budget reached after {si} seeds")
            break
        rng = np.random.default_rng(seed)
        gen = torch.Generator(device=device)
        gen.manual_seed(seed)
        for _ in range(per_seed):
            offsets, data, spec = synthetic_array(rng, gen, device, max_elements=max_elements)
            n, N = spec["n"], spec["N"]
            base = device_features(offsets, n, data.element_size(), 0)
            reps = int(min(8, max(2, 2e7 // max(N, 1))))
            for op in ops:
                f = base.copy()
                f[14] = np.float32(op / 3.0)
                feats.append(f)
                times.append(time_variants(op, offsets, data, n, reps))
                specs.append(dict(spec, op=op))
            del offsets, data
        if log and (si % 10 == 9 or si == len(seeds) - 1):
            log(f"collect: {si + 1}/{len(seeds)} seeds, {len(feats)} rows, {time.time() - t0:.1f} s")
    return np.array(feats, dtype=np.float32), np.array(times, dtype=np.float64), specs


# ---------------------------------------------------------------------------
# training (HIP kernels)
# ---------------------------------------------------------------------------


def targets(times_ms):
    lg = np.log2(np.maximum(times_ms, 1e-6) * 1000.0)
    mean = lg.mean(axis=1, keepdims=True)
    return np.concatenate([lg - mean, mean], axis=1).astype(np.float32)


def normalisation(X):
    mu = X.mean(axis=0).astype(np.float32)
    sigma = X.std(axis=0).astype(np.float32)
    inv = np.where(sigma > 1e-6, 1.0 / np.maximum(sigma, 1e-6), 1.0).astype(np.float32)
    Z = (X - mu) * inv
    span = Z.max(axis=0) - Z.min(axis=0)
    lo = (Z.min(axis=0) - 0.1 * span - 0.25).astype(np.float32)
    hi = (Z.max(axis=0) + 0.1 * span + 0.25).astype(np.float32)
    return mu, inv, lo, hi


def init_params(rng, lay):
    F, H, O = lay["features"], lay["hidden"], lay["outputs"]
    parts = [
        rng.normal(0, math.sqrt(2.0 / F), (F, H)), np.zeros(H),
        rng.normal(0, math.sqrt(2.0 / H), (H, H)), np.zeros(H),
        rng.normal(0, math.sqrt(1.0 / H), (H, O)), np.zeros(O),
    ]
    p = np.concatenate([x.reshape(-1) for x in parts]).astype(np.float32)
    assert p.size == lay["params"]
    return p


class Trainer:
    """Adam on the device over normalised rows; deterministic (fixed
    permutations, fixed-order gradient sums)."""

    def __init__(self, Xn, T, params, device, ow=OUTPUT_WEIGHTS):
        self.lay = layout()
        self.device = device
        self.R = Xn.shape[0]
        self.X = torch.from_numpy(np.ascontiguousarray(Xn, dtype=np.float32)).to(device)
        self.T = torch.from_numpy(np.ascontiguousarray(T, dtype=np.float32)).to(device)
        self.ow = torch.tensor(ow, dtype=torch.float32, device=device)
        self.P = torch.from_numpy(np.ascontiguousarray(params, dtype=np.float32)).to(device)
        self.m = torch.zeros_like(self.P)
        self.v = torch.zeros_like(self.P)
        self.step_no = 0
        self.loss_acc = torch.zeros(1, dtype=torch.float32, device=device)
        self._grads = None

    def _buffers(self, bsz):
        parts = -(-bsz // self.lay["rows"])
        if self._grads is None or self._grads.shape[0] < parts:
            self._grads = torch.empty((parts, self.lay["params"]), dtype=torch.float32, device=self.device)
            self._losses = torch.empty(parts, dtype=torch.float32, device=self.device)
        return parts

    def step(self, pa, pb, row0, bsz, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        rows = min(bsz, self.R - row0)
        parts = self._buffers(rows)
        set_stream()
        _check(_fn("awkward_hip_mlp_train_step", [P, P, P, I64, I64, I64, I64, I64, P, P, P])(
            self.X.data_ptr(), self.T.data_ptr(), self.ow.data_ptr(), self.R, pa, pb, row0, bsz, self.P.data_ptr(),
            self._grads.data_ptr(), self._losses.data_ptr()))
        self.step_no += 1
        _check(_fn("awkward_hip_mlp_adam", [P, P, P, P, P, I32, F32, F32, F32, F32, I64, P])(
            self.P.data_ptr(), self.m.data_ptr(), self.v.data_ptr(), self._grads.data_ptr(),
            self._losses.data_ptr(), -(-rows // self.lay["rows"]), lr, beta1, beta2, eps, self.step_no,
            self.loss_acc.data_ptr()))

    def epoch(self, rng, bsz, lr):
        R = self.R
        while True:
            pa = int(rng.integers(1, R)) if R > 1 else 0
            if math.gcd(pa, R) == 1:
                break
        pb = int(rng.integers(0, R))
        self.loss_acc.zero_()
        steps = 0
        for row0 in range(0, R, bsz):
            self.step(pa, pb, row0, bsz, lr)
            steps += 1
        return float(self.loss_acc.item()) / steps

    def params(self):
        return self.P.cpu().numpy()


def forward(X, params_dev):
    """Outputs (rows x 11) and the inside-range flags for raw feature rows."""
    lay = layout()
    device = params_dev.device
    Xd = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float32)).to(device)
    Y = torch.empty((X.shape[0], lay["outputs"]), dtype=torch.float32, device=device)
    inside = torch.empty(X.shape[0], dtype=torch.int32, device=device)
    set_stream()
    _check(_fn("awkward_hip_mlp_forward", [P, I64, P, P, P])(Xd.data_ptr(), X.shape[0], params_dev.data_ptr(),
                                                              Y.data_ptr(), inside.data_ptr()))
    return Y.cpu().numpy(), inside.cpu().numpy().astype(bool)


def build_table(X, times_ms):
    """The autotuner table: per bucket (op, log2 lists, 2 log2 mean length,
    itemsize), the tile with the least total regret over the bucket's rows."""
    rows = {}
    for f, t in zip(X, times_ms):
        n = int(round(2 ** f[0] - 1))
        N = int(round(2 ** f[1] - 1))
        op = int(round(f[14] * 3))
        itemsize = int(round(f[13] * 8))
        rows.setdefault(bucket_key(op, n, N, itemsize), []).append(t / t.min())
    return {k: int(np.argmin(np.sum(v, axis=0))) for k, v in rows.items()}


def train(X, times_ms, device, epochs=80, bsz=512, lr=3e-3, seed=0, log=None, val_frac=0.1):
    lay = layout()
    rng = np.random.default_rng(seed)
    T = targets(times_ms)
    order = rng.permutation(X.shape[0])
    nval = int(len(order) * val_frac)
    val, tr = order[:nval], order[nval:]
    mu, inv, lo, hi = normalisation(X[tr])
    Xn = ((X - mu) * inv).astype(np.float32)
    p0 = init_params(rng, lay)
    trainer = Trainer(Xn[tr], T[tr], p0, device)
    history = []
    for e in range(epochs):
        lr_e = 1e-4 + 0.5 * (lr - 1e-4) * (1 + math.cos(math.pi * e / epochs))
        loss = trainer.epoch(rng, bsz, lr_e)
        history.append(loss)
        if log and (e % 10 == 9 or e == epochs - 1):
            log(f"train: epoch {e + 1}/{epochs} loss {loss:.4f}")
    params = np.concatenate([trainer.params(), mu, inv, lo, hi]).astype(np.float32)
    assert params.size == lay["total"]
    model = Model(params, build_table(X[tr], times_ms[tr]),
                  {"rows": int(len(tr)), "val_rows": int(nval), "epochs": epochs, "batch": bsz, "lr": lr,
                   "loss": history[-1]})
    if nval:
        r = regret(model, X[val], times_ms[val], device)
        model.info["val_regret"] = r["learned"]["aggregate"]
    return model


# ---------------------------------------------------------------------------
# evaluation
# ---------------------------------------------------------------------------


def choices(model, X, device):
    """Learned choice per row (the device network), the fallback applied
    where a feature is outside the training range."""
    Y, inside = forward(X, model.device_params(device))
    pick = np.argmin(Y[:, :N_VARIANTS], axis=1)
    out = pick.copy()
    for i in np.nonzero(~inside)[0]:
        f = X[i]
        out[i] = model.table_choice(int(round(f[14] * 3)), int(round(2 ** f[0] - 1)), int(round(2 ** f[1] - 1)),
                                    int(round(f[13] * 8)))
    return out, inside


def regret(model, X, times_ms, device):
    """Aggregate regret (sum of the chosen tiles' times over the sum of the
    best times, minus one) and per-row statistics for the learned choice,
    the autotuner table, the fixed heuristic and the plain default tile."""
    from awkward_amd import dispatch

    best = times_ms.min(axis=1)
    learned, inside = choices(model, X, device)
    n = np.round(2 ** X[:, 0] - 1).astype(np.int64)
    N = np.round(2 ** X[:, 1] - 1).astype(np.int64)
    ops = np.round(X[:, 14] * 3).astype(int)
    its = np.round(X[:, 13] * 8).astype(int)
    table = np.array([model.table_choice(int(o), int(a), int(b), int(c)) for o, a, b, c in zip(ops, n, N, its)])
    heur = np.array([dispatch.heuristic_variant(int(a), int(b), int(c)) for a, b, c in zip(n, N, its)])
    fixed = np.full(len(best), 3)
    idx = np.arange(len(best))

    def stats(pick):
        t = times_ms[idx, pick]
        ratio = t / best
        return {"aggregate": round(float(t.sum() / best.sum() - 1.0), 4),
                "mean_row": round(float(ratio.mean() - 1.0), 4),
                "p95_row": round(float(np.percentile(ratio, 95) - 1.0), 4),
                "exact": round(float((pick == times_ms.argmin(axis=1)).mean()), 4)}

    out = {"rows": int(len(best)), "best_total_ms": round(float(best.sum()), 3),
           "learned": stats(learned), "table": stats(table), "heuristic": stats(heur), "fixed_variant3": stats(fixed),
           "fallback_rows": int((~inside).sum())}
    per_op = {}
    for o in sorted(set(ops.tolist())):
        m = ops == o
        b = best[m].sum()
        per_op[str(o)] = {k: round(float(times_ms[idx[m], p[m]].sum() / b - 1.0), 4)
                          for k, p in (("learned", learned), ("table", table), ("heuristic", heur))}
    out["per_op"] = per_op
    return out


def _seeds(spec):
    a, b = spec.split(":")
    return list(range(int(a), int(b)))


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m awkward_amd.learned")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--seeds", default="1000:2000")
    c.add_argument("--per-seed", type=int, default=20)
    c.add_argument("--max-elements", type=int, default=120_000_000)
    c.add_argument("--budget-s", type=float, default=None)
    c.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--data", required=True, nargs="+", help="one or more collected .npz files (concatenated)")
    t.add_argument("--out", default=DEFAULT_MODEL)
    t.add_argument("--epochs", type=int, default=80)
    t.add_argument("--batch", type=int, default=512)
    t.add_argument("--lr", type=float, default=3e-3)
    e = sub.add_parser("evaluate")
    e.add_argument("--model", default=DEFAULT_MODEL)
    e.add_argument("--data", required=True)
    e.add_argument("--out")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    device = torch.device("cuda", torch.cuda.current_device())
    if args.cmd == "collect":
        X, Tm, specs = collect(_seeds(args.seeds), args.per_seed, device, log=log, max_elements=args.max_elements,
                               budget_s=args.budget_s)
        np.savez_compressed(args.out, X=X, times=Tm, kinds=np.array([s["kind"] for s in specs]),
                            n=np.array([s["n"] for s in specs]), N=np.array([s["N"] for s in specs]))
        log(f"wrote {args.out}: {X.shape[0]} rows")
    elif args.cmd == "train":
        parts = [np.load(path) for path in args.data]
        X = np.concatenate([d["X"] for d in parts])
        Tm = np.concatenate([d["times"] for d in parts])
        t0 = time.time()
        model = train(X, Tm, device, epochs=args.epochs, bsz=args.batch, lr=args.lr, log=log)
        model.info["train_s"] = round(time.time() - t0, 2)
        model.save(args.out)
        log(json.dumps(model.info))
    else:
        d = np.load(args.data)
        model = Model.load(args.model)
        r = regret(model, d["X"], d["times"], device)
        log(json.dumps(r))
        if args.out:
            with open(args.out, "w") as f:
                json.dump(r, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
