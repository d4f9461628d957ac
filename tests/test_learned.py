"""The learned tile dispatcher (csrc/learned.hip, awkward_amd/learned.py).

The dispatcher changes which tile runs, never a result: the learned batch
path is checked against the oracle like every fixed tile.  Its own kernels
are pinned independently of their implementation:
* statistics/features: numpy over the same list lengths (integer counts,
  then the same double formulas);
* the network: a torch fp32 forward of the same parameters (tolerance 1e-4,
  different summation order), gradients: torch autograd in float64
  (tolerance 1e-3 relative), Adam: torch.optim.Adam (tolerance 1e-6);
* the fused inference equals the batched forward bit for bit (one code
  path), and out-of-range features take the caller's fallback.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from awkward_amd import learned  # noqa: E402

gpu = pytest.mark.gpu


def _lay():
    lay = learned.layout()
    F, H, O = lay["features"], lay["hidden"], lay["outputs"]
    k = {"W1": 0}
    k["b1"] = F * H
    k["W2"] = k["b1"] + H
    k["b2"] = k["W2"] + H * H
    k["W3"] = k["b2"] + H
    k["b3"] = k["W3"] + H * O
    k["params"] = k["b3"] + O
    k["mu"] = k["params"]
    k["inv"] = k["mu"] + F
    k["lo"] = k["inv"] + F
    k["hi"] = k["lo"] + F
    return lay, k


def test_layout_constants():
    lay, k = _lay()
    assert (lay["features"], lay["hidden"], lay["outputs"]) == (16, 64, 11)
    assert lay["params"] == k["params"] == 16 * 64 + 64 + 64 * 64 + 64 + 64 * 11 + 11
    assert lay["total"] == k["hi"] + 16


def test_targets_and_table_host_logic():
    t = np.array([[1.0, 2.0, 4.0] + [8.0] * 7, [0.5] * 10])
    T = learned.targets(t)
    assert T.shape == (2, 11)
    np.testing.assert_allclose(T[:, :10].sum(axis=1), 0.0, atol=1e-5)
    np.testing.assert_allclose(T[1, 10], np.log2(500.0), rtol=1e-6)
    # table: per bucket the tile with the least summed regret
    X = np.zeros((3, 16), dtype=np.float32)
    X[:, 0] = np.log2(1 + 1000)
    X[:, 1] = np.log2(1 + 10000)
    X[:, 13] = 1.0
    times = np.ones((3, 10))
    times[:, 7] = 0.5
    times[2, 7] = 5.0
    times[:, 2] = 0.6
    table = learned.build_table(X, times)
    assert list(table.values()) == [2]


def _np_features(counts, itemsize, op):
    counts = np.asarray(counts, dtype=np.int64)
    n, N = len(counts), int(counts.sum())
    nn, NN = max(n, 1), max(N, 1)
    mean = N / nn
    c = np.minimum(counts, 1 << 21).astype(np.float64)
    var = max(float((c * c).sum()) / nn - mean * mean, 0.0)
    bins = np.zeros(7)
    for length in counts[counts > 0]:
        if length <= 8:
            b = 0
        else:
            b = min(6, (int(np.ceil(np.log2(length))) - 2) // 2)
        bins[b] += length
    f = [np.log2(1 + n), np.log2(1 + N), np.log2(1 + mean), np.log2(1 + np.sqrt(var)),
         np.log2(1 + (counts.max() if n else 0)), (counts == 0).sum() / nn]
    f += list(bins / NN)
    f += [itemsize / 8.0, op / 3.0, (counts > 32).sum() / nn]
    return np.array(f, dtype=np.float32)


def _offsets(counts, device="cuda"):
    off = np.zeros(len(counts) + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    return torch.from_numpy(off).to(device)


@gpu
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_features_match_numpy(seed):
    rng = np.random.default_rng(seed)
    n = [5, 3000, 70000][seed]
    counts = np.concatenate([rng.poisson(3, n // 2), np.minimum(rng.zipf(1.4, n - n // 2), 50000)])
    counts[rng.random(n) < 0.1] = 0
    # bin edges exactly
    counts[:6] = [8, 9, 32, 33, 8192, 8193][: len(counts[:6])]
    off = _offsets(counts)
    got = learned.device_features(off, n, 4, 2)
    np.testing.assert_allclose(got, _np_features(counts, 4, 2), rtol=2e-6, atol=1e-7)


@gpu
def test_features_accumulate_slots():
    rng = np.random.default_rng(5)
    parts = [rng.poisson(4, 5000), rng.poisson(60, 9000), np.minimum(rng.zipf(1.6, 7000), 3000)]
    offs = [_offsets(c) for c in parts]
    lay = learned.layout()
    plan = learned.StatsPlan([(offs[0], 5000, 1), (offs[1], 9000, 0), (offs[2], 7000, 1)], 2, offs[0].device)
    plan.run()
    meta = torch.tensor([8, 3, 0, 0, 8, 3, 0, 0], dtype=torch.int32, device="cuda")
    params = torch.zeros(lay["total"], dtype=torch.float32, device="cuda")
    feats = torch.empty(2 * 16, dtype=torch.float32, device="cuda")
    choice = torch.empty(4, dtype=torch.int32, device="cuda")
    learned.infer(plan.acc, meta, 2, params, features=feats, choice=choice)
    f = feats.cpu().numpy().reshape(2, 16)
    np.testing.assert_allclose(f[0], _np_features(parts[1], 8, 3), rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(f[1], _np_features(np.concatenate([parts[0], parts[2]]), 8, 3), rtol=2e-6, atol=1e-7)


def _random_params(rng, lay, k, lo=-1e30, hi=1e30):
    p = np.zeros(lay["total"], dtype=np.float32)
    p[: k["params"]] = rng.normal(0, 0.3, k["params"])
    p[k["mu"]: k["mu"] + 16] = rng.normal(0, 1, 16)
    p[k["inv"]: k["inv"] + 16] = rng.uniform(0.5, 2.0, 16)
    p[k["lo"]: k["lo"] + 16] = lo
    p[k["hi"]: k["hi"] + 16] = hi
    return p


def _torch_forward(p, X, k, dtype=torch.float32):
    t = torch.from_numpy(p).to(dtype)
    W1 = t[k["W1"]: k["b1"]].reshape(16, 64)
    b1 = t[k["b1"]: k["W2"]]
    W2 = t[k["W2"]: k["b2"]].reshape(64, 64)
    b2 = t[k["b2"]: k["W3"]]
    W3 = t[k["W3"]: k["b3"]].reshape(64, 11)
    b3 = t[k["b3"]: k["params"]]
    x = (torch.from_numpy(X).to(dtype) - t[k["mu"]: k["mu"] + 16]) * t[k["inv"]: k["inv"] + 16]
    h1 = torch.relu(x @ W1 + b1)
    h2 = torch.relu(h1 @ W2 + b2)
    return h2 @ W3 + b3


@gpu
def test_forward_matches_torch():
    lay, k = _lay()
    rng = np.random.default_rng(11)
    p = _random_params(rng, lay, k)
    X = rng.normal(0, 2, (333, 16)).astype(np.float32)
    Y, inside = learned.forward(X, torch.from_numpy(p).cuda())
    ref = _torch_forward(p, X, k).numpy()
    np.testing.assert_allclose(Y, ref, rtol=1e-4, atol=1e-4)
    assert inside.all()


@gpu
def test_inside_flags_and_fallback():
    lay, k = _lay()
    rng = np.random.default_rng(12)
    p = _random_params(rng, lay, k, lo=-1.0, hi=1.0)
    X = rng.normal(0, 1, (200, 16)).astype(np.float32)
    _, inside = learned.forward(X, torch.from_numpy(p).cuda())
    Z = (X - p[k["mu"]: k["mu"] + 16]) * p[k["inv"]: k["inv"] + 16]
    np.testing.assert_array_equal(inside, ((Z >= -1.0) & (Z <= 1.0)).all(axis=1))


@gpu
def test_infer_equals_forward_and_takes_fallback():
    lay, k = _lay()
    rng = np.random.default_rng(13)
    counts = [rng.poisson(3, 4000), rng.poisson(40, 3000)]
    offs = [_offsets(c) for c in counts]
    for lo, hi, expect_fallback in ((-1e30, 1e30, False), (5.0, 6.0, True)):
        p = _random_params(rng, lay, k, lo=lo, hi=hi)
        pd = torch.from_numpy(p).cuda()
        plan = learned.StatsPlan([(offs[0], 4000, 0), (offs[1], 3000, 1)], 2, offs[0].device)
        plan.run()
        meta = torch.tensor([4, 0, 7, 0, 8, 2, 9, 0], dtype=torch.int32, device="cuda")
        feats = torch.empty(32, dtype=torch.float32, device="cuda")
        costs = torch.empty(22, dtype=torch.float32, device="cuda")
        choice = torch.empty(4, dtype=torch.int32, device="cuda")
        learned.infer(plan.acc, meta, 2, pd, features=feats, costs=costs, choice=choice)
        f = feats.cpu().numpy().reshape(2, 16)
        Y, inside = learned.forward(f, pd)
        np.testing.assert_array_equal(costs.cpu().numpy().reshape(2, 11), Y)
        c = choice.cpu().numpy().reshape(2, 2)
        if expect_fallback:
            assert not inside.any()
            np.testing.assert_array_equal(c, [[7, 1], [9, 1]])
        else:
            assert inside.all()
            np.testing.assert_array_equal(c[:, 0], np.argmin(Y[:, :10], axis=1))
            np.testing.assert_array_equal(c[:, 1], [0, 0])


def _two_launch(plan, meta, params):
    """stats_kernel then infer_kernel: features, costs and choice."""
    nslots = plan.nslots
    feats = torch.full((nslots * 16,), -1.0, dtype=torch.float32, device="cuda")
    costs = torch.full((nslots * 11,), -1.0, dtype=torch.float32, device="cuda")
    choice = torch.full((2 * nslots,), -1, dtype=torch.int32, device="cuda")
    plan.run()
    learned.infer(plan.acc, meta, nslots, params, features=feats, costs=costs, choice=choice)
    return feats.cpu().numpy(), costs.cpu().numpy(), choice.cpu().numpy()


def _one_launch(plan, meta, params, out=None):
    nslots = plan.nslots
    if out is None:
        out = (torch.full((nslots * 16,), -1.0, dtype=torch.float32, device="cuda"),
               torch.full((nslots * 11,), -1.0, dtype=torch.float32, device="cuda"),
               torch.full((2 * nslots,), -1, dtype=torch.int32, device="cuda"))
    plan.choose(meta, params, features=out[0], costs=out[1], choice=out[2])
    return out


@gpu
@pytest.mark.parametrize("case", ["many_tiles", "one_tile", "empty_slot", "no_lists", "fallback"])
def test_one_launch_choice_equals_stats_then_infer(case):
    """awkward_hip_dispatch_choose (statistics whose last workgroup runs the
    network) gives the features, costs and choice of the two-kernel form bit
    for bit, leaves its accumulators at zero, and so gives the same answer
    when called again (and after the two-kernel form left its sums behind)."""
    lay, k = _lay()
    rng = np.random.default_rng(21)
    parts = [rng.poisson(4, 300000), rng.poisson(60, 9000), np.minimum(rng.zipf(1.6, 70000), 3000),
             rng.poisson(2, 5)]
    offs = [_offsets(c) for c in parts]
    dev = offs[0].device
    entries = {"many_tiles": [(offs[0], 300000, 1), (offs[1], 9000, 0), (offs[2], 70000, 1)],
               "one_tile": [(offs[3], 5, 0)],
               "empty_slot": [(offs[1], 9000, 0), (offs[2], 70000, 2)],
               "no_lists": [(offs[3], 0, 0)],
               "fallback": [(offs[0], 300000, 0), (offs[2], 70000, 1)]}[case]
    nslots = {"many_tiles": 2, "one_tile": 1, "empty_slot": 3, "no_lists": 1, "fallback": 2}[case]
    p = _random_params(rng, lay, k, **({"lo": 5.0, "hi": 6.0} if case == "fallback" else {}))
    params = torch.from_numpy(p).cuda()
    meta = torch.tensor([[8, s % 4, 3 + s, 0] for s in range(nslots)], dtype=torch.int32).reshape(-1).cuda()
    plan = learned.StatsPlan(entries, nslots, dev)
    for _ in range(2):
        f, c, ch = (t.cpu().numpy() for t in _one_launch(plan, meta, params))
        assert not plan.acc.any().item()
        fr, cr, chr_ = _two_launch(learned.StatsPlan(entries, nslots, dev), meta, params)
        np.testing.assert_array_equal(f, fr)
        np.testing.assert_array_equal(c, cr)
        np.testing.assert_array_equal(ch, chr_)
    # the two-kernel form on the same plan leaves sums behind; the next one
    # launch clears them first
    _two_launch(plan, meta, params)
    assert plan.acc.any().item() == (case not in ("no_lists",))
    f2, c2, ch2 = (t.cpu().numpy() for t in _one_launch(plan, meta, params))
    np.testing.assert_array_equal(f2, f)
    np.testing.assert_array_equal(ch2, ch)
    ch = ch.reshape(nslots, 2)
    if case == "fallback":
        np.testing.assert_array_equal(ch, [[3, 1], [4, 1]])
    if case == "empty_slot":
        assert tuple(ch[1]) == (4, 1)  # no lists: the fallback
    if case == "no_lists":
        assert tuple(ch[0]) == (3, 1)


@gpu
def test_one_launch_choice_in_a_graph():
    """The one-launch choice captured in a hipGraph: the accumulators end
    every replay at zero, so every replay (eager library work and
    synchronisations in between) equals the eager answer."""
    from awkward_amd.fused import Graph

    lay, k = _lay()
    rng = np.random.default_rng(22)
    parts = [rng.poisson(30, 200000), np.minimum(rng.zipf(1.5, 50000), 5000)]
    offs = [_offsets(c) for c in parts]
    p = _random_params(rng, lay, k)
    params = torch.from_numpy(p).cuda()
    meta = torch.tensor([4, 0, 2, 0, 8, 3, 5, 0], dtype=torch.int32, device="cuda")
    plan = learned.StatsPlan([(offs[0], 200000, 0), (offs[1], 50000, 1)], 2, offs[0].device)
    out = _one_launch(plan, meta, params)
    torch.cuda.synchronize()
    eager = [t.clone() for t in out]
    g = Graph(lambda: _one_launch(plan, meta, params, out))
    other = learned.StatsPlan([(offs[1], 50000, 0)], 1, offs[0].device)
    for _ in range(3):
        for t in out:
            t.fill_(-7)
        _two_launch(other, meta[:4], params)  # eager work between the replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(out, eager):
            assert torch.equal(a, b)
        assert not plan.acc.any().item()


@gpu
def test_train_step_gradient_matches_autograd():
    lay, k = _lay()
    rng = np.random.default_rng(21)
    p = _random_params(rng, lay, k)
    R = 150  # three workgroups, the last one partial
    Xn = rng.normal(0, 1, (R, 16)).astype(np.float32)
    T = rng.normal(0, 1, (R, 11)).astype(np.float32)
    ow = np.array(learned.OUTPUT_WEIGHTS, dtype=np.float32)
    tr = learned.Trainer(Xn, T, p[: k["params"]], torch.device("cuda"))
    tr._buffers(R)
    learned._check(learned._fn("awkward_hip_mlp_train_step", [learned.P, learned.P, learned.P] + [learned.I64] * 5
                               + [learned.P] * 3)(
        tr.X.data_ptr(), tr.T.data_ptr(), tr.ow.data_ptr(), R, 1, 0, 0, R, tr.P.data_ptr(), tr._grads.data_ptr(),
        tr._losses.data_ptr()))
    grads = tr._grads[:3].sum(dim=0).cpu().numpy().astype(np.float64)
    loss = float(tr._losses[:3].sum().item())

    # the training kernel takes features already normalised (no mu / inv_sigma)
    q = torch.from_numpy(p[: k["params"]].astype(np.float64)).requires_grad_(True)
    W1 = q[k["W1"]: k["b1"]].reshape(16, 64)
    b1 = q[k["b1"]: k["W2"]]
    W2 = q[k["W2"]: k["b2"]].reshape(64, 64)
    b2 = q[k["b2"]: k["W3"]]
    W3 = q[k["W3"]: k["b3"]].reshape(64, 11)
    b3 = q[k["b3"]: k["params"]]
    x = torch.from_numpy(Xn.astype(np.float64))
    y = torch.relu(torch.relu(x @ W1 + b1) @ W2 + b2) @ W3 + b3
    ref_loss = ((y - torch.from_numpy(T.astype(np.float64))) ** 2 * torch.from_numpy(ow.astype(np.float64))).sum() / R
    ref_loss.backward()
    np.testing.assert_allclose(loss, float(ref_loss.detach()), rtol=1e-4)
    np.testing.assert_allclose(grads, q.grad.numpy(), rtol=1e-3, atol=2e-5)


@gpu
def test_adam_matches_torch():
    lay, k = _lay()
    rng = np.random.default_rng(22)
    P0 = rng.normal(0, 0.5, k["params"]).astype(np.float32)
    dev = torch.device("cuda")
    Pd = torch.from_numpy(P0.copy()).to(dev)
    m = torch.zeros_like(Pd)
    v = torch.zeros_like(Pd)
    ref = torch.nn.Parameter(torch.from_numpy(P0.copy()))
    opt = torch.optim.Adam([ref], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    loss_acc = torch.zeros(1, dtype=torch.float32, device=dev)
    for step in range(1, 4):
        g = rng.normal(0, 1, (3, k["params"])).astype(np.float32)
        losses = torch.tensor([0.5, 0.25, 0.25], dtype=torch.float32, device=dev)
        gd = torch.from_numpy(g).to(dev)
        learned._check(learned._fn("awkward_hip_mlp_adam", [learned.P] * 5 + [learned.I32] + [learned.F32] * 4
                                   + [learned.I64, learned.P])(
            Pd.data_ptr(), m.data_ptr(), v.data_ptr(), gd.data_ptr(), losses.data_ptr(), 3, 1e-2, 0.9, 0.999, 1e-8,
            step, loss_acc.data_ptr()))
        ref.grad = torch.from_numpy(g.sum(axis=0))
        opt.step()
    np.testing.assert_allclose(Pd.cpu().numpy(), ref.detach().numpy(), rtol=0, atol=2e-6)
    assert abs(float(loss_acc.item()) - 3.0) < 1e-6


@gpu
def test_training_reduces_loss():
    rng = np.random.default_rng(23)
    X = rng.normal(0, 1, (2000, 16)).astype(np.float32)
    w = rng.normal(0, 1, (16, 10))
    times = 2.0 ** (X @ w * 0.3 + 3.0)  # a learnable cost surface
    # the relative targets have variance ~1.4 per output (~14 summed): a
    # trained network must get well below that
    model = learned.train(X, times, torch.device("cuda"), epochs=60, bsz=256, lr=3e-3)
    assert model.info["loss"] < 1.5
    r = learned.regret(model, X[:300], times[:300], torch.device("cuda"))
    assert r["learned"]["aggregate"] < r["fixed_variant3"]["aggregate"]


@gpu
@pytest.mark.parametrize("op", [0, 1, 2, 3])
def test_time_variants_sweep(op):
    rng = np.random.default_rng(31)
    counts = rng.poisson(5, 20000)
    off = _offsets(counts)
    data = torch.randn(int(counts.sum()), device="cuda", dtype=torch.float64)
    ms = learned.time_variants(op, off, data, 20000, 2)
    assert ms.shape == (10,)
    assert (ms > 0).all() and (ms < 100).all()


@gpu
def test_batch_learned_matches_oracle():
    from test_batch import check_array, make_batch

    from awkward_amd.batch import BatchRunner
    from awkward_amd.fused import Graph

    lay, k = _lay()
    rng = np.random.default_rng(41)
    # an untrained model: the choice is arbitrary, the results must not change
    disp = learned.Dispatcher(learned.Model(_random_params(rng, lay, k), {}))
    batch = make_batch(7, narrays=12)
    dev = [(torch.from_numpy(o).cuda(), torch.from_numpy(x).cuda()) for o, x in batch]
    runner = BatchRunner(dev, variants="learned", dispatcher=disp)
    graph = Graph(runner.run)
    for rep in range(3):
        if rep == 0:
            runner.run()
        else:
            graph.replay()
        torch.cuda.synchronize()
        for i, (o, x) in enumerate(batch):
            check_array(o, x, runner.result(i))
    for g in runner.groups:
        assert 0 <= g.chosen() < 10


@gpu
def test_enable_routes_tuned_reducers():
    import awkward_amd as ak
    from awkward_amd.backend import HipBackend
    from oracle.backend import OracleBackend
    from parity import same_bits
    from test_pipeline import flat

    lay, k = _lay()
    rng = np.random.default_rng(42)
    model = learned.Model(_random_params(rng, lay, k), {})
    counts = rng.poisson(7, 5000)
    off = np.zeros(5001, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    x = rng.standard_normal(int(off[-1]))
    xi = rng.integers(-50, 50, int(off[-1]))
    d = learned.enable(model)
    try:
        hb = HipBackend(fused=True)
        for data in (x, xi):
            a = ak.from_offsets(off, data, hb)
            b = ak.from_offsets(off, data, OracleBackend())
            for fn in (ak.sum, ak.max, ak.argmax, ak.min, ak.argmin):
                for u, v in zip(flat(fn(b, axis=-1)), flat(fn(a, axis=-1))):
                    assert same_bits(u, v), fn.__name__
        assert d.last_fallback is False
    finally:
        learned.disable()


def test_committed_model():
    """The trained dispatcher shipped with the package: the device layout's
    size, finite parameters, a sane range box, the autotuner table, and the
    validation regret recorded at training time within the 10% target."""
    lay, k = _lay()
    model = learned.Model.load(learned.DEFAULT_MODEL)
    assert model.params.shape == (lay["total"],)
    assert np.isfinite(model.params).all()
    lo = model.params[k["lo"]: k["lo"] + 16]
    hi = model.params[k["hi"]: k["hi"] + 16]
    assert (lo < hi).all()
    assert len(model.table) > 100
    assert all(0 <= int(v) < 10 for v in model.table.values())
    assert model.info["val_regret"] < 0.10
