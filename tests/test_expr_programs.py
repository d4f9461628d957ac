"""The host side of the expression programs (awkward_amd/elementwise.py),
on the CPU: every compiled program satisfies the device's validity rule
(expr.h expr_program_ok: every register number below the program's
nregs <= 16),
and running the compiled op list with a small numpy interpreter gives what
numpy gives for the same ufunc chain -- so register allocation, common
subexpressions, immediates and operand swaps are checked without a GPU."""

import numpy as np
import pytest
import torch

from awkward_amd import elementwise as E

TYPES = {v: k for k, v in E.TYPE_CODE.items()}


def leafx(t, leg=0):
    return E.DeviceExpr(None, E.leaf(t, leg=leg), t.shape[0])


def build(uf, *args):
    return E.build(None, uf, list(args))


def valid(prog):
    assert 0 <= prog.nleaves <= E.MAX_LEAVES and 0 <= prog.nops <= E.MAX_OPS
    assert max(prog.nleaves, 1) <= prog.nregs <= E.MAX_REGS and 0 <= prog.out < prog.nregs
    for k in range(prog.nops):
        o = prog.op[k]
        assert o.dst < prog.nregs and o.a < prog.nregs and o.b < prog.nregs, (k, o.dst, o.a, o.b)


_NP = {E.ADD: np.add, E.SUB: np.subtract, E.MUL: np.multiply, E.DIV: np.true_divide, E.GT: np.greater,
       E.GE: np.greater_equal, E.LT: np.less, E.LE: np.less_equal, E.EQ: np.equal, E.NE: np.not_equal,
       E.AND: np.bitwise_and, E.OR: np.bitwise_or, E.XOR: np.bitwise_xor, E.MAX: np.maximum, E.MIN: np.minimum,
       E.NEG: np.negative, E.ABS: np.absolute, E.COSH: np.cosh, E.COS: np.cos, E.SINH: np.sinh, E.SIN: np.sin,
       E.EXP: np.exp, E.SQRT: np.sqrt, E.LAND: np.logical_and, E.LOR: np.logical_or, E.LNOT: np.logical_not,
       E.INVERT: np.invert}


def _imm(bits, dt):
    a = np.array([bits], dtype=np.uint64)
    if dt == np.float32:
        return a.astype(np.uint32).view(np.float32)[0]
    if dt == np.float64:
        return a.view(np.float64)[0]
    if dt.kind == "b":
        return np.bool_(bits)
    return a.view(np.int64).astype(dt)[0] if dt.kind == "i" else a.astype(dt)[0]


def interpret(prog, columns):
    """The op list run on numpy arrays: registers 0 .. nleaves - 1 start
    with the leaves; every op computes in its `type`."""
    regs = {}
    for k in range(prog.nleaves):
        regs[k] = columns[prog.leaf[k].data]
    for k in range(prog.nops):
        o = prog.op[k]
        t = TYPES[o.type]
        if o.code == E.CONST:
            regs[o.dst] = _imm(o.imm, t)
            continue
        a = np.asarray(regs[o.a]).astype(t)
        if o.code == E.CAST:
            regs[o.dst] = a.astype(TYPES[o.to])
            continue
        f = _NP[o.code]
        if f.nin == 1:
            regs[o.dst] = f(a)
        else:
            b = _imm(o.imm, t) if o.bimm else np.asarray(regs[o.b]).astype(t)
            regs[o.dst] = f(a, b)
    return np.asarray(regs[prog.out]).astype(TYPES[prog.out_type])


def check(expr, columns, want):
    prog, _ = expr.program()
    valid(prog)
    got = interpret(prog, columns)
    assert got.dtype == want.dtype
    with np.errstate(all="ignore"):
        assert np.array_equal(got, want, equal_nan=want.dtype.kind == "f")
    return prog


@pytest.fixture
def cols():
    rng = np.random.default_rng(1)
    arrays = {name: rng.standard_normal(64).astype(np.float32) for name in ("pt", "eta", "phi")}
    arrays["q"] = rng.integers(-3, 4, 64).astype(np.int8)
    tensors = {k: torch.from_numpy(v) for k, v in arrays.items()}
    by_ptr = {t.data_ptr(): arrays[k] for k, t in tensors.items()}
    return arrays, tensors, by_ptr


def test_pair_mass_program(cols):
    a, t, by_ptr = cols
    p0 = {k: leafx(t[k], 0) for k in ("pt", "eta", "phi")}
    p1 = {k: leafx(t[k], 1) for k in ("pt", "eta", "phi")}
    expr = build(np.multiply, build(np.multiply, build(np.multiply, 2, p0["pt"]), p1["pt"]),
                 build(np.subtract, build(np.cosh, build(np.subtract, p0["eta"], p1["eta"])),
                       build(np.cos, build(np.subtract, p0["phi"], p1["phi"]))))
    prog, _ = expr.program()
    valid(prog)
    assert prog.nleaves == 6 and sorted(prog.leaf[k].leg for k in range(6)) == [0, 0, 0, 1, 1, 1]
    assert prog.op[0].bimm == 1  # the constant 2 became an immediate second operand
    assert prog.nregs <= 8  # fits the 4-slot register file


def test_one_leg_programs_interpret_like_numpy(cols):
    a, t, by_ptr = cols
    pt, eta, phi, q = (leafx(t[k]) for k in ("pt", "eta", "phi", "q"))
    with np.errstate(all="ignore"):
        check(build(np.multiply, build(np.multiply, 2, pt), build(np.subtract, build(np.cosh, eta), build(np.cos, phi))),
              by_ptr, 2 * a["pt"] * (np.cosh(a["eta"]) - np.cos(a["phi"])))
        # the same subtree twice is evaluated once
        s = build(np.add, pt, eta)
        prog = check(build(np.multiply, s, s), by_ptr, (a["pt"] + a["eta"]) * (a["pt"] + a["eta"]))
        assert prog.nops == 2
        # a constant first: the operands swap (commutative) or the comparison mirrors
        check(build(np.less, 0.5, pt), by_ptr, 0.5 < a["pt"])
        check(build(np.subtract, 1, pt), by_ptr, 1 - a["pt"])
        # mixed types: int8 * float32 -> float32 through a cast
        check(build(np.multiply, q, pt), by_ptr, a["q"] * a["pt"])
        check(build(np.logical_and, build(np.greater, pt, 0), build(np.less, q, 2)), by_ptr,
              (a["pt"] > 0) & (a["q"] < 2))


def test_deep_trees_fit_the_registers(cols):
    """A balanced tree over 8 leaves and a
    left-leaning chain of 40 ops both fit the sixteen registers."""
    a, t, by_ptr = cols
    rng = np.random.default_rng(2)
    arrays = [rng.standard_normal(16).astype(np.float32) for _ in range(8)]
    ts = [torch.from_numpy(x) for x in arrays]
    by = {x.data_ptr(): arr for x, arr in zip(ts, arrays)}
    level = [leafx(x) for x in ts]
    want = list(arrays)
    while len(level) > 1:
        level = [build(np.multiply, build(np.cos, level[i]), level[i + 1]) for i in range(0, len(level), 2)]
        want = [np.cos(want[i]) * want[i + 1] for i in range(0, len(want), 2)]
    check(level[0], by, want[0])
    x, w = leafx(ts[0]), arrays[0]
    for k in range(39):
        x = build(np.add, build(np.multiply, x, 0.5), leafx(ts[k % 8]))
        w = w * np.float32(0.5) + arrays[k % 8]
        if 2 * (k + 1) >= E.MAX_OPS:
            break
    check(x, by, w)


def test_too_many_registers_raise(cols):
    """Seventeen live values at once exceed the sixteen registers."""
    rng = np.random.default_rng(3)
    ts = [torch.from_numpy(rng.standard_normal(8).astype(np.float32)) for _ in range(8)]
    parts = [build(np.cos, leafx(x)) for x in ts] + [build(np.sin, leafx(x)) for x in ts] + \
        [build(np.exp, leafx(ts[0]))]
    total = parts[-1]
    for p in reversed(parts[:-1]):
        total = build(np.add, p, total)  # right-leaning: every value is live until the end
    with pytest.raises(OverflowError):
        total.program()
    half = parts[8]
    for p in reversed(parts[:8]):
        half = build(np.add, p, half)  # nine live results fit in sixteen registers
    prog, _ = half.program()
    valid(prog)
    assert prog.nregs > 8


def test_selection_cuts_recognised(cols):
    """The library's cut matcher (host code, no device work) takes exactly
    the conjunctions of float32 column comparisons with constants, through
    abs or not; anything else stays with the expression interpreter."""
    import ctypes

    from awkward_amd._lib import lib

    f = lib().awkward_hip_expr_cut_terms
    f.argtypes = [ctypes.c_void_p]
    f.restype = ctypes.c_int
    a, t, by_ptr = cols
    pt, eta, phi, q = (leafx(t[k]) for k in ("pt", "eta", "phi", "q"))

    def terms(expr):
        prog, _ = expr.program()
        return f(ctypes.addressof(prog))

    gt, lt, ab, band = np.greater, np.less, np.absolute, np.bitwise_and
    assert terms(build(gt, pt, 25)) == 1
    assert terms(build(band, build(gt, pt, 25), build(lt, build(ab, eta), 2.4))) == 2
    assert terms(build(np.logical_and, build(lt, 0.5, pt), build(np.not_equal, phi, 0))) == 2
    eight = build(gt, pt, 0)
    for k in range(7):
        eight = build(band, eight, build(lt, [pt, eta, phi][k % 3], 10 + k))
    assert terms(eight) == 8
    nine = build(band, eight, build(gt, phi, -9))
    assert terms(nine) == 0  # more terms than the cut kernel takes
    assert terms(build(np.bitwise_or, build(gt, pt, 25), build(lt, eta, 1))) == 0
    assert terms(build(np.invert, build(gt, pt, 25))) == 0
    assert terms(build(gt, build(np.add, pt, eta), 0.5)) == 0
    assert terms(build(lt, pt, eta)) == 0  # column against column
    assert terms(build(lt, q, 2)) == 0  # an int8 column
    # a Python float is weak (NEP 50): compared in float32, against +inf here
    assert terms(build(np.less, pt, 2.5e300)) == 1
    d = leafx(torch.from_numpy(a["pt"].astype(np.float64)))
    assert terms(build(lt, d, 1.0)) == 0  # a float64 column
    assert terms(build(lt, pt, np.float64(1.0))) == 0  # a float64 scalar makes it a float64 comparison
