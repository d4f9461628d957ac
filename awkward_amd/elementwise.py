"""Elementwise ufuncs of the hip nplike as device expressions.

The reference applies numpy/cupy ufuncs one at a time, each writing a
temporary (``src/awkward/_nplikes/array_module.py:254-332``).  Here a ufunc
on device columns returns a ``DeviceExpr``: a node of a recorded expression
tree whose leaves are columns -- optionally gathered through an int64 index,
so a lazy carry (``IndexedArray``) is read in place rather than copied.  The
tree becomes one program of ``csrc/expr.h`` when it is needed:

* ``materialize()`` writes it with ``awkward_hip_expr_eval`` (one pass over
  the leaves, one write of the result);
* ``awkward_hip_mask_getitem`` evaluates a boolean tree inside the jagged
  getitem's scan, so a mask such as ``(ev.pt > 25) & (abs(ev.eta) < 2.4)``
  is never written at all.

Types follow numpy exactly: every ufunc's loop and result dtypes come from
``ufunc.resolve_dtypes`` of the operands' dtypes, with Python scalars passed
as the weak ``int`` / ``float`` / ``bool`` of NEP 50, and the inputs are cast
to the loop dtypes before the op, as numpy's loops do.  Comparisons with a
Python integer outside the loop type's range give numpy's exact (uniform)
answer; arithmetic with one raises numpy's OverflowError.
"""

from __future__ import annotations

import ctypes
import heapq

import numpy as np

#: expr.h's ExprType codes
TYPE_CODE = {
    np.dtype(np.bool_): 0, np.dtype(np.int8): 1, np.dtype(np.uint8): 2, np.dtype(np.int16): 3,
    np.dtype(np.uint16): 4, np.dtype(np.int32): 5, np.dtype(np.uint32): 6, np.dtype(np.int64): 7,
    np.dtype(np.uint64): 8, np.dtype(np.float32): 9, np.dtype(np.float64): 10,
}

#: expr.h's ExprCode values
CONST, CAST, ADD, SUB, MUL, DIV, GT, GE, LT, LE, EQ, NE, AND, OR, XOR, LAND, LOR, LNOT, INVERT, NEG, ABS, \
    COSH, COS, SINH, SIN, EXP, SQRT, MAX, MIN = range(29)

BINARY = {
    "add": ADD, "subtract": SUB, "multiply": MUL, "true_divide": DIV, "divide": DIV,
    "greater": GT, "greater_equal": GE, "less": LT, "less_equal": LE, "equal": EQ, "not_equal": NE,
    "bitwise_and": AND, "bitwise_or": OR, "bitwise_xor": XOR, "logical_and": LAND, "logical_or": LOR,
    "maximum": MAX, "minimum": MIN,
}
UNARY = {
    "absolute": ABS, "negative": NEG, "invert": INVERT, "logical_not": LNOT, "cosh": COSH, "cos": COS,
    "sinh": SINH, "sin": SIN, "exp": EXP, "sqrt": SQRT,
}
COMPARISONS = {GT, GE, LT, LE, EQ, NE}

MAX_LEAVES = 8
MAX_OPS = 40
MAX_REGS = 16


class _Leaf(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("index", ctypes.c_void_p), ("type", ctypes.c_int32),
                ("leg", ctypes.c_int32)]


class _Op(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("type", ctypes.c_uint8), ("dst", ctypes.c_uint8), ("a", ctypes.c_uint8),
                ("b", ctypes.c_uint8), ("to", ctypes.c_uint8), ("bimm", ctypes.c_uint8), ("pad1", ctypes.c_uint8),
                ("imm", ctypes.c_uint64)]


class Program(ctypes.Structure):
    """expr.h's ExprProgram."""

    _fields_ = [("nleaves", ctypes.c_int32), ("nops", ctypes.c_int32), ("out", ctypes.c_int32),
                ("out_type", ctypes.c_int32), ("wide", ctypes.c_int32), ("nregs", ctypes.c_int32),
                ("leaf", _Leaf * MAX_LEAVES), ("op", _Op * MAX_OPS)]


def _bits(value, dtype: np.dtype) -> int:
    """Canonical register bits of a scalar of `dtype` (expr.h's layout)."""
    x = np.array(value).astype(dtype) if not isinstance(value, np.ndarray) else value.astype(dtype)
    if dtype.kind == "f":
        return int(x.view(np.uint32 if dtype.itemsize == 4 else np.uint64))
    if dtype.kind == "b":
        return int(bool(x))
    v = int(x)
    return v & (2**64 - 1)


class Node:
    """One node of an expression tree (a leaf column, a constant or an op)."""

    __slots__ = ("kind", "dtype", "code", "children", "data", "index", "value", "loop", "leg")

    def __init__(self, kind, dtype, code=None, children=(), data=None, index=None, value=None, loop=None, leg=0):
        self.kind, self.dtype, self.code, self.children = kind, np.dtype(dtype), code, tuple(children)
        self.data, self.index, self.value, self.loop, self.leg = data, index, value, loop, leg

    def key(self):
        """Identity of a leaf or constant (ops are identified by their
        children's positions when a program is compiled)."""
        if self.kind == "leaf":
            return ("leaf", self.data.data_ptr(), None if self.index is None else self.index.data_ptr(),
                    self.dtype.str, self.leg)
        return ("const", self.value, self.dtype.str)


def leaf(data, index=None, leg=0) -> Node:
    """A column read at each element's position (through `index` when given);
    in a pair program (awkward_hip_pair_reduce) `leg` 0 / 1 reads it at the
    pair's first / second element."""
    from awkward_amd.nplike import to_numpy_dtype

    return Node("leaf", to_numpy_dtype(data.dtype), data=data, index=index, leg=leg)


def const(value, dtype) -> Node:
    return Node("const", dtype, value=_bits(value, np.dtype(dtype)))


def _value(node: Node):
    """The Python value of a constant node."""
    bits = np.array([node.value], dtype=np.uint64)
    if node.dtype == np.float32:
        return float(bits.astype(np.uint32).view(np.float32)[0])
    if node.dtype == np.float64:
        return float(bits.view(np.float64)[0])
    if node.dtype.kind == "b":
        return bool(node.value)
    if node.dtype.kind == "i":
        return int(bits.view(np.int64)[0])
    return int(node.value)


def cast(node: Node, dtype) -> Node:
    dtype = np.dtype(dtype)
    if node.dtype == dtype:
        return node
    if node.kind == "const":  # folded with numpy's cast
        return const(np.array(_value(node), dtype=node.dtype).astype(dtype), dtype)
    return Node("op", dtype, CAST, (node,), loop=node.dtype)


class DeviceExpr:
    """A lazy elementwise result of the hip nplike: an expression tree over
    device columns, with numpy's dtype and the leaves' length."""

    def __init__(self, nplike, node: Node, length: int):
        self.nplike = nplike
        self.node = node
        self.length = int(length)
        self._value = None

    @property
    def dtype(self):
        return self.node.dtype

    @property
    def shape(self):
        return (self.length,)

    def program(self) -> tuple[Program, list]:
        return compile_program(self.node)

    def materialize(self):
        """The expression written to a new device tensor (awkward_hip_expr_eval)."""
        if self._value is None:
            self._value = self.nplike.eval_expr(self)
        return self._value


def _weak(x):
    """numpy's NEP 50 view of a Python scalar for resolve_dtypes."""
    if isinstance(x, bool):
        return bool
    if isinstance(x, int):
        return int
    if isinstance(x, float):
        return float
    raise TypeError(f"unsupported scalar {x!r}")


def build(nplike, ufunc, args) -> DeviceExpr:
    """The DeviceExpr of ufunc(*args); args are DeviceExpr, device tensors,
    (data, index) gathers, or Python / numpy scalars."""
    import torch

    name = ufunc.__name__
    if name in BINARY:
        code = BINARY[name]
        if len(args) != 2:
            raise TypeError(f"{name} takes two operands")
    elif name in UNARY:
        code = UNARY[name]
        if len(args) != 1:
            raise TypeError(f"{name} takes one operand")
    else:
        raise NotImplementedError(f"ufunc {name} on the hip backend")
    nodes, dtypes, length = [], [], None
    for a in args:
        if isinstance(a, DeviceExpr):
            nodes.append(a.node)
            dtypes.append(a.dtype)
            length = a.length if length is None else length
        elif isinstance(a, torch.Tensor):
            n = leaf(a)
            nodes.append(n)
            dtypes.append(n.dtype)
            length = a.shape[0] if length is None else length
        elif isinstance(a, tuple):  # (data, index): data[index]
            n = leaf(a[0], a[1])
            nodes.append(n)
            dtypes.append(n.dtype)
            length = a[1].shape[0] if length is None else length
        elif isinstance(a, np.generic):
            nodes.append(("scalar", a))
            dtypes.append(a.dtype)
        else:
            nodes.append(("scalar", a))
            dtypes.append(_weak(a))
    if length is None:
        raise TypeError("an elementwise operation needs at least one array operand")
    # numpy's own answer on one-element operands: the result dtype, and the
    # errors (OverflowError for out-of-range Python integers in arithmetic)
    probe = [np.zeros(1, d) if isinstance(d, np.dtype) else n[1] for n, d in zip(nodes, dtypes)]
    out_dtype = np.dtype(ufunc(*probe).dtype)
    loop = ufunc.resolve_dtypes(tuple(dtypes) + (None,))
    in_loop = [np.dtype(t) for t in loop[:-1]]
    if out_dtype == np.float16 or np.dtype(np.float16) in in_loop:
        raise NotImplementedError(f"{name} computes in float16 here, which the hip backend does not support")
    operands = []
    for k, (n, lt) in enumerate(zip(nodes, in_loop)):
        if isinstance(n, tuple):
            v = n[1]
            if code in COMPARISONS and isinstance(v, int) and not isinstance(v, bool) and lt.kind in "iu":
                info = np.iinfo(lt)
                if not info.min <= v <= info.max:
                    # numpy compares exactly: the answer is the same for every element
                    other = in_loop[1 - k]
                    ends = np.array([np.iinfo(other).min, np.iinfo(other).max], dtype=other)
                    res = ufunc(ends, v) if k == 1 else ufunc(v, ends)
                    return DeviceExpr(nplike, const(bool(res[0]), np.bool_), length) if res[0] == res[1] else \
                        _raise(name)
            operands.append(const(v, lt))
        else:
            operands.append(cast(n, lt))
    if code in COMPARISONS and in_loop[0] != in_loop[1]:
        # numpy's mixed uint64 / int64 comparison loops compare exactly
        node = Node("op", out_dtype, code, operands, loop=in_loop[0])
        node.value = 1 if in_loop[0] == np.uint64 else 2
        return DeviceExpr(nplike, node, length)
    return DeviceExpr(nplike, Node("op", out_dtype, code, operands, loop=in_loop[0]), length)


def _raise(name):
    raise NotImplementedError(f"{name} with this Python integer on the hip backend")


COMMUTATIVE = {ADD, MUL, AND, OR, XOR, MAX, MIN, EQ, NE, LAND, LOR}
MIRROR = {GT: LT, GE: LE, LT: GT, LE: GE}


def compile_program(root: Node):
    """(Program, keep): the tree as expr.h's program, common subtrees
    evaluated once, constants folded into their ops as immediate second
    operands (swapping the operands of commutative ops and mirroring
    comparisons where the constant comes first), registers reused after
    their last use.  `keep` holds the tensors the program points to."""
    order, index = [], {}

    def visit(n):
        kids = [visit(c) for c in n.children]
        k = n.key() if n.kind != "op" else ("op", n.code, n.dtype.str, n.loop.str, n.value, tuple(kids))
        if k in index:
            return index[k]
        index[k] = len(order)
        order.append([n, kids, n.code, n.value, False])  # node, kids, code, mixed side, immediate b
        return index[k]

    out_id = visit(root)
    for e in order:
        n, kids = e[0], e[1]
        if n.kind != "op" or len(kids) != 2:
            continue
        if order[kids[0]][0].kind == "const" and order[kids[1]][0].kind != "const":
            if e[2] in COMMUTATIVE or e[2] in MIRROR:
                e[1] = [kids[1], kids[0]]
                e[2] = MIRROR.get(e[2], e[2])
                if e[3] in (1, 2):
                    e[3] = 3 - e[3]
        if order[e[1][1]][0].kind == "const":
            e[4] = True
    leaves = [i for i, e in enumerate(order) if e[0].kind == "leaf"]
    if len(leaves) > MAX_LEAVES:
        raise OverflowError("expression has too many columns")
    last_use = {}
    for i, e in enumerate(order):
        for c in (e[1][:1] if e[4] else e[1]):
            last_use[c] = i
    last_use[out_id] = len(order)
    # registers 0 .. nleaves - 1 start with the leaves; the lowest free
    # register takes each result, so `nregs` (and with it the device's
    # register file) stays small
    reg, free = {}, list(range(len(leaves), MAX_REGS))
    prog = Program()
    keep = []
    for r, i in enumerate(leaves):
        n = order[i][0]
        reg[i] = r
        prog.leaf[r].data = n.data.data_ptr()
        prog.leaf[r].index = 0 if n.index is None else n.index.data_ptr()
        prog.leaf[r].type = TYPE_CODE[n.dtype]
        prog.leaf[r].leg = n.leg
        keep.extend([n.data, n.index])
    nops = 0
    for i, (n, kids, code, side, bimm) in enumerate(order):
        if n.kind == "leaf" or (n.kind == "const" and i not in last_use):
            continue  # leaves sit in their registers; constants used only as immediates need none
        # operands whose last use is this op free their registers for its result
        for c in set(kids[:1] if bimm else kids):
            if last_use.get(c) == i:
                heapq.heappush(free, reg[c])
        if not free:
            raise OverflowError("expression needs too many registers")
        dst = heapq.heappop(free)
        reg[i] = dst
        if nops >= MAX_OPS:
            raise OverflowError("expression has too many operations")
        o = prog.op[nops]
        nops += 1
        o.dst = dst
        if n.kind == "const":
            o.code, o.type, o.imm = CONST, TYPE_CODE[n.dtype], n.value
            continue
        o.code = code
        o.type = TYPE_CODE[n.loop]
        o.a = reg[kids[0]]
        if bimm:
            o.bimm, o.imm = 1, order[kids[1]][0].value
        else:
            o.b = reg[kids[1]] if len(kids) > 1 else reg[kids[0]]
        if code == CAST:
            o.to = TYPE_CODE[n.dtype]
        elif code in COMPARISONS and side in (1, 2):
            o.to = side
    prog.nleaves = len(leaves)
    prog.nops = nops
    # 64-bit registers only when some value of the program needs them
    prog.wide = int(any(e[0].dtype.itemsize == 8 or (e[0].loop is not None and e[0].loop.itemsize == 8)
                        for e in order))
    prog.out = reg[out_id]
    prog.out_type = TYPE_CODE[root.dtype]
    prog.nregs = max([prog.out, len(leaves) - 1] + [prog.op[k].dst for k in range(nops)]) + 1
    return prog, keep
