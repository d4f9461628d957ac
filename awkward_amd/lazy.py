"""Lazy front end: record a graph of jagged-array operations, then run it
either node by node through the host layer or, where a part of the graph is
one of the fused forms, as one kernel (and, for repeated runs on the same
inputs, from a captured hipGraph).

Mirrors the reference's operator IR (``src/awkward/_connect/lazy/_ir.py``:
``OpType`` FILTER / REDUCE / COMBINATIONS / GETITEM and the element-wise and
comparison nodes; ``_lazy_impl.py``: ``LazyAwkwardArray`` with ``compute``,
``compile``, ``fusion_stats``, ``invalidate``, ``visualize``), with the
fusion pass (``_fusion.py``) replaced by pattern matching onto this
backend's fused kernels:

* ``ak.num`` / ``ak.count`` / ``ak.sum`` / ``ak.max`` / ``ak.argmax`` of
  ``ev[(ev.a OP c) & (abs(ev.b) OP d) & ...][:, start:stop:step]`` (or of a
  float32 field of it), on one list-of-records array, are evaluated together
  by ``awkward_hip_event_pipeline``: every such reduction that shares the
  input, predicate and slice is one output of the same launch.

Everything else runs eagerly through the host layer, one node at a time,
memoised per node.  The fused and eager results are the same bits (the
fused kernel reproduces the per-kernel path; ``tests/test_lazy.py``).
"""

from __future__ import annotations

import itertools

import numpy as np

from awkward_amd import contents as C

__all__ = ("lazy", "compute", "LazyArray", "Plan")

_ids = itertools.count(1)

_CMP = {"gt": ">", "ge": ">=", "lt": "<", "le": "<=", "eq": "==", "ne": "!="}
_NP_CMP = {"gt": np.greater, "ge": np.greater_equal, "lt": np.less, "le": np.less_equal, "eq": np.equal,
           "ne": np.not_equal}
_NP_BIN = {"add": np.add, "sub": np.subtract, "mul": np.multiply, "div": np.true_divide, "and": np.bitwise_and,
           "or": np.bitwise_or}
_NP_UN = {"abs": np.absolute, "neg": np.negative}


def _ufunc(op, table):
    """The numpy ufunc of a Binary / Unary node's op: the operator names
    above, or a numpy ufunc's own name (nodes recorded by __array_ufunc__)."""
    f = table.get(op)
    if f is None:
        f = getattr(np, op, None)
        if not isinstance(f, np.ufunc):
            raise NotImplementedError(f"lazy op {op!r}")
    return f



# ---- nodes ---------------------------------------------------------------------------
class Node:
    kind = "node"

    def __init__(self, *inputs, **params):
        self.id = next(_ids)
        self.inputs = inputs
        self.params = params

    def label(self):
        extra = ", ".join(f"{k}={v!r}" for k, v in self.params.items() if k != "array")
        return f"{self.kind}#{self.id}({extra})"


class Input(Node):
    kind = "input"


class Constant(Node):
    kind = "constant"


class Field(Node):  # GETITEM with a field name
    kind = "field"


class Filter(Node):  # GETITEM with a jagged boolean mask
    kind = "filter"


class Slice(Node):  # GETITEM with [:, start:stop:step]
    kind = "slice"


class Compare(Node):
    kind = "compare"


class Binary(Node):
    kind = "binary"


class Unary(Node):
    kind = "unary"


class Reduce(Node):
    kind = "reduce"


class Combinations(Node):
    kind = "combinations"


# ---- the lazy array ------------------------------------------------------------------------
def _node(value):
    if isinstance(value, LazyArray):
        return value.node
    if isinstance(value, (int, float, bool, np.integer, np.floating)):
        return Constant(value=value)
    return Input(array=value)


class LazyArray:
    """A node of the graph, with the reference LazyAwkwardArray's operators."""

    def __init__(self, node, context=None):
        self.node = node
        self.context = context if context is not None else {}

    def _wrap(self, node):
        return LazyArray(node, self.context)

    # element-wise and comparisons
    def _cmp(self, op, other):
        return self._wrap(Compare(self.node, _node(other), op=op))

    def __gt__(self, o):
        return self._cmp("gt", o)

    def __ge__(self, o):
        return self._cmp("ge", o)

    def __lt__(self, o):
        return self._cmp("lt", o)

    def __le__(self, o):
        return self._cmp("le", o)

    def __eq__(self, o):
        return self._cmp("eq", o)

    def __ne__(self, o):
        return self._cmp("ne", o)

    def _bin(self, op, other, reflected=False):
        a, b = (_node(other), self.node) if reflected else (self.node, _node(other))
        return self._wrap(Binary(a, b, op=op))

    def __add__(self, o):
        return self._bin("add", o)

    def __radd__(self, o):
        return self._bin("add", o, True)

    def __sub__(self, o):
        return self._bin("sub", o)

    def __rsub__(self, o):
        return self._bin("sub", o, True)

    def __mul__(self, o):
        return self._bin("mul", o)

    def __rmul__(self, o):
        return self._bin("mul", o, True)

    def __truediv__(self, o):
        return self._bin("div", o)

    def __and__(self, o):
        return self._bin("and", o)

    def __or__(self, o):
        return self._bin("or", o)

    def __abs__(self):
        return self._wrap(Unary(self.node, op="abs"))

    def __neg__(self):
        return self._wrap(Unary(self.node, op="neg"))

    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        """numpy ufuncs on lazy arrays record a node (np.cosh(lz.eta), ...)."""
        if method != "__call__" or kwargs or ufunc.nout != 1:
            return NotImplemented
        name = ufunc.__name__
        nodes = [_node(x) for x in inputs]
        if len(nodes) == 1:
            return self._wrap(Unary(nodes[0], op=name))
        if len(nodes) == 2:
            cmp = {v.__name__: k for k, v in _NP_CMP.items()}
            if name in cmp:
                return self._wrap(Compare(nodes[0], nodes[1], op=cmp[name]))
            return self._wrap(Binary(nodes[0], nodes[1], op=name))
        return NotImplemented

    __hash__ = None

    def __bool__(self):
        raise TypeError("a lazy array has no truth value before compute()")

    # list operations
    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        return self[name]

    def __getitem__(self, key):
        if isinstance(key, str):
            return self._wrap(Field(self.node, name=key))
        if isinstance(key, LazyArray):
            return self._wrap(Filter(self.node, key.node))
        if isinstance(key, tuple) and len(key) == 2 and key[0] == slice(None) and isinstance(key[1], slice):
            s = key[1]
            return self._wrap(Slice(self.node, start=s.start, stop=s.stop, step=s.step))
        raise NotImplementedError(f"lazy getitem {key!r}")

    def filter(self, condition):
        return self[condition]

    def combinations(self, n, replacement=False, fields=None):
        return self._wrap(Combinations(self.node, n=n, replacement=replacement, fields=fields))

    def _reduce(self, op):
        return self._wrap(Reduce(self.node, op=op))

    def sum(self):
        return self._reduce("sum")

    def max(self):
        return self._reduce("max")

    def min(self):
        return self._reduce("min")

    def argmax(self):
        return self._reduce("argmax")

    def argmin(self):
        return self._reduce("argmin")

    def count(self):
        return self._reduce("count")

    def num(self):
        return self._reduce("num")

    # execution
    def compute(self, fuse=True):
        return compute(self, fuse=fuse)[0]

    def compile(self):
        return Plan([self])

    def fusion_stats(self):
        return Plan([self]).stats()

    def invalidate(self):
        self.context.clear()

    def visualize(self):
        lines = []

        def walk(node, depth):
            lines.append("  " * depth + node.label())
            for child in node.inputs:
                walk(child, depth + 1)

        walk(self.node, 0)
        return "\n".join(lines)

    def __repr__(self):
        return f"<LazyArray {self.node.label()}>"


def lazy(array):
    """A lazy view of an Array (``_lazy_impl.py:357``)."""
    return LazyArray(Input(array=array))


# ---- eager evaluation --------------------------------------------------------------------
def _eager(node, memo):
    import awkward_amd as ak

    if node.id in memo:
        return memo[node.id]
    vals = [_eager(child, memo) for child in node.inputs]
    p = node.params
    if isinstance(node, Input):
        out = p["array"]
    elif isinstance(node, Constant):
        out = p["value"]
    elif isinstance(node, Field):
        out = vals[0][p["name"]]
    elif isinstance(node, Filter):
        out = vals[0][vals[1]]
    elif isinstance(node, Slice):
        out = vals[0][:, slice(p["start"], p["stop"], p["step"])]
    elif isinstance(node, Compare):
        out = _NP_CMP[p["op"]](vals[0], vals[1])
    elif isinstance(node, Binary):
        out = _ufunc(p["op"], _NP_BIN)(vals[0], vals[1])
    elif isinstance(node, Unary):
        out = _ufunc(p["op"], _NP_UN)(vals[0])
    elif isinstance(node, Reduce):
        out = getattr(ak, p["op"])(vals[0], axis=-1) if p["op"] != "num" else ak.num(vals[0], axis=1)
    elif isinstance(node, Combinations):
        out = ak.combinations(vals[0], p["n"], replacement=p["replacement"], fields=p["fields"])
    else:
        raise NotImplementedError(node.kind)
    memo[node.id] = out
    return out


# ---- fusion: pattern matching onto the event pipeline ----------------------------------------
def _terms(pred, source):
    """[(field, op, value)] if `pred` is a conjunction of comparisons of
    fields (or their absolute values) of `source` with constants."""
    if isinstance(pred, Binary) and pred.params["op"] == "and":
        a, b = _terms(pred.inputs[0], source), _terms(pred.inputs[1], source)
        return None if a is None or b is None else a + b
    if not isinstance(pred, Compare) or not isinstance(pred.inputs[1], Constant):
        return None
    lhs, op = pred.inputs[0], _CMP[pred.params["op"]]
    if isinstance(lhs, Unary) and lhs.params["op"] == "abs":
        lhs, op = lhs.inputs[0], "abs" + op
    if isinstance(lhs, Field) and lhs.inputs[0] is source:
        value = pred.inputs[1].params["value"]
        if isinstance(value, (bool, np.bool_)):
            return None
        return [(lhs.params["name"], op, float(value))]
    return None


def _match(node):
    """(input node, terms, slice, output name, reduce field) when `node` is a
    reduction the event pipeline computes, else None."""
    if not isinstance(node, Reduce) or node.params["op"] not in ("num", "count", "sum", "max", "argmax"):
        return None
    x = node.inputs[0]
    field = None
    if isinstance(x, Field):
        field = x.params["name"]
        x = x.inputs[0]
    sl = slice(None)
    if isinstance(x, Slice):
        sl = slice(x.params["start"], x.params["stop"], x.params["step"])
        if sl.step is not None and sl.step <= 0:
            return None
        x = x.inputs[0]
    if not isinstance(x, Filter) or not isinstance(x.inputs[0], Input):
        return None
    source = x.inputs[0]
    terms = _terms(x.inputs[1], source)
    if not terms:
        return None
    op = node.params["op"]
    if op in ("num", "count"):
        if op == "count" and field is None:
            return None
        return source, tuple(terms), sl, "num", field
    if field is None:
        return None
    return source, tuple(terms), sl, op, field


_FUSED_TYPES = (np.dtype(np.float32), np.dtype(np.float64), np.dtype(np.int8), np.dtype(np.int32),
                np.dtype(np.int64))


def _record_dtypes(array):
    """{field: dtype} of a hip list-of-records array the event pipeline can
    read, else None."""
    layout = array.layout
    if not isinstance(layout, C.ListOffsetArray) or not isinstance(layout.content, C.RecordArray):
        return None
    if not getattr(array.backend, "fused", False) or array.backend.nplike.name != "hip":
        return None
    rec = layout.content
    out = {}
    for name, col in zip(rec.field_names(), rec.contents):
        if isinstance(col, C.NumpyArray):
            out[name] = col.dtype
    return out


# ---- fusion: reductions over the pairs of COMBINATIONS(n=2) ---------------------------------
_PAIR_OPS = ("max", "min", "argmax", "argmin", "sum")


class _NoPair(Exception):
    pass


def _list_source(node):
    """The ListOffsetArray layout of a hip array (fused backend, int64
    offsets) that `node` -- an Input -- holds, else None."""
    if not isinstance(node, Input):
        return None
    array = node.params["array"]
    layout = getattr(array, "layout", None)
    be = getattr(array, "backend", None)
    if not isinstance(layout, C.ListOffsetArray) or not getattr(be, "fused", False) or be.nplike.name != "hip":
        return None
    if be.nplike.dtype(layout.offsets) != np.int64:
        return None
    return layout


def _pair_legs(comb):
    """The two leg names of a COMBINATIONS(n=2, replacement=False) node whose
    input the pair kernel reads, else None."""
    p = comb.params
    if p["n"] != 2 or p["replacement"] or _list_source(comb.inputs[0]) is None:
        return None
    return tuple(p["fields"]) if p["fields"] is not None else ("0", "1")


def _pair_combination(tree):
    """The one COMBINATIONS node whose legs the elementwise `tree` reads."""
    found = set()

    def walk(n):
        if isinstance(n, Combinations):
            found.add(n)
        elif isinstance(n, (Field, Binary, Unary, Compare)):
            for c in n.inputs:
                walk(c)
        elif not isinstance(n, Constant):
            raise _NoPair

    walk(tree)
    if len(found) != 1:
        raise _NoPair
    return found.pop()


def _pair_expr(node, comb, legs, content, nplike, memo):
    """An elementwise tree over the legs of `comb` as a DeviceExpr whose
    leaves read leg 0 or 1 of each pair (Python scalars stay scalars, so the
    types are the eager ufuncs' own); _NoPair for anything else."""
    import torch

    from awkward_amd import elementwise as E

    if node.id in memo:
        return memo[node.id]
    if isinstance(node, Constant):
        out = node.params["value"]
    elif isinstance(node, Field):
        x, name = node.inputs[0], node.params["name"]
        if x is comb and isinstance(content, C.NumpyArray):  # pairs["0"] of a list of numbers
            leg, col = name, content
        elif isinstance(x, Field) and x.inputs[0] is comb and isinstance(content, C.RecordArray):
            leg = x.params["name"]
            if name not in content.field_names():
                raise _NoPair
            col = content._getitem_field(name)
        else:
            raise _NoPair
        if leg not in legs or not isinstance(col, C.NumpyArray) or not isinstance(col.raw, torch.Tensor):
            raise _NoPair
        out = E.DeviceExpr(nplike, E.leaf(col.raw, leg=legs.index(leg)), 1)
    elif isinstance(node, (Binary, Unary, Compare)):
        args = [_pair_expr(c, comb, legs, content, nplike, memo) for c in node.inputs]
        if not any(isinstance(a, E.DeviceExpr) for a in args):
            raise _NoPair
        op = node.params["op"]
        uf = _NP_CMP[op] if isinstance(node, Compare) else _ufunc(op, _NP_BIN if isinstance(node, Binary) else _NP_UN)
        try:
            out = E.build(nplike, uf, args)
        except (NotImplementedError, TypeError, OverflowError):
            raise _NoPair from None
    else:
        raise _NoPair
    memo[node.id] = out
    return out


def _match_pair(node):
    """(COMBINATIONS node, expression node, op, DeviceExpr) when `node` is a
    reduction at axis=-1 of a float expression over the pairs, else None."""
    from awkward_amd import elementwise as E

    if not isinstance(node, Reduce) or node.params["op"] not in _PAIR_OPS:
        return None
    tree = node.inputs[0]
    try:
        comb = _pair_combination(tree)
        legs = _pair_legs(comb)
        if legs is None:
            return None
        layout = _list_source(comb.inputs[0])
        expr = _pair_expr(tree, comb, legs, layout.content, layout.backend.nplike, {})
    except _NoPair:
        return None
    if not isinstance(expr, E.DeviceExpr) or np.dtype(expr.dtype) not in (np.float32, np.float64):
        return None
    return comb, tree, node.params["op"], expr


class Plan:
    """The execution plan of one or more lazy outputs: fused event-pipeline
    launches for the reductions that match, one pair-kernel launch per
    expression over the pairs of COMBINATIONS(n=2) for the reductions of it
    (``awkward_hip_pair_reduce``), the carries of COMBINATIONS outputs from
    the offsets alone, and eager host-layer evaluation for the rest."""

    def __init__(self, outputs, fuse=True):
        self.outputs = list(outputs)
        self.groups = {}  # (source id, terms, slice, reduce field) -> group
        self.fused_nodes = {}
        self._pipelines = {}
        self.pair_groups = {}  # (combinations id, expression id) -> group
        self.combos = {}  # combinations id -> node
        self._pairs = {}
        self._combos = {}
        self.units = {}  # input id -> {"source", "ns": {n: [combinations keys]}, "pairs": key or None}
        self._unit_of = {}  # combinations / pair key -> input id
        self._runners = {}
        if not fuse:
            return
        f32 = np.dtype(np.float32)
        pending_num = []
        for out in self.outputs:
            m = _match(out.node)
            if m is None:
                self._add_other(out)
                continue
            source, terms, sl, name, field = m
            dtypes = _record_dtypes(source.params["array"])
            if dtypes is None or any(dtypes.get(t[0]) not in _FUSED_TYPES for t in terms):
                continue
            if name == "num":
                pending_num.append((out, source, terms, sl))
                continue
            if dtypes.get(field) != f32:  # the kernel reduces float32 values
                continue
            self._add(out, source, terms, sl, field, name)
        self._plan_units(fuse)
        for out, source, terms, sl in pending_num:
            base = (source.id, terms, (sl.start, sl.stop, sl.step))
            key = next((k for k in self.groups if k[:3] == base), None)
            if key is not None:
                self._add(out, source, terms, sl, key[3], "num")
                continue
            dtypes = _record_dtypes(source.params["array"])
            red = next((t[0] for t in terms if dtypes[t[0]] == f32), None)
            if red is None:
                red = next((f for f, dt in dtypes.items() if dt == f32), None)
            if red is not None:
                self._add(out, source, terms, sl, red, "num")

    def _add_other(self, out):
        node = out.node
        m = _match_pair(node)
        if m is not None:
            comb, tree, op, expr = m
            key = ("pairs", comb.id, tree.id)
            g = self.pair_groups.setdefault(key, {"comb": comb, "expr": expr, "ops": set(), "nodes": []})
            g["ops"].add(op)
            g["nodes"].append((node.id, op))
            self.fused_nodes[node.id] = key
        elif isinstance(node, Combinations) and _list_source(node.inputs[0]) is not None:
            key = ("combinations", node.id)
            self.combos[key] = node
            self.fused_nodes[node.id] = key

    def _plan_units(self, fuse):
        """COMBINATIONS outputs with n in {2, 3} (no replacement) on one input,
        and one pair reduction group over that input's pairs, become one
        awkward_hip_combinations_pairs launch."""
        if not fuse:
            return
        for key, node in self.combos.items():
            n = node.params["n"]
            if n not in (2, 3) or node.params["replacement"]:
                continue
            src = node.inputs[0]
            u = self.units.setdefault(src.id, {"source": src, "ns": {}, "pairs": None})
            u["ns"].setdefault(n, []).append(key)
            self._unit_of[key] = src.id
        for key, g in self.pair_groups.items():
            u = self.units.get(g["comb"].inputs[0].id)
            if u is not None and u["pairs"] is None:
                u["pairs"] = key
                self._unit_of[key] = g["comb"].inputs[0].id

    def _unit(self, uid):
        from awkward_amd.fused import CombinationsPairs

        r = self._runners.get(uid)
        if r is None:
            u = self.units[uid]
            expr, outputs = None, ()
            if u["pairs"] is not None:
                g = self.pair_groups[u["pairs"]]
                expr, outputs = g["expr"], tuple(o for o in _PAIR_OPS if o in g["ops"])
            r = CombinationsPairs(u["source"].params["array"], sorted(u["ns"]), expr, outputs)
            self._runners[uid] = r
        return r

    def _pair_reduce(self, key):
        from awkward_amd.fused import PairReduce

        if key in self._unit_of:
            return self._unit(self._unit_of[key])
        p = self._pairs.get(key)
        if p is None:
            g = self.pair_groups[key]
            offsets = _list_source(g["comb"].inputs[0]).offsets
            p = PairReduce(offsets, g["expr"], outputs=tuple(o for o in _PAIR_OPS if o in g["ops"]))
            self._pairs[key] = p
        return p

    def _combinations(self, key):
        from awkward_amd.fused import Combinations as Decode

        if key in self._unit_of:
            return self._unit(self._unit_of[key]).tuples[self.combos[key].params["n"]]
        d = self._combos.get(key)
        if d is None:
            node = self.combos[key]
            d = Decode(node.inputs[0].params["array"], node.params["n"], node.params["replacement"])
            self._combos[key] = d
        return d

    def _combinations_array(self, key):
        from awkward_amd.highlevel import Array

        node, d = self.combos[key], self._combinations(key)
        layout = _list_source(node.inputs[0])
        be = layout.backend
        contents = [layout.content._carry(c[:d.total], True) for c in d.carry]
        record = C.RecordArray(contents, node.params["fields"], d.total, be)
        return Array(C.ListOffsetArray(d.tooffsets, record, be))

    def _add(self, out, source, terms, sl, red, name):
        key = (source.id, terms, (sl.start, sl.stop, sl.step), red)
        g = self.groups.setdefault(key, {"source": source, "terms": terms, "slice": sl, "field": red,
                                         "outputs": set(), "nodes": []})
        g["outputs"].add(name)
        g["nodes"].append((out.node.id, name))
        self.fused_nodes[out.node.id] = key

    def _pipeline(self, key):
        from awkward_amd.fused import EventPipeline

        p = self._pipelines.get(key)
        if p is None:
            g = self.groups[key]
            outputs = tuple(o for o in ("num", "sum", "argmax", "max") if o in g["outputs"])
            p = EventPipeline(g["source"].params["array"], list(g["terms"]), g["slice"], g["field"], outputs=outputs)
            self._pipelines[key] = p
        return p

    def run(self):
        """The outputs' values, in order."""
        from awkward_amd.highlevel import Array

        memo = {}
        results = {}
        for key in self.groups:
            p = self._pipeline(key)
            o = p.run()
            be = self.groups[key]["source"].params["array"].backend
            results[key] = {
                "num": lambda o=o, be=be: Array(C.NumpyArray(o["num"], be)),
                "sum": lambda o=o, be=be: Array(C.NumpyArray(o["sum"], be)),
                "argmax": lambda o=o, be=be: Array(C.ByteMaskedArray(o["argmax_mask"], C.NumpyArray(o["argmax"], be),
                                                                     False, be)),
                "max": lambda o=o, be=be: Array(C.ByteMaskedArray(o["max_mask"], C.NumpyArray(o["max"], be), False,
                                                                  be)),
            }
        for uid in self.units:
            self._unit(uid).run()
        for key in self.pair_groups:
            r = self._pair_reduce(key)
            if key not in self._unit_of:
                r.run()
            o = r.out
            be = _list_source(self.pair_groups[key]["comb"].inputs[0]).backend
            masked = {k: lambda o=o, be=be, k=k: Array(C.ByteMaskedArray(o["mask"], C.NumpyArray(o[k], be), False, be))
                      for k in ("max", "min", "argmax", "argmin")}
            results[key] = dict(masked, sum=lambda o=o, be=be: Array(C.NumpyArray(o["sum"], be)))
        for key in self.combos:
            if key not in self._unit_of:
                self._combinations(key).run()
        out = []
        for lz in self.outputs:
            key = self.fused_nodes.get(lz.node.id)
            if key is None:
                out.append(_eager(lz.node, memo))
            elif key in self.combos:
                out.append(self._combinations_array(key))
            elif key in self.pair_groups:
                name = dict(self.pair_groups[key]["nodes"])[lz.node.id]
                out.append(results[key][name]())
            else:
                name = dict(self.groups[key]["nodes"])[lz.node.id]
                out.append(results[key][name]())
        return out

    def graph(self):
        """The fused launches captured into a hipGraph (fused.Graph): replay()
        reruns them on the same inputs; the outputs' buffers are reused."""
        from awkward_amd.fused import Graph

        from awkward_amd.fused import Concurrent

        if not self.fused_nodes or any(lz.node.id not in self.fused_nodes for lz in self.outputs):
            raise ValueError("only plans whose every output is fused can be captured")
        reductions = [self._pipeline(key) for key in self.groups] + \
            [self._pair_reduce(k) for k in self.pair_groups if k not in self._unit_of] + \
            [self._unit(uid) for uid in self.units]
        decodes = [self._combinations(key) for key in self.combos if key not in self._unit_of]
        if reductions and decodes:
            # the reductions (latency-bound) and the carries (write-bound)
            # are independent: two branches of the graph
            return Graph(Concurrent(lambda: [d.run() for d in decodes], lambda: [p.run() for p in reductions]))
        return Graph(lambda: [p.run() for p in reductions + decodes])

    def stats(self):
        """Counts in the spirit of the reference's fusion_stats: graph nodes,
        reductions fused, the kernel launches they become, and nodes left to
        eager evaluation."""
        seen = set()

        def walk(node):
            if node.id in seen:
                return
            seen.add(node.id)
            for child in node.inputs:
                walk(child)

        for lz in self.outputs:
            walk(lz.node)
        fused = len(self.fused_nodes)
        in_units = sum(1 for k in list(self.pair_groups) + list(self.combos) if k in self._unit_of)
        launches = len(self.groups) + len(self.pair_groups) + len(self.combos) - in_units + len(self.units)
        return {"nodes": len(seen), "fused_reductions": fused - len(self.combos), "fused_launches": launches,
                "eager_outputs": len(self.outputs) - fused}


def compute(*outputs, fuse=True):
    """Evaluates lazy outputs together: reductions that share one fused
    launch are computed by it once (``_lazy_impl.py`` ``compute``)."""
    return Plan(outputs, fuse=fuse).run()
