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
        out = _NP_BIN[p["op"]](vals[0], vals[1])
    elif isinstance(node, Unary):
        out = abs(vals[0]) if p["op"] == "abs" else -vals[0]
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


class Plan:
    """The execution plan of one or more lazy outputs: fused event-pipeline
    launches for the reductions that match, eager host-layer evaluation for
    the rest."""

    def __init__(self, outputs, fuse=True):
        self.outputs = list(outputs)
        self.groups = {}  # (source id, terms, slice, reduce field) -> group
        self.fused_nodes = {}
        self._pipelines = {}
        if not fuse:
            return
        f32 = np.dtype(np.float32)
        pending_num = []
        for out in self.outputs:
            m = _match(out.node)
            if m is None:
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
        out = []
        for lz in self.outputs:
            key = self.fused_nodes.get(lz.node.id)
            if key is None:
                out.append(_eager(lz.node, memo))
            else:
                name = dict(self.groups[key]["nodes"])[lz.node.id]
                out.append(results[key][name]())
        return out

    def graph(self):
        """The fused launches captured into a hipGraph (fused.Graph): replay()
        reruns them on the same inputs; the outputs' buffers are reused."""
        from awkward_amd.fused import Graph

        if not self.groups or any(lz.node.id not in self.fused_nodes for lz in self.outputs):
            raise ValueError("only plans whose every output is fused can be captured")
        pipes = [self._pipeline(key) for key in self.groups]
        return Graph(lambda: [p.run() for p in pipes])

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
        return {"nodes": len(seen), "fused_reductions": fused, "fused_launches": len(self.groups),
                "eager_outputs": len(self.outputs) - fused}


def compute(*outputs, fuse=True):
    """Evaluates lazy outputs together: reductions that share one fused
    launch are computed by it once (``_lazy_impl.py`` ``compute``)."""
    return Plan(outputs, fuse=fuse).run()
