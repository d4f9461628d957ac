"""Layout nodes of the jagged-array path and the kernel call sequences they
issue, following the reference's layout code:

* ``ListOffsetArray`` -- ``src/awkward/contents/listoffsetarray.py``
  (``_reduce_next`` local branch :1394-1430, ``_combinations`` :1152-1261,
  ``_argsort_next`` / ``_sort_next`` local branches, ``_local_index`` :664-697,
  ``_carry`` :380-391);
* ``ListArray`` -- ``contents/listarray.py`` (``_getitem_next_jagged`` NumpyArray
  branch :549-618, range slice :782-870, ``_compact_offsets64`` :409-433);
* ``NumpyArray`` -- ``contents/numpyarray.py`` (``_carry`` :367-373,
  ``_reduce_next`` :1046-1114, ``_argsort_next`` :904-991, ``_sort_next`` :993-1032);
* ``RecordArray`` -- ``contents/recordarray.py`` (lazy ``_carry`` :552-591);
* ``IndexedArray`` -- ``contents/indexedarray.py`` (``_reduce_next`` :802-825);
* ``ByteMaskedArray`` -- the masked output of max/min/argmax/argmin.

Each node holds its backend; every buffer is an array of that backend's
nplike, and every data-parallel step is a registry kernel or an nplike
operation of the backend.  Only the axis=-1 paths of the configs are
implemented; everything else raises NotImplementedError.
"""

from __future__ import annotations

import numpy as np

builtins_min = min

from awkward_amd import reducers as _reducers


class Content:
    backend = None

    @property
    def nplike(self):
        return self.backend.nplike

    def _kernel(self, *key):
        return self.backend[key]

    def _check(self, err):
        self.backend.maybe_kernel_error(err)

    def to_list(self):
        raise NotImplementedError

    @property
    def is_list(self):
        return False

    @property
    def depth(self):
        """List nesting depth counted from the numbers (NumpyArray: 1),
        Content.branch_depth's depth for a non-branching layout."""
        return self.content.depth

    def _reduce_axis(self, reducer, negaxis, starts, shifts, offsets, outlength, mask, keepdims):
        """Content._reduce_next of the reference for any negaxis: the
        numbers' level is the local reduction (negaxis 1)."""
        if negaxis != 1:
            raise NotImplementedError(f"a reduction {negaxis} levels above the numbers through a "
                                      f"{type(self).__name__}")
        return self._reduce_next(reducer, starts, offsets, outlength, mask, keepdims, shifts)


# ---- leaves ---------------------------------------------------------------------
class NumpyArray(Content):
    """1-D numbers.  On a fusing nplike the buffer may be a lazy
    elementwise expression (``raw``); ``data`` writes it out on first use."""

    def __init__(self, data, backend):
        self._data = data
        self.backend = backend

    @property
    def data(self):
        d = self._data
        if hasattr(d, "materialize"):
            d = self._data = d.materialize()
        return d

    @property
    def raw(self):
        return self._data

    @property
    def depth(self):
        return 1

    @property
    def length(self):
        return self.nplike.length(self._data)

    @property
    def dtype(self):
        return self.nplike.dtype(self._data)

    def __repr__(self):
        return f"NumpyArray(len={self.length}, dtype={self.dtype})"

    def slice(self, start, stop):
        return NumpyArray(self.data[start:stop], self.backend)

    def _carry(self, carry, allow_lazy=True):
        return NumpyArray(self.nplike.take(self.data, carry), self.backend)

    def project(self):
        return self

    def _reduce_next(self, reducer, starts, offsets, outlength, mask, keepdims, shifts=None):
        out, outmask = reducer.apply(self, offsets, starts, shifts, outlength, mask)
        if mask and outmask is None:
            outmask = self.nplike.empty(outlength, np.int8)
            self._check(
                self._kernel("awkward_NumpyArray_reduce_mask_ByteMaskedArray_64", np.int8, np.int64)(
                    outmask, offsets, outlength
                )
            )
        if mask:
            out = ByteMaskedArray(outmask, out, False, self.backend)
        if keepdims:
            out = RegularArray(out, 1, outlength, self.backend)
        return out

    def _argsort_next(self, offsets, ascending, stable):
        nextcarry = self.nplike.empty(self.length, np.int64)
        self._check(
            self._kernel("awkward_argsort", np.int64, self.dtype, self.nplike.dtype(offsets))(
                nextcarry, self.data, self.length, offsets, self.nplike.length(offsets), ascending, stable
            )
        )
        return NumpyArray(nextcarry, self.backend)

    def _sort_next(self, offsets, ascending, stable):
        out = self.nplike.empty(self.length, self.dtype)
        self._check(
            self._kernel("awkward_sort", self.dtype, self.dtype, self.nplike.dtype(offsets))(
                out, self.data, self.length, offsets, self.nplike.length(offsets), ascending, stable
            )
        )
        return NumpyArray(out, self.backend)

    def to_list(self):
        return self.nplike.to_numpy(self.data).tolist()


class RecordArray(Content):
    @property
    def depth(self):
        """The fields' list depth (their minimum, as purelist_depth takes it)."""
        return builtins_min((c.depth for c in self.contents), default=1)

    def _reduce_axis(self, reducer, negaxis, starts, shifts, offsets, outlength, mask, keepdims):
        raise NotImplementedError("reductions of records are not on the hip path (reduce a field)")

    def _reduce_next(self, reducer, starts, offsets, outlength, mask, keepdims, shifts=None):
        raise NotImplementedError("reductions of records are not on the hip path (reduce a field)")

    def __init__(self, contents, fields, length, backend):
        self.contents = list(contents)
        self.fields = None if fields is None else list(fields)
        self._length = length
        self.backend = backend

    @property
    def length(self):
        return self._length

    def __repr__(self):
        return f"RecordArray(fields={self.fields}, len={self.length})"

    def field_names(self):
        return self.fields if self.fields is not None else [str(i) for i in range(len(self.contents))]

    def _getitem_field(self, where):
        names = self.field_names()
        return self.contents[names.index(str(where))]

    def slice(self, start, stop):
        return RecordArray([c.slice(start, stop) for c in self.contents], self.fields,
                           max(0, min(stop, self.length) - start), self.backend)

    def _carry(self, carry, allow_lazy=True):
        if allow_lazy:
            return IndexedArray(carry, self, self.backend)
        return RecordArray([c._carry(carry, False) for c in self.contents], self.fields,
                           self.nplike.length(carry), self.backend)

    def project(self):
        return self

    def to_list(self):
        cols = [c.to_list() for c in self.contents]
        names = self.field_names()
        if self.fields is None:
            return [tuple(col[i] for col in cols) for i in range(self.length)]
        return [{k: col[i] for k, col in zip(names, cols)} for i in range(self.length)]


class IndexedArray(Content):
    """A lazy carry: element i is content[index[i]].  Buffers are immutable,
    so a field's lazy carry and a projection are made once per node and
    reused: ak.sum(sel.pt) then ak.argmax(sel.pt) gather sel.pt once."""

    def __init__(self, index, content, backend):
        self.index = index
        self.content = content
        self.backend = backend
        self._fields = {}
        self._projected = None

    @property
    def length(self):
        return self.nplike.length(self.index)

    def __repr__(self):
        return f"IndexedArray(len={self.length}, {self.content!r})"

    def slice(self, start, stop):
        if start == 0 and stop == self.length:
            return self
        return IndexedArray(self.index[start:stop], self.content, self.backend)

    def _getitem_field(self, where):
        field = self._fields.get(where)
        if field is None:
            field = self._fields[where] = IndexedArray(self.index, self.content._getitem_field(where), self.backend)
        return field

    def _carry(self, carry, allow_lazy=True):
        return IndexedArray(self.nplike.take(self.index, carry), self.content, self.backend)

    def project(self):
        if self._projected is None:
            self._projected = self.content._carry(self.index, False)
        return self._projected

    def _reduce_next(self, reducer, starts, offsets, outlength, mask, keepdims, shifts=None):
        return self.project()._reduce_next(reducer, starts, offsets, outlength, mask, keepdims, shifts)

    def _argsort_next(self, offsets, ascending, stable):
        return self.project()._argsort_next(offsets, ascending, stable)

    def _sort_next(self, offsets, ascending, stable):
        return self.project()._sort_next(offsets, ascending, stable)

    @property
    def dtype(self):
        return self.content.dtype

    @property
    def data(self):
        return self.project().data

    def to_list(self):
        return self.project().to_list()


def _option_reduce_next(layout, reducer, starts, offsets, outlength, mask, keepdims, shifts, numnull_key,
                        reduce_key, shift_keys, args):
    """The local (axis=-1) branch of ByteMaskedArray._reduce_next
    (contents/bytemaskedarray.py:848-975) and IndexedOptionArray._reduce_next
    (contents/indexedoptionarray.py:1394-1431 with _rearrange_prepare_next
    :1177-1227 and _rearrange_nextshifts :1136-1175): count the nulls, compact
    the valid elements of each bin into the next carry and offsets, and for
    positional reducers pass the nulls before each valid element down as
    shifts, so argmin/argmax count them in their local index."""
    nplike = layout.nplike
    n = layout.length
    numnull = nplike.empty(1, np.int64)
    layout._check(layout._kernel(*numnull_key)(numnull, *args["numnull"]))
    next_length = n - nplike.item(numnull, 0)
    nextcarry = nplike.empty(next_length, np.int64)
    nextoffsets = nplike.empty(outlength + 1, np.int64)
    outindex = nplike.empty(n, np.int64)
    layout._check(layout._kernel(*reduce_key)(nextcarry, nextoffsets, outindex, *args["reduce"](offsets, outlength)))
    nextshifts = None
    if reducer.needs_position:
        nextshifts = nplike.empty(next_length, np.int64)
        if shifts is None:
            layout._check(layout._kernel(*shift_keys[0])(nextshifts, *args["shifts"]))
        else:
            layout._check(layout._kernel(*shift_keys[1])(nextshifts, *args["shifts"], shifts))
    nxt = layout.content._carry(nextcarry, False)
    return nxt._reduce_next(reducer, starts, nextoffsets, outlength, mask, keepdims, nextshifts)


class ByteMaskedArray(Content):
    def __init__(self, mask, content, valid_when, backend):
        self.mask = mask
        self.content = content
        self.valid_when = valid_when
        self.backend = backend

    @property
    def length(self):
        return self.nplike.length(self.mask)

    def __repr__(self):
        return f"ByteMaskedArray(valid_when={self.valid_when}, {self.content!r})"

    def slice(self, start, stop):
        return ByteMaskedArray(self.mask[start:stop], self.content.slice(start, stop), self.valid_when, self.backend)

    def _carry(self, carry, allow_lazy=True):
        return ByteMaskedArray(self.nplike.take(self.mask, carry), self.content._carry(carry, allow_lazy),
                               self.valid_when, self.backend)

    def _project_valid(self):
        """(carry of the valid elements, outindex with -1 at the nulls):
        awkward_ByteMaskedArray_getitem_nextcarry_outindex_64."""
        nplike = self.nplike
        m = self._mask8()
        n = self.length
        numnull = nplike.empty(1, np.int64)
        self._check(self._kernel("awkward_ByteMaskedArray_numnull", np.int64, np.int8)(numnull, m, n,
                                                                                      bool(self.valid_when)))
        nextcarry = nplike.empty(n - nplike.item(numnull, 0), np.int64)
        outindex = nplike.empty(n, np.int64)
        self._check(self._kernel("awkward_ByteMaskedArray_getitem_nextcarry_outindex", np.int64, np.int64, np.int8)(
            nextcarry, outindex, m, n, bool(self.valid_when)))
        return nextcarry, outindex

    def _mask8(self):
        m = self.mask
        return self.nplike.view(m, np.int8) if self.nplike.dtype(m) == np.dtype(np.bool_) else m

    def _reduce_next(self, reducer, starts, offsets, outlength, mask, keepdims, shifts=None):
        m = self._mask8()
        n = self.length
        vw = bool(self.valid_when)
        return _option_reduce_next(
            self, reducer, starts, offsets, outlength, mask, keepdims, shifts,
            ("awkward_ByteMaskedArray_numnull", np.int64, np.int8),
            ("awkward_ByteMaskedArray_reduce_next_64", np.int64, np.int64, np.int64, np.int8, np.int64),
            (("awkward_ByteMaskedArray_reduce_next_nonlocal_nextshifts_64", np.int64, np.int8),
             ("awkward_ByteMaskedArray_reduce_next_nonlocal_nextshifts_fromshifts_64", np.int64, np.int8, np.int64)),
            {"numnull": (m, n, vw), "reduce": lambda off, ol: (m, off, ol, vw), "shifts": (m, n, vw)},
        )

    def to_list(self):
        mask = self.nplike.to_numpy(self.mask).astype(bool)
        values = self.content.to_list()
        valid = mask if self.valid_when else ~mask
        return [v if ok else None for v, ok in zip(values, valid)]


class IndexedOptionArray(Content):
    """Option type by index: element i is content[index[i]], or None where
    index[i] < 0 (contents/indexedoptionarray.py)."""

    def __init__(self, index, content, backend):
        self.index = index
        self.content = content
        self.backend = backend

    @property
    def length(self):
        return self.nplike.length(self.index)

    def __repr__(self):
        return f"IndexedOptionArray(len={self.length}, {self.content!r})"

    def slice(self, start, stop):
        return IndexedOptionArray(self.index[start:stop], self.content, self.backend)

    def _carry(self, carry, allow_lazy=True):
        return IndexedOptionArray(self.nplike.take(self.index, carry), self.content, self.backend)

    def _reduce_next(self, reducer, starts, offsets, outlength, mask, keepdims, shifts=None):
        ix = self.index
        it = self.nplike.dtype(ix)
        n = self.length
        return _option_reduce_next(
            self, reducer, starts, offsets, outlength, mask, keepdims, shifts,
            ("awkward_IndexedArray_numnull", np.int64, it),
            ("awkward_IndexedArray_reduce_next_64", np.int64, np.int64, np.int64, it, np.int64),
            (("awkward_IndexedArray_reduce_next_nonlocal_nextshifts_64", np.int64, it),
             ("awkward_IndexedArray_reduce_next_nonlocal_nextshifts_fromshifts_64", np.int64, it, np.int64)),
            {"numnull": (ix, n), "reduce": lambda off, ol: (ix, off, ol), "shifts": (ix, n)},
        )

    def _project_valid(self):
        """(carry of the valid elements, outindex with -1 at the nulls):
        awkward_IndexedArray_getitem_nextcarry_outindex_64 (an index past the
        content is an IndexError, as in the reference)."""
        nplike = self.nplike
        ix = self.index
        it = nplike.dtype(ix)
        n = self.length
        numnull = nplike.empty(1, np.int64)
        self._check(self._kernel("awkward_IndexedArray_numnull", np.int64, it)(numnull, ix, n))
        nextcarry = nplike.empty(n - nplike.item(numnull, 0), np.int64)
        outindex = nplike.empty(n, it)
        self._check(self._kernel("awkward_IndexedArray_getitem_nextcarry_outindex", np.int64, it, it)(
            nextcarry, outindex, ix, n, self.content.length))
        return nextcarry, outindex

    def to_list(self):
        index = self.nplike.to_numpy(self.index).astype(np.int64)
        values = self.content.to_list()
        return [values[i] if i >= 0 else None for i in index]


class RegularArray(Content):
    def __init__(self, content, size, length, backend):
        self.content = content
        self.size = size
        self._length = length
        self.backend = backend

    @property
    def depth(self):
        return 1 + self.content.depth

    @property
    def is_list(self):
        return True

    def slice(self, start, stop):
        return RegularArray(self.content.slice(start * self.size, stop * self.size), self.size, stop - start,
                            self.backend)

    def to_ListOffsetArray64(self):
        offsets = self.nplike.arange(self.length + 1, np.int64) * self.size
        return ListOffsetArray(offsets, self.content.slice(0, self.length * self.size), self.backend)

    def _reduce_axis(self, reducer, negaxis, starts, shifts, offsets, outlength, mask, keepdims):
        return self.to_ListOffsetArray64()._reduce_axis(reducer, negaxis, starts, shifts, offsets, outlength, mask,
                                                        keepdims)

    @property
    def length(self):
        return self._length

    def to_list(self):
        flat = self.content.to_list()
        return [flat[i * self.size:(i + 1) * self.size] for i in range(self.length)]


# ---- lists ------------------------------------------------------------------------
class ListOffsetArray(Content):
    def __init__(self, offsets, content, backend, exact=False):
        self.offsets = offsets
        self.content = content
        self.backend = backend
        # exact: offsets made by this library's own scans -- non-decreasing,
        # from 0 to the content's length -- so no list selects more than the
        # content holds (a range slice's carry fits in content.length)
        self.exact = exact

    @property
    def is_list(self):
        return True

    @property
    def length(self):
        return self.nplike.length(self.offsets) - 1

    @property
    def starts(self):
        return self.offsets[:-1]

    @property
    def stops(self):
        return self.offsets[1:]

    def __repr__(self):
        return f"ListOffsetArray(len={self.length}, {self.content!r})"

    def _getitem_field(self, where):
        return ListOffsetArray(self.offsets, self.content._getitem_field(where), self.backend, self.exact)

    def _carry(self, carry, allow_lazy=True):
        return ListArray(self.nplike.take(self.starts.contiguous() if hasattr(self.starts, "contiguous") else self.starts, carry),
                         self.nplike.take(self.stops.contiguous() if hasattr(self.stops, "contiguous") else self.stops, carry),
                         self.content, self.backend)

    def to_ListArray(self):
        return ListArray(self.starts, self.stops, self.content, self.backend)

    def _zero_based(self):
        """(offsets starting at 0, content trimmed to them): the reference's
        to_ListOffsetArray64(True) / trim step (one host read of offsets[0])."""
        nplike = self.nplike
        offsets = self.offsets
        if nplike.dtype(offsets) != np.dtype(np.int64):
            offsets = nplike.asarray(offsets, np.int64)
        if self.exact:  # from 0 to the content's length: nothing to read or trim
            return offsets, self.content
        first = nplike.item(offsets, 0)
        last = nplike.item(offsets, nplike.length(offsets) - 1)
        if first != 0:
            offsets = offsets - first
        return offsets, self.content.slice(first, last)

    @property
    def depth(self):
        return 1 + self.content.depth

    def to_ListOffsetArray64(self):
        return self

    def slice(self, start, stop):
        return ListOffsetArray(self.offsets[start:stop + 1], self.content, self.backend)

    # -- reduce at any axis: listoffsetarray.py:1263-1431 --
    def _reduce_axis(self, reducer, negaxis, starts, shifts, offsets, outlength, mask, keepdims):
        nplike = self.nplike
        selfoffsets, trimmed = self._zero_based()
        depth = self.depth
        if negaxis == depth:
            # non-local: this level's lists are regrouped by column
            (distincts, maxcount, maxnextparents, nextcarry, nextoffsets,
             nextstarts) = self._rearrange_prepare_next(selfoffsets, trimmed, outlength, offsets)
            outstarts = nplike.empty(outlength, np.int64)
            outstops = nplike.empty(outlength, np.int64)
            self._check(self._kernel("awkward_ListOffsetArray_reduce_nonlocal_outstartsstops_64", np.int64, np.int64,
                                     np.int64)(outstarts, outstops, distincts, nplike.length(distincts), outlength))
            nextshifts = None
            if reducer.needs_position:
                nextlen = nplike.length(nextcarry)
                nextshifts = nplike.empty(nextlen, np.int64)
                nummissing = nplike.empty(maxcount, np.int64)
                missing = nplike.empty(nplike.item(selfoffsets, self.length), np.int64)
                self._check(self._kernel("awkward_ListOffsetArray_reduce_nonlocal_nextshifts_64", np.int64, np.int64,
                                         np.int64, np.int64, np.int64, np.int64, np.int64)(
                    nummissing, missing, nextshifts, selfoffsets, self.length, starts, offsets, outlength, maxcount,
                    nextlen, nextcarry))
            nextcontent = trimmed._carry(nextcarry, False)
            outcontent = nextcontent._reduce_axis(reducer, negaxis - 1, nextstarts, nextshifts, nextoffsets,
                                                  maxnextparents + 1, mask, False)
            out = ListArray(outstarts, outstops, outcontent, self.backend)
            if keepdims:
                out = RegularArray(out, 1, self.length, self.backend)
            return out
        # local: the reduction is below this level
        outcontent = trimmed._reduce_axis(reducer, negaxis, selfoffsets[:-1], shifts, selfoffsets, self.length, mask,
                                          keepdims)
        if not (keepdims and depth == negaxis + 1) and depth >= negaxis + 2:
            outcontent = outcontent.to_ListOffsetArray64()
        return ListOffsetArray(offsets, outcontent, self.backend)

    def _rearrange_prepare_next(self, selfoffsets, trimmed, outlength, offsets):
        """listoffsetarray.py:1432-1526 on the zero-based lists."""
        nplike = self.nplike
        nextlen = builtins_min(nplike.item(selfoffsets, self.length), trimmed.length)
        maxcount_ = nplike.empty(1, np.int64)
        offsetscopy = nplike.empty(self.length + 1, np.int64)
        self._check(self._kernel("awkward_ListOffsetArray_reduce_nonlocal_maxcount_offsetscopy_64", np.int64, np.int64,
                                 np.int64)(maxcount_, offsetscopy, selfoffsets, self.length))
        maxcount = nplike.item(maxcount_, 0)
        nextcarry = nplike.empty(nextlen, np.int64)
        maxnextparents_ = nplike.empty(1, np.int64)
        distincts = nplike.empty(outlength * maxcount, np.int64)
        nextoffsets = nplike.empty(outlength * maxcount + 1, np.int64)
        self._check(self._kernel("awkward_ListOffsetArray_reduce_nonlocal_preparenext_64", np.int64, np.int64,
                                 np.int64, np.int64, np.int64, np.int64, np.int64)(
            nextcarry, nextoffsets, nextlen, maxnextparents_, distincts, outlength * maxcount, offsetscopy,
            selfoffsets, self.length, offsets, outlength, maxcount))
        maxnextparents = nplike.item(maxnextparents_, 0)
        nextoffsets = nextoffsets[: maxnextparents + 2]
        return distincts, maxcount, maxnextparents, nextcarry, nextoffsets, nextoffsets[:-1]

    # -- reduce, axis=-1: listoffsetarray.py:1394-1430 --
    def _reduce_local(self, reducer, mask, keepdims):
        nextoffsets, trimmed = self._zero_based()
        nextstarts = nextoffsets[:-1]
        return trimmed._reduce_next(reducer, nextstarts, nextoffsets, self.length, mask, keepdims)

    # -- argsort / sort, axis=-1 --
    def _argsort_local(self, ascending, stable):
        nextoffsets, trimmed = self._zero_based()
        out = trimmed._argsort_next(nextoffsets, ascending, stable)
        return ListOffsetArray(nextoffsets, out, self.backend)

    def _sort_local(self, ascending, stable):
        nextoffsets, trimmed = self._zero_based()
        out = trimmed._sort_next(nextoffsets, ascending, stable)
        return ListOffsetArray(nextoffsets, out, self.backend)

    # -- local index: listoffsetarray.py:664-697 --
    def _local_index(self):
        nplike = self.nplike
        offsets = self.offsets
        if nplike.dtype(offsets) != np.dtype(np.int64):
            offsets = nplike.asarray(offsets, np.int64)
        innerlength = nplike.item(offsets, nplike.length(offsets) - 1)
        localindex = nplike.empty(innerlength, np.int64)
        self._check(
            self._kernel("awkward_ListArray_localindex", np.int64, np.int64)(localindex, offsets, self.length)
        )
        return ListOffsetArray(offsets, NumpyArray(localindex, self.backend), self.backend)

    # -- combinations, axis=1: listoffsetarray.py:1152-1261 --
    def _combinations(self, n, replacement, fields):
        nplike = self.nplike
        starts, stops = self.starts, self.stops
        if hasattr(starts, "contiguous"):
            starts, stops = starts.contiguous(), stops.contiguous()
        totallen_buf = nplike.empty(1, np.int64)
        offsets = nplike.empty(self.length + 1, np.int64)
        cdt = nplike.dtype(starts)
        self._check(
            self._kernel("awkward_ListArray_combinations_length", np.int64, np.int64, cdt, cdt)(
                totallen_buf, offsets, n, replacement, starts, stops, self.length
            )
        )
        totallen = nplike.index_as_shape_item(totallen_buf)
        tocarry = [nplike.empty(totallen, np.int64) for _ in range(n)]
        toindex = nplike.empty(n, np.int64)
        fromindex = nplike.empty(n, np.int64)
        self._check(
            self._kernel("awkward_ListArray_combinations", np.int64, np.int64, np.int64, cdt, cdt)(
                self.backend.pointer_array(tocarry), toindex, fromindex, n, replacement, starts, stops, self.length
            )
        )
        contents = [self.content._carry(ptr, True) for ptr in tocarry]
        record = RecordArray(contents, fields, totallen, self.backend)
        return ListOffsetArray(offsets, record, self.backend)

    def _getitem_range(self, start, stop, step):
        return self.to_ListArray()._getitem_range(start, stop, step, self.content.length if self.exact else None)

    def _getitem_jagged(self, item):
        return self.to_ListArray()._getitem_jagged(item)

    def to_list(self):
        offs = self.nplike.to_numpy(self.offsets)
        flat = self.content.to_list()
        return [flat[offs[i]:offs[i + 1]] for i in range(len(offs) - 1)]


class ListArray(Content):
    def __init__(self, starts, stops, content, backend):
        self.starts = starts
        self.stops = stops
        self.content = content
        self.backend = backend

    @property
    def is_list(self):
        return True

    @property
    def length(self):
        return self.nplike.length(self.starts)

    def __repr__(self):
        return f"ListArray(len={self.length}, {self.content!r})"

    def _contig(self):
        s, e = self.starts, self.stops
        if hasattr(s, "contiguous"):
            s, e = s.contiguous(), e.contiguous()
        return s, e

    def _getitem_field(self, where):
        return ListArray(self.starts, self.stops, self.content._getitem_field(where), self.backend)

    # -- [:, start:stop:step]: listarray.py:782-870 --
    def _getitem_range(self, start, stop, step, bound=None):
        """bound: an upper bound of the carry's length known without a
        device read (lists of exact offsets: their content's length).  With
        it, and a nplike that fuses, the carry is allocated to the bound and
        trimmed to next_range's own total, so the carrylength pass (and its
        host read) is skipped; otherwise the reference's two kernels run."""
        nplike = self.nplike
        step = 1 if step is None else step
        if step == 0:
            raise ValueError("slice step cannot be zero")
        kstart = _kSliceNone if start is None else start
        kstop = _kSliceNone if stop is None else stop
        starts, stops = self._contig()
        cdt = nplike.dtype(starts)
        lenstarts = self.length
        bounded = bound is not None and getattr(nplike, "fuse", False)
        if bounded:
            nextcarry = nplike.empty(max(bound, 1), np.int64)
        else:
            carrylength = nplike.empty(1, np.int64)
            self._check(
                self._kernel("awkward_ListArray_getitem_next_range_carrylength", np.int64, cdt, cdt)(
                    carrylength, starts, stops, lenstarts, kstart, kstop, step
                )
            )
            nextcarry = nplike.empty(nplike.index_as_shape_item(carrylength), np.int64)
        nextoffsets = nplike.empty(lenstarts + 1, cdt)
        self._check(
            self._kernel("awkward_ListArray_getitem_next_range", cdt, np.int64, cdt, cdt)(
                nextoffsets, nextcarry, starts, stops, lenstarts, kstart, kstop, step
            )
        )
        if bounded:
            nextcarry = nextcarry[: nplike.item(nextoffsets, lenstarts)]
        return ListOffsetArray(nextoffsets, self.content._carry(nextcarry, True), self.backend, exact=True)

    # -- jagged integer index, NumpyArray slice content: listarray.py:549-618 --
    def _getitem_next_jagged(self, slicestarts, slicestops, slicecontent):
        """self sliced list by list by [slicestarts[i], slicestops[i]) of
        slicecontent: an int64 NumpyArray (jagged_carrylen + jagged_apply),
        or lists of a deeper slice (listarray.py:489-547: jagged_descend
        checks every list's length, then the content is sliced by the
        slice's inner lists)."""
        nplike = self.nplike
        starts, stops = self._contig()
        cdt = nplike.dtype(starts)
        if isinstance(slicecontent, ListOffsetArray):
            n = nplike.length(slicestarts)
            outoffsets = nplike.empty(n + 1, np.int64)
            err = self._kernel("awkward_ListArray_getitem_jagged_descend", np.int64, np.int64, np.int64, cdt, cdt)(
                outoffsets, slicestarts, slicestops, n, starts, stops)
            if err.str is not None:
                raise IndexError(self.backend.format_kernel_error(err))
            compact = self.to_ListOffsetArray64()
            offs = compact.offsets
            first, last = nplike.item(offs, 0), nplike.item(offs, nplike.length(offs) - 1)
            next_content = compact.content.slice(first, last)
            if not next_content.is_list:
                raise IndexError("too many jagged slice dimensions for array")
            if isinstance(next_content, ListOffsetArray):
                next_content = next_content.to_ListArray()
            so = slicecontent.offsets
            m = nplike.length(so)
            outcontent = next_content._getitem_next_jagged(so[: m - 1], so[1:], slicecontent.content)
            return ListOffsetArray(outoffsets, outcontent, self.backend)
        sliceindex = slicecontent.data
        carrylen_buf = nplike.empty(1, np.int64)
        self._check(
            self._kernel("awkward_ListArray_getitem_jagged_carrylen", np.int64, np.int64, np.int64)(
                carrylen_buf, slicestarts, slicestops, nplike.length(slicestarts)
            )
        )
        carrylen = nplike.index_as_shape_item(carrylen_buf)
        outoffsets = nplike.empty(nplike.length(slicestarts) + 1, np.int64)
        nextcarry = nplike.empty(carrylen, np.int64)
        err = self._kernel("awkward_ListArray_getitem_jagged_apply", np.int64, np.int64, np.int64, np.int64,
                           np.int64, cdt, cdt)(
            outoffsets, nextcarry, slicestarts, slicestops, nplike.length(slicestarts), sliceindex,
            nplike.length(sliceindex), starts, stops, self.content.length
        )
        if err.str is not None:
            raise IndexError(self.backend.format_kernel_error(err))
        return ListOffsetArray(outoffsets, self.content._carry(nextcarry, True), self.backend)

    def _getitem_jagged(self, item):
        """self[item] for a jagged integer index `item` (a ListOffsetArray of
        int64 with as many lists as self), as Content._getitem wraps it:
        RegularArray(self, len, 1) sliced by the jagged head
        (regulararray.py:731-775 -> awkward_RegularArray_getitem_jagged_expand)."""
        nplike = self.nplike
        if item.length != self.length:
            raise IndexError(
                f"cannot fit jagged slice with length {item.length} into array with length {self.length}")
        multistarts = nplike.empty(item.length, np.int64)
        multistops = nplike.empty(item.length, np.int64)
        self._check(
            self._kernel("awkward_RegularArray_getitem_jagged_expand", np.int64, np.int64, np.int64)(
                multistarts, multistops, item.offsets, item.length, 1
            )
        )
        return self._getitem_next_jagged(multistarts, multistops, item.content)

    @property
    def depth(self):
        return 1 + self.content.depth

    def _reduce_axis(self, reducer, negaxis, starts, shifts, offsets, outlength, mask, keepdims):
        return self.to_ListOffsetArray64()._reduce_axis(reducer, negaxis, starts, shifts, offsets, outlength, mask,
                                                        keepdims)

    def slice(self, start, stop):
        return ListArray(self.starts[start:stop], self.stops[start:stop], self.content, self.backend)

    def to_ListOffsetArray64(self):
        nplike = self.nplike
        starts, stops = self._contig()
        cdt = nplike.dtype(starts)
        offsets = nplike.empty(self.length + 1, np.int64)
        self._check(
            self._kernel("awkward_ListArray_compact_offsets", np.int64, cdt, cdt)(offsets, starts, stops, self.length)
        )
        # content of the compacted lists: the ranges [starts[i], stops[i]) in order
        carry = _ranges_carry(self, offsets)
        return ListOffsetArray(offsets, self.content._carry(carry, True), self.backend)

    def to_list(self):
        s = self.nplike.to_numpy(self.starts)
        e = self.nplike.to_numpy(self.stops)
        flat = self.content.to_list()
        return [flat[a:b] for a, b in zip(s, e)]


def _ranges_carry(listarray, offsets):
    """Index of every element of a ListArray in list order (the reference's
    ListArray._compact_offsets64 + broadcast carry): localindex over the
    compact offsets plus each list's start."""
    nplike = listarray.nplike
    total = nplike.item(offsets, nplike.length(offsets) - 1)
    local = nplike.empty(total, np.int64)
    listarray._check(
        listarray._kernel("awkward_ListArray_localindex", np.int64, np.int64)(local, offsets, listarray.length)
    )
    starts, _ = listarray._contig()
    nextcarry = nplike.empty(total, np.int64)
    # start of each element's list: gather starts by the parents of `local`
    parents = nplike.parents(offsets, total)
    nextcarry = nplike.take(nplike.asarray(starts, np.int64), parents) + local
    return nextcarry


_kSliceNone = 2**63 - 1


def _first(out):
    """Element 0 of a length-1 list level (_do.reduce's ``next[0]``)."""
    nplike = out.nplike
    if isinstance(out, RegularArray):
        return out.content.slice(0, out.size)
    if isinstance(out, ListOffsetArray):
        return out.content.slice(nplike.item(out.offsets, 0), nplike.item(out.offsets, 1))
    if isinstance(out, ListArray):
        return out.content.slice(nplike.item(out.starts, 0), nplike.item(out.stops, 0))
    raise NotImplementedError(type(out).__name__)


def completely_flatten(layout):
    """The numbers of every list, in order, missing values dropped -- what
    ``Content.completely_flatten`` gives the reference's axis=None reductions
    (``_do.py:225-340``; ``contents/listoffsetarray.py``,
    ``indexedoptionarray.py``, ``bytemaskedarray.py``, ``regulararray.py``)."""
    if isinstance(layout, NumpyArray):
        return layout
    if isinstance(layout, (ListArray, RegularArray)):
        layout = layout.to_ListOffsetArray64()
    if isinstance(layout, ListOffsetArray):
        return completely_flatten(layout._zero_based()[1])
    if isinstance(layout, (IndexedOptionArray, ByteMaskedArray)):
        nextcarry, _ = layout._project_valid()
        return completely_flatten(layout.content._carry(nextcarry, False))
    if isinstance(layout, IndexedArray):
        return completely_flatten(layout.project())
    raise TypeError(f"cannot reduce {type(layout).__name__} with axis=None (records have no single number type)")


def reduce(layout, reducer, mask, keepdims, axis=-1):
    """_do.reduce (src/awkward/_do.py:225-340): returns the reduced layout
    (the outer length-1 level removed).  axis=-1 of a list of numbers (or of
    optional numbers) is the local path; higher axes regroup the lists by
    column first (the non-local branch).  Missing lists (an option type over
    the lists) reduce their valid lists and stay missing in the result, as
    the reference's option types do above the reduced axis
    (contents/indexedoptionarray.py:1433-1470)."""
    depth = layout.depth
    negaxis = -axis if axis < 0 else depth - axis
    if not 0 < negaxis <= depth:
        raise ValueError(f"axis={axis} exceeds the depth of the nested list structure (which is {depth})")
    if isinstance(layout, (IndexedOptionArray, ByteMaskedArray)) and getattr(layout.content, "is_list", False):
        nextcarry, outindex = layout._project_valid()
        out = reduce(layout.content._carry(nextcarry, False), reducer, mask, keepdims, axis)
        if negaxis == depth:  # the missing lists are the ones reduced over: they contribute nothing
            return out
        return IndexedOptionArray(outindex, out, layout.backend)
    if depth == 1:
        # a flat array (negaxis == depth == 1): one list holding every element;
        # the result has length 1 and the caller takes its value
        nplike = layout.nplike
        starts = nplike.asarray(np.zeros(1, dtype=np.int64))
        offsets = nplike.asarray(np.array([0, layout.length], dtype=np.int64))
        return layout._reduce_next(reducer, starts, offsets, 1, mask, False)
    if negaxis != 1 or depth > 2:
        nplike = layout.nplike
        starts = nplike.asarray(np.zeros(1, dtype=np.int64))
        offsets = nplike.asarray(np.array([0, layout.length], dtype=np.int64))
        return _first(layout._reduce_axis(reducer, negaxis, starts, None, offsets, 1, mask, keepdims))
    if isinstance(layout, ListArray):
        layout = layout.to_ListOffsetArray64()
    if not isinstance(layout, ListOffsetArray):
        raise NotImplementedError("hip reductions are implemented at axis=-1 of a list array")
    inner = layout.content
    if isinstance(inner, (ListOffsetArray, ListArray, RecordArray)):
        raise NotImplementedError("hip reductions are implemented for one level of lists over numbers")
    return layout._reduce_local(reducer, mask, keepdims)


__all__ = [
    "Content",
    "NumpyArray",
    "RecordArray",
    "IndexedArray",
    "ByteMaskedArray",
    "IndexedOptionArray",
    "RegularArray",
    "ListOffsetArray",
    "ListArray",
    "reduce",
    "completely_flatten",
    "_reducers",
]
