"""Reducers: kernel choice and dtype rules of the reference's
``src/awkward/_reducers.py`` for the axis=-1 path.

* ``Sum`` (:332-455): bool, int8/16/32 accumulate in int64 and unsigned
  narrower types in uint64 (``_promote_integer_rank`` :73-82 with the
  Linux/numpy-2 answer ``_use32 = False``); float32 stays float32.
* ``Max`` / ``Min`` (:861-1103): identity ``-inf`` / ``+inf`` or the integer
  type's min / max unless ``initial`` is given; result keeps the dtype.
* ``ArgMax`` / ``ArgMin`` (:132-251): int64 global indices, then
  ``apply_positional_corrections`` (:85-129) subtracts the list starts; bool
  data is viewed as int8.
* ``Count`` (:254-279).
* ``Prod`` (:676-758): the types of ``Sum``; booleans reduce with
  ``awkward_reduce_prod_bool`` and the result is cast to int64.
* ``Any`` / ``All`` (:760-858): ``awkward_reduce_sum_bool`` /
  ``awkward_reduce_prod_bool`` of any dtype, a bool result.  ``Max`` /
  ``Min`` of booleans are these two kernels (:1046-1062, :899-913), and keep
  the bool dtype.

``apply`` returns ``(result, mask)``.  On a backend with ``fused = True``
(the hip backend) the result comes from one tuned kernel per reduction --
with argmin/argmax's start correction and mask_identity's empty-list mask
written in the same pass, and the tile variant chosen by
``awkward_amd.dispatch`` -- instead of the reference's sequence of separate
kernels (reduce, adjust_starts, reduce_mask), which other backends run.
Both give the same bits.
"""

from __future__ import annotations

import numpy as np


def _promote_integer_rank(dtype: np.dtype) -> np.dtype:
    if dtype in (np.dtype(np.bool_), np.dtype(np.int8), np.dtype(np.int16), np.dtype(np.int32)):
        return np.dtype(np.int64)
    if dtype in (np.dtype(np.uint8), np.dtype(np.uint16), np.dtype(np.uint32)):
        return np.dtype(np.uint64)
    return dtype


class Reducer:
    name = "reducer"
    needs_position = False


def _fused(array):
    return getattr(array.backend, "fused", False)


def _variant(op, array, outlength, offsets=None):
    from awkward_amd import dispatch

    return dispatch.choose_variant(op, outlength, array.length, array.dtype.itemsize, offsets)


class Sum(Reducer):
    name = "sum"

    def apply(self, array, offsets, starts, shifts, outlength, mask=False):
        nplike = array.nplike
        dtype = array.dtype
        out_dtype = np.dtype(np.int64) if dtype == np.dtype(np.bool_) else _promote_integer_rank(dtype)
        result = nplike.empty(outlength, out_dtype)
        if _fused(array):
            from awkward_amd import dispatch

            array._check(dispatch.reduce_sum(result, array.data, offsets, outlength,
                                             _variant("sum", array, outlength, offsets)))
        elif dtype == np.dtype(np.bool_):
            array._check(
                array.backend["awkward_reduce_sum_int64_bool_64", np.int64, np.bool_, np.int64](
                    result, array.data, offsets, outlength
                )
            )
        else:
            array._check(
                array.backend["awkward_reduce_sum", out_dtype, dtype, np.int64](result, array.data, offsets,
                                                                                outlength)
            )
        return type(array)(result, array.backend), None


class _Extremum(Reducer):
    kernel = ""
    is_max = True

    def __init__(self, initial=None):
        self.initial = initial

    def identity_for(self, dtype: np.dtype):
        if self.initial is not None:
            return self.initial
        if dtype.kind in "iu":
            info = np.iinfo(dtype)
            return int(info.min if self.is_max else info.max)
        return -np.inf if self.is_max else np.inf

    def apply(self, array, offsets, starts, shifts, outlength, mask=False):
        dtype = array.dtype
        if dtype == np.dtype(np.bool_):
            # _reducers.py:1046-1062 (max: any) and :899-913 (min: all)
            return _any_all("any" if self.is_max else "all", array, offsets, outlength, mask)
        result = array.nplike.empty(outlength, dtype)
        identity = self.identity_for(dtype)
        if dtype.kind == "f":
            identity = float(identity)
        if _fused(array):
            from awkward_amd import dispatch

            outmask = array.nplike.empty(outlength, np.int8) if mask else None
            array._check(dispatch.reduce_extremum(self.name, result, outmask, array.data, offsets, outlength,
                                                  identity, _variant(self.name, array, outlength, offsets)))
            return type(array)(result, array.backend), outmask
        array._check(
            array.backend[self.kernel, dtype, dtype, np.int64](result, array.data, offsets, outlength, identity)
        )
        return type(array)(result, array.backend), None


class Max(_Extremum):
    name = "max"
    kernel = "awkward_reduce_max"
    is_max = True


class Min(_Extremum):
    name = "min"
    kernel = "awkward_reduce_min"
    is_max = False


class _Arg(Reducer):
    kernel = ""
    needs_position = True

    def apply(self, array, offsets, starts, shifts, outlength, mask=False):
        nplike = array.nplike
        data = array.data
        dtype = array.dtype
        if dtype == np.dtype(np.bool_):
            data = nplike.view(data, np.int8)
            dtype = np.dtype(np.int8)
        result = nplike.empty(outlength, np.int64)
        if _fused(array):
            from awkward_amd import dispatch

            outmask = nplike.empty(outlength, np.int8) if mask else None
            array._check(dispatch.reduce_arg(self.name, result, outmask, data, offsets, outlength,
                                             _variant(self.name, array, outlength, offsets),
                                             shifts, starts))
            return type(array)(result, array.backend), outmask
        array._check(
            array.backend[self.kernel, np.int64, dtype, np.int64, np.int64](result, data, offsets, starts, outlength)
        )
        # apply_positional_corrections: below an option type the shifts count
        # the nulls removed before each element
        if shifts is None:
            array._check(
                array.backend["awkward_NumpyArray_reduce_adjust_starts_64", np.int64, np.int64, np.int64](
                    result, outlength, offsets, starts
                )
            )
        else:
            array._check(
                array.backend["awkward_NumpyArray_reduce_adjust_starts_shifts_64", np.int64, np.int64, np.int64,
                              np.int64](result, outlength, offsets, starts, shifts)
            )
        return type(array)(result, array.backend), None


class ArgMax(_Arg):
    name = "argmax"
    kernel = "awkward_reduce_argmax"


class ArgMin(_Arg):
    name = "argmin"
    kernel = "awkward_reduce_argmin"


class Count(Reducer):
    name = "count"

    def apply(self, array, offsets, starts, shifts, outlength, mask=False):
        result = array.nplike.empty(outlength, np.int64)
        array._check(array.backend["awkward_reduce_count_64", np.int64, np.int64](result, offsets, outlength))
        return type(array)(result, array.backend), None


def _any_all(kind, array, offsets, outlength, mask):
    """awkward_reduce_sum_bool (kind "any") or awkward_reduce_prod_bool
    ("all") into a bool result; on the fused backend with the empty-list
    mask of mask_identity=True from the same pass."""
    result = array.nplike.empty(outlength, np.bool_)
    if _fused(array):
        from awkward_amd import dispatch

        outmask = array.nplike.empty(outlength, np.int8) if mask else None
        array._check(dispatch.reduce_anyall(kind, result, outmask, array.data, offsets, outlength,
                                            _variant(kind, array, outlength, offsets)))
        return type(array)(result, array.backend), outmask
    family = "awkward_reduce_sum_bool" if kind == "any" else "awkward_reduce_prod_bool"
    array._check(array.backend[family, np.bool_, array.dtype, np.int64](result, array.data, offsets, outlength))
    return type(array)(result, array.backend), None


class Prod(Reducer):
    name = "prod"

    def apply(self, array, offsets, starts, shifts, outlength, mask=False):
        nplike = array.nplike
        dtype = array.dtype
        if dtype == np.dtype(np.bool_):
            # _reducers.py:691-714: all-nonzero per list, cast to the promoted integer type
            flags, _ = _any_all("all", array, offsets, outlength, False)
            return type(array)(nplike.asarray(flags.data, _promote_integer_rank(dtype)), array.backend), None
        out_dtype = _promote_integer_rank(dtype)
        result = nplike.empty(outlength, out_dtype)
        if _fused(array):
            from awkward_amd import dispatch

            array._check(dispatch.reduce_prod(result, array.data, offsets, outlength,
                                              _variant("prod", array, outlength, offsets)))
        else:
            array._check(array.backend["awkward_reduce_prod", out_dtype, dtype, np.int64](result, array.data, offsets,
                                                                                         outlength))
        return type(array)(result, array.backend), None


class Any(Reducer):
    name = "any"

    def apply(self, array, offsets, starts, shifts, outlength, mask=False):
        return _any_all("any", array, offsets, outlength, mask)


class All(Reducer):
    name = "all"

    def apply(self, array, offsets, starts, shifts, outlength, mask=False):
        return _any_all("all", array, offsets, outlength, mask)
