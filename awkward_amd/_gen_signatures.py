"""Generator for the hip backend's kernel table and C header.

Reads Awkward's kernel contract (``kernel-specification.yml``) and emits, for
the jagged-array kernel families this backend serves:

* ``awkward_amd/_kernel_signatures.py`` -- one row per specialization: the
  family name, the argument list (name, spec type, direction) and the
  registry key ``(family, *dtype names of the pointer arguments)``.  The key
  rule is the one the reference's generator applies
  (``dev/generate-kernel-signatures.py:225-236`` and ``:312-331``): every
  ``List[...]`` argument contributes its element dtype once, in argument
  order.
* ``include/awkward_hip.h`` -- the C-ABI prototypes of the shared library,
  each citing the specification entry and the CPU kernel it replaces.

Run it where the reference checkout is available::

    python awkward_amd/_gen_signatures.py /root/reference

The outputs are committed; nothing at run time reads the specification.
"""

from __future__ import annotations

import os
import sys

import yaml

#: kernel families served by the hip backend (the jagged-array hot path)
FAMILIES = [
    "awkward_reduce_sum",
    "awkward_reduce_sum_int64_bool_64",
    "awkward_reduce_sum_int32_bool_64",
    "awkward_reduce_max",
    "awkward_reduce_min",
    # ak.prod, ak.any / ak.all, and ak.max / ak.min of booleans (_reducers.py:676-758, 760-858, 899-913, 1046-1062)
    "awkward_reduce_prod",
    "awkward_reduce_sum_bool",
    "awkward_reduce_prod_bool",
    "awkward_reduce_argmax",
    "awkward_reduce_argmin",
    "awkward_reduce_count_64",
    "awkward_NumpyArray_reduce_adjust_starts_64",
    "awkward_NumpyArray_reduce_mask_ByteMaskedArray_64",
    "awkward_localindex",
    "awkward_ListArray_localindex",
    "awkward_RegularArray_getitem_jagged_expand",
    "awkward_ListArray_getitem_jagged_carrylen",
    "awkward_ListArray_getitem_jagged_apply",
    "awkward_ListArray_getitem_jagged_descend",
    "awkward_ListArray_getitem_next_range",
    "awkward_ListArray_getitem_next_range_carrylength",
    "awkward_ListArray_getitem_next_range_counts",
    "awkward_RegularArray_getitem_next_range",
    "awkward_ListArray_getitem_next_at",
    "awkward_ListArray_compact_offsets",
    "awkward_ListArray_combinations_length",
    "awkward_ListArray_combinations",
    "awkward_IndexedArray_getitem_nextcarry",
    "awkward_ListOffsetArray_local_preparenext_64",
    "awkward_sort",
    "awkward_argsort",
    # option-type contents under the reducers and getitem
    "awkward_ByteMaskedArray_numnull",
    "awkward_ByteMaskedArray_reduce_next_64",
    "awkward_ByteMaskedArray_reduce_next_nonlocal_nextshifts_64",
    "awkward_ByteMaskedArray_reduce_next_nonlocal_nextshifts_fromshifts_64",
    "awkward_ByteMaskedArray_getitem_nextcarry",
    "awkward_ByteMaskedArray_getitem_nextcarry_outindex",
    "awkward_IndexedArray_numnull",
    "awkward_IndexedArray_reduce_next_64",
    "awkward_IndexedArray_reduce_next_nonlocal_nextshifts_64",
    "awkward_IndexedArray_reduce_next_nonlocal_nextshifts_fromshifts_64",
    "awkward_IndexedArray_getitem_nextcarry_outindex",
    "awkward_NumpyArray_reduce_adjust_starts_shifts_64",
    # non-local (axis != -1) reductions
    "awkward_ListOffsetArray_reduce_nonlocal_maxcount_offsetscopy_64",
    "awkward_ListOffsetArray_reduce_nonlocal_preparenext_64",
    "awkward_ListOffsetArray_reduce_nonlocal_nextshifts_64",
    "awkward_ListOffsetArray_reduce_nonlocal_outstartsstops_64",
]

_DTYPE = {
    "bool": "bool",
    "int8": "int8",
    "uint8": "uint8",
    "int16": "int16",
    "uint16": "uint16",
    "int32": "int32",
    "uint32": "uint32",
    "int64": "int64",
    "uint64": "uint64",
    "float": "float32",
    "double": "float64",
}


def _parse_type(typename: str):
    """Return (const, depth, base) for a spec type string."""
    const = typename.startswith("Const[")
    if const:
        typename = typename[len("Const[") : -1]
    depth = 0
    while typename.startswith("List["):
        depth += 1
        typename = typename[len("List[") : -1]
    base = typename[:-2] if typename.endswith("_t") else typename
    return const, depth, base


def _c_type(typename: str) -> str:
    const, depth, base = _parse_type(typename)
    cbase = {"float": "float", "double": "double", "bool": "bool"}.get(base, base + "_t")
    return ("const " if const else "") + cbase + "*" * depth


def _cpu_citation(reference: str, family: str, spec: str) -> str:
    src = os.path.join(reference, "awkward-cpp", "src", "cpu-kernels")
    for candidate in (family + ".cpp", spec + ".cpp"):
        path = os.path.join(src, candidate)
        if os.path.exists(path):
            with open(path) as f:
                for lineno, line in enumerate(f, 1):
                    if spec in line:
                        return f"awkward-cpp/src/cpu-kernels/{candidate}:{lineno}"
            return f"awkward-cpp/src/cpu-kernels/{candidate}"
    return "(no CPU source)"


def main(reference: str) -> None:
    spec_path = os.path.join(reference, "kernel-specification.yml")
    with open(spec_path) as f:
        text = f.read()
    spec = yaml.safe_load(text)
    lines = text.splitlines()
    line_of = {}
    for lineno, line in enumerate(lines, 1):
        stripped = line.strip()
        if stripped.startswith("- name: "):
            line_of.setdefault(stripped[len("- name: ") :], lineno)

    by_name = {k["name"]: k for k in spec["kernels"]}
    rows = []
    for family in FAMILIES:
        kernel = by_name[family]
        for sp in kernel["specializations"]:
            args = [(a["name"], a["type"], a["dir"]) for a in sp["args"]]
            key = [family]
            for _, t, _ in args:
                _, depth, base = _parse_type(t)
                if depth:
                    key.append(_DTYPE[base])
            rows.append((family, sp["name"], args, tuple(key), line_of[sp["name"]]))

    here = os.path.dirname(os.path.abspath(__file__))
    repo = os.path.dirname(here)

    with open(os.path.join(here, "_kernel_signatures.py"), "w") as out:
        out.write(
            '"""Kernel table of the hip backend (generated by _gen_signatures.py; do not edit).\n\n'
            "Each row: specialization name -> (family, args, key).  ``args`` lists\n"
            "(name, spec type, direction) in call order; ``key`` is the registry key\n"
            "(family, *dtype names of the pointer arguments), built with the rule of\n"
            "dev/generate-kernel-signatures.py:225-236.\n"
            '"""\n\n'
            "# fmt: off\n"
            "SPECIALIZATIONS = {\n"
        )
        for family, name, args, key, _ in rows:
            out.write(f"    {name!r}: ({family!r}, {args!r}, {key!r}),\n")
        out.write("}\n")

    with open(os.path.join(repo, "include", "awkward_hip.h"), "w") as out:
        out.write(HEADER_PROLOGUE)
        current = None
        for family, name, args, key, lineno in rows:
            if family != current:
                current = family
                out.write(f"\n/* ---- {family} ---- */\n")
            cite = _cpu_citation(reference, family, name)
            params = ", ".join(f"{_c_type(t)} {n}" for n, t, _ in args)
            out.write(f"/* replaces {cite}; contract kernel-specification.yml:{lineno} */\n")
            out.write(f"struct Error {name}({params});\n")
        out.write(HEADER_EPILOGUE)
    print(f"wrote {len(rows)} specializations")


HEADER_PROLOGUE = r"""/* awkward_hip.h -- C ABI of libawkward_hip.so (generated by
 * awkward_amd/_gen_signatures.py from kernel-specification.yml; the runtime
 * section at the end is written by hand in the generator).
 *
 * Every kernel entry point keeps the name, argument order and meaning of the
 * awkward-cpp CPU kernel it replaces, and returns the same struct Error by
 * value (awkward-cpp/include/awkward/common.h:39-78: str == NULL means
 * success; identity / attempt carry the failing position, kSliceNone when
 * absent).  The differences from the CPU ABI:
 *
 *   - every pointer argument is a DEVICE pointer on the current HIP device,
 *     including the array of output pointers of the combinations kernel
 *     (tocarry is a device array of n device pointers);
 *   - kernels are enqueued on the stream set by awkward_hip_set_stream()
 *     (default: the null stream) and return before they complete, except
 *     where a result is needed on the host to size an allocation (none of
 *     the entry points here do that: the *_carrylen / *_carrylength /
 *     totallen outputs stay in device memory, as in the reference);
 *   - data-dependent errors (bad indices, stops < starts) are detected on
 *     the device.  In the default synchronous-error mode the entry point
 *     waits for its kernels and returns the first error in the CPU kernel's
 *     loop order, with the same message, identity and attempt.  In deferred
 *     mode (awkward_hip_set_deferred_errors(1)) it returns success at once
 *     and awkward_hip_error_poll() reports the first recorded error later,
 *     like the reference CUDA backend's error word
 *     (src/awkward/_connect/cuda/__init__.py:257-294).
 */
#ifndef AWKWARD_HIP_H_
#define AWKWARD_HIP_H_

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct Error {
  const char* str;
  const char* filename;
  int64_t identity;
  int64_t attempt;
};
"""

HEADER_EPILOGUE = r"""
/* ---- runtime (no counterpart in awkward-cpp, which is synchronous CPU code) ---- */

/* Stream that subsequent entry points enqueue on (a hipStream_t; NULL = the
   null stream).  Mirrors the reference CUDA backend's use of the current
   CuPy stream (src/awkward/_kernels.py:188-232). */
void awkward_hip_set_stream(void* stream);
void* awkward_hip_get_stream(void);

/* 0 (default): entry points with device-detected errors synchronise and
   return the error by value.  1: they return at once; errors accumulate in
   a device error word read by awkward_hip_error_poll. */
void awkward_hip_set_deferred_errors(int deferred);

/* Synchronises the current stream and returns the first error recorded in
   deferred mode since the last poll (then clears it). */
struct Error awkward_hip_error_poll(void);

/* Library version string and the gfx target it was built for. */
const char* awkward_hip_version(void);

#ifdef __cplusplus
}
#endif

#endif /* AWKWARD_HIP_H_ */
"""

if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
