"""Extract the reference's golden kernel vectors for the hip backend's kernels.

Reads ``kernel-test-data.json`` (the ``unit-tests`` section) of an Awkward
checkout and writes ``tests/golden/kernel_vectors.json`` with the vectors of
every kernel family in ``awkward_amd._gen_signatures.FAMILIES``.  The file is
data only: inputs, expected outputs and expected error messages, exactly as
the reference holds them.

    python tests/golden/make_fixtures.py /root/reference
"""

from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from awkward_amd._gen_signatures import FAMILIES  # noqa: E402


def main(reference: str) -> None:
    with open(os.path.join(reference, "kernel-test-data.json")) as f:
        data = json.load(f)
    out = {}
    for entry in data["unit-tests"]:
        if entry["name"] in FAMILIES:
            out[entry["name"]] = entry["tests"]
    with open(os.path.join(HERE, "kernel_vectors.json"), "w") as f:
        json.dump(out, f, indent=0, sort_keys=True)
    print({k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
