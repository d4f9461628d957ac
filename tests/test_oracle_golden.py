"""The CPU oracle against every golden vector the reference holds."""

import pytest

import oracle
from kernel_harness import Case, HostMemory, check_case, load_vectors, specializations

VECTORS = load_vectors()
PARAMS = [
    (family, i)
    for family, tests in sorted(VECTORS.items())
    for i in range(len(tests))
]


@pytest.mark.parametrize("family,index", PARAMS)
def test_oracle_vector(family, index):
    vector = VECTORS[family][index]
    ran = 0
    for name in specializations(family):
        case = Case(name, vector)
        if not case.ok:
            continue
        check_case(oracle.kernel(name), case, HostMemory())
        ran += 1
    assert ran > 0, f"no specialization of {family} can hold vector {index}"
