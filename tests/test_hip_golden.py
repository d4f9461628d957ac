"""Every golden vector of the reference, through libawkward_hip.so on the GPU."""

import pytest

from kernel_harness import Case, DeviceMemory, check_case, load_vectors, specializations

pytestmark = pytest.mark.gpu

VECTORS = load_vectors()
PARAMS = [(family, i) for family, tests in sorted(VECTORS.items()) for i in range(len(tests))]


@pytest.fixture(scope="module")
def hip():
    from awkward_amd._kernels import HipKernel

    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = HipKernel(name)._impl
        return cache[name]

    return get


@pytest.mark.parametrize("family,index", PARAMS)
def test_hip_vector(hip, family, index):
    vector = VECTORS[family][index]
    ran = 0
    for name in specializations(family):
        case = Case(name, vector)
        if not case.ok:
            continue
        check_case(hip(name), case, DeviceMemory())
        ran += 1
    assert ran > 0
