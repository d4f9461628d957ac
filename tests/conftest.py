import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


_GPU_ORDER = ["test_hip_golden", "test_hip_parity", "test_pipeline", "test_fused", "test_batch"]


def pytest_collection_modifyitems(session, config, items):
    """GPU files run core kernel parity first, fused/graph forms later, and
    test_zz_* files after everything else."""

    def rank(item):
        mod = item.module.__name__.rsplit(".", 1)[-1]
        if mod.startswith("test_zz_"):
            return len(_GPU_ORDER)
        return _GPU_ORDER.index(mod) if mod in _GPU_ORDER else -1

    items.sort(key=rank)
