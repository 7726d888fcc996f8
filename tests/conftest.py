import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def small_set():
    """20k x 96 normalised mixture, 200 queries, R <= 32 graph, exact truth."""
    from parallel_anns_amd import datasets
    return datasets.prepare("t20k", verbose=False)


@pytest.fixture(scope="session")
def tiny_set():
    from parallel_anns_amd import datasets
    return datasets.prepare("t5k", verbose=False)


@pytest.fixture(scope="session")
def gpu_index(small_set):
    from parallel_anns_amd.search import Index
    idx = Index(0)
    idx.set_base(small_set["base"])
    idx.set_queries(small_set["queries"])
    idx.set_graph(small_set["adj"], small_set["deg"], small_set["ep"])
    yield idx
    idx.close()
