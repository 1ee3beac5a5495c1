import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import __graft_entry__  # noqa: E402

__graft_entry__.load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def rmat():
    from gpu4gst_1_amd import datasets
    return datasets.make_workload("rmat32k")


@pytest.fixture(scope="session")
def rmat_inst(rmat):
    import pyoracle
    return pyoracle.Instance(rmat.graph, rmat.groups)


@pytest.fixture(scope="session")
def gpu_engine(rmat):
    from gpu4gst_1_amd import engine
    eng = engine.Engine(rmat.graph, rmat.groups, device=0)
    yield eng
    eng.close()
