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


def fingerprint(wl):
    """sha256 prefix of a workload's graph and group table (as in tests/golden/make_scale_golden.py)."""
    import hashlib

    import numpy as np

    h = hashlib.sha256()
    for a in (wl.graph.beg_pos, wl.graph.adj, wl.graph.w, wl.groups.offs, wl.groups.verts):
        h.update(np.ascontiguousarray(a).view(np.uint8))
    return h.hexdigest()[:24]


def load_scale(name):
    """(workload, fixture) for tests/golden/scale_<name>.json; the generated graph must
    match the one the oracle solved."""
    import json

    from gpu4gst_1_amd import datasets

    with open(os.path.join(os.path.dirname(__file__), "golden", f"scale_{name}.json")) as f:
        fx = json.load(f)
    if fx.get("partial"):
        pytest.skip(f"scale_{name}.json is incomplete (its generator was interrupted; run it again to finish it)")
    wl = datasets.make_workload(name)
    assert fingerprint(wl) == fx["fingerprint"], f"{name}: generated graph differs from the fixture's"
    return wl, fx
