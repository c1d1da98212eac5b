import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, HERE)
sys.path.insert(0, ROOT)

import graphs  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def golden_index():
    with open(os.path.join(GOLDEN, "index.json")) as f:
        return json.load(f)


def load_golden_graph(name):
    z = np.load(os.path.join(GOLDEN, f"graph_{name}.npz"))
    return graphs.CSR(z["degree"], z["adj"], z["weight"], z["label"]), z["sources"]


def golden_case_args(case):
    ex = case["extra"]
    return dict(stop_prob=float(ex.get("-sp", 0.2)), p=float(ex.get("-p", 2.0)), q=float(ex.get("-q", 0.5)),
                schema=[int(x) for x in ex["-s"].split(",")] if "-s" in ex else None)


@pytest.fixture(scope="session")
def golden_graphs(golden_index):
    return {name: load_golden_graph(name) for name in golden_index["graphs"]}


@pytest.fixture(scope="session")
def gpu_ctx():
    import thunderrw_amd as T

    if T.device_count() < 1:
        pytest.fail("no HIP device visible to a test marked gpu")
    ctx = T.Context(0)
    yield ctx
    ctx.close()
