import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build()
    O.lib()
    O.use_reference(False)
    return O


@pytest.fixture(scope="session")
def golden():
    d = os.path.join(ROOT, "tests", "golden")
    with open(os.path.join(d, "golden.json")) as f:
        meta = json.load(f)
    arrays = np.load(os.path.join(d, "golden.npz"))
    return [(c, arrays[c["name"]]) for c in meta["cases"]]


@pytest.fixture(scope="session")
def ctx():
    import mdz_amd as M
    c = M.Context(0)
    yield c
    c.close()
