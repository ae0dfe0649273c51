import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libtpr_hip.so")


@pytest.fixture(scope="session")
def golden():
    import json

    import numpy as np
    d = os.path.join(ROOT, "tests", "golden")
    with open(os.path.join(d, "ref_flops.json")) as f:
        flops = json.load(f)
    return {
        "flops": flops,
        "ans1024": np.load(os.path.join(d, "ans1024.npy")),
        "nstab": dict(np.load(os.path.join(d, "nstab.npz"))),
        "ref": dict(np.load(os.path.join(d, "ref_vectors.npz"))),
    }
