import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")


@pytest.fixture(scope="session")
def golden():
    path = os.path.join(ROOT, "tests", "golden", "golden.npz")
    with np.load(path, allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}
    return data


def golden_cases(data, prefix):
    names = sorted({k.split("/")[1] for k in data if k.startswith(prefix + "/")})
    out = {}
    for n in names:
        p = f"{prefix}/{n}/"
        out[n] = {k[len(p):]: v for k, v in data.items() if k.startswith(p)}
    return out


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.build()
    return oracle
