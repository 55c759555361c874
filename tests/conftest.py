import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def orc():
    """The CPU oracle (oracle/dim2d_oracle.c), built on first use."""
    from oracle import oracle
    oracle.lib()
    return oracle


def max_rel_err(a, b, region=(slice(2, -2), slice(2, -2)), ks=range(4)):
    """Per-variable max |a-b| / max |b| over `region` (the interior by default)."""
    out = []
    for k in ks:
        ref = b[k][region]
        scale = np.abs(ref).max()
        out.append(float(np.abs(a[k][region] - ref).max() / (scale if scale > 0 else 1.0)))
    return out
