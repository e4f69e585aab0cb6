import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, GOLDEN):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtcx.so on cuda:0)")


def golden_names():
    """Oracle golden vectors (qe_example*.npz hold the reference's QE examples instead)."""
    names = (os.path.splitext(os.path.basename(p))[0] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))
    return sorted(n for n in names if not n.startswith("qe_example"))


@pytest.fixture(scope="session")
def golden():
    out = {}
    for name in golden_names():
        with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
            out[name] = {k: z[k] for k in z.files}
    return out


@pytest.fixture(scope="session")
def engine():
    """One GPU context for the whole GPU session (one process, one device)."""
    from tc_amd import tcx
    eng = tcx.Engine(0)
    yield eng
    eng.close()
