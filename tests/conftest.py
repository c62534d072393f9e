import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def golden(name):
    """Load a committed golden fixture (numbers only, no pickles)."""
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def unpack_list(cnt, packed, width):
    """Fixed-stride list from the packed (count, entries) form of a fixture."""
    out = np.zeros((len(cnt), width), np.int32)
    off = 0
    for i, c in enumerate(cnt):
        out[i, :c] = packed[off:off + c]
        off += c
    return out


def rel_max(a, b):
    """max |a - b| / max |b|: the force / virial error norm (no per-component
    relative error, since net forces near equilibrium approach 0)."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="session")
def fmd_lib():
    from fuller_md_en_amd import _native
    return _native.lib()
