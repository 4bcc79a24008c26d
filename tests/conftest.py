import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def rel_l2(a, b) -> float:
    """||a - b||_2 / ||b||_2 (b is the FP64 oracle)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b.ravel())
    if nb == 0.0:
        return float(np.linalg.norm(a.ravel()))
    return float(np.linalg.norm((a - b).ravel()) / nb)


@pytest.fixture(scope="session")
def gpu_ctx():
    from luxum_amd.sim import Context
    return Context(0)


def reference_yaml(base: str, dst, **over) -> str:
    """Copy a reference-format parameters.yaml from tests/golden with some
    "Key": value entries replaced (keys as the reference spells them)."""
    import re
    text = open(os.path.join(ROOT, "tests", "golden", base), encoding="utf-8-sig").read()
    for key, val in over.items():
        if isinstance(val, bool):
            val = "true" if val else "false"
        text, n = re.subn(r'("%s"\s*:\s*)[^,\n]+' % re.escape(key), r"\g<1>%s" % val, text)
        assert n == 1, key
    with open(dst, "w") as f:
        f.write(text)
    return str(dst)
