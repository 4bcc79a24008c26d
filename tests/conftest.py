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


# Parity of each field component under the mirror planes of a workload that
# is symmetric about the grid centre in y and z (node index i <-> n-1-i, and
# a component staggered along the axis maps its index m <-> n-2-m: in both
# cases the stored array is reversed along that axis).  A plane wave along
# +x polarised along z (E_inc = z g, H_inc = -y g/eta) on a geometry
# symmetric under y -> -y is unchanged by that reflection: the polar E keeps
# Ex, Ez and flips Ey, the axial H keeps Hy and flips Hx, Hz.  The reflection
# z -> -z turns the incident wave into its negative, so the field is the
# negative of its mirror image: Ex, Ey, Hz are odd and Ez, Hx, Hy even.
MIRROR_PARITY = {  # component -> (parity about the y centre plane, about the z centre plane)
    0: (+1, -1),   # Ex
    1: (-1, -1),   # Ey
    2: (+1, +1),   # Ez
    3: (-1, +1),   # Hx
    4: (+1, +1),   # Hy
    5: (-1, -1),   # Hz
}


def mirror_asymmetry(field: np.ndarray, axis: int, parity: int) -> float:
    """||F - parity * reverse(F, axis)|| / ||F|| (0 for an exact mirror image)."""
    f = np.asarray(field, dtype=np.float64)
    nf = float(np.linalg.norm(f.ravel()))
    if nf == 0.0:
        return 0.0
    return float(np.linalg.norm((f - parity * np.flip(f, axis=axis)).ravel()) / nf)
