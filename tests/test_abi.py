"""The C ABI library loads on the CPU and exports every symbol the header
declares; parameter handling (no device calls) works without a GPU."""
import ctypes as C
import os
import re

import pytest

from luxum_amd import _lib
from luxum_amd.params import LX_EY, LX_FP16, LX_INC_AUX1D, LxParams, default_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "luxum.h")


def header_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert set(syms) == set(_lib.EXPORTS)


def test_params_struct_matches_c_defaults():
    L = _lib.lib()
    p = LxParams()
    assert L.lx_params_default(C.byref(p)) == 0
    assert p.as_dict() == default_params().as_dict()
    assert C.sizeof(LxParams) % 8 == 0


def _check(p):
    return _lib.lib().lx_params_check(C.byref(p))


def test_checkdata_rules_of_the_reference():
    # reader::checkData (reader.cc:141-162)
    assert _check(default_params(nx=34, npx=4)) == 0
    assert _check(default_params(nx=35, npx=4)) == _lib.LX_EINVAL
    assert _check(default_params(ny=33, npy=2)) == _lib.LX_EINVAL
    assert _check(default_params(is_planar=1, ny=2)) == _lib.LX_EINVAL
    assert _check(default_params(is_planar=1, ny=1, npy=2)) == _lib.LX_EINVAL
    assert _check(default_params(is_planar=1, ny=1, dy=0.0)) == 0


def test_solver_preconditions():
    ok = default_params(use_pml=1, pml_thickness=6 * 5e-8)
    assert _check(ok) == 0
    assert _check(default_params(use_pml=1, pml_thickness=1 * 5e-8)) == _lib.LX_EINVAL       # < 2 cells
    assert _check(default_params(use_pml=1, pml_thickness=17 * 5e-8)) == _lib.LX_EINVAL      # slabs overlap
    pw = default_params(use_pw=1, tf_lo=(5, 5, 5), tf_hi=(20, 20, 20))
    assert _check(pw) == 0
    bad = default_params(use_pw=1, tf_lo=(0, 5, 5), tf_hi=(20, 20, 20))
    assert _check(bad) == _lib.LX_EINVAL                                                     # face on the wall
    aux = default_params(use_pw=1, incident=LX_INC_AUX1D, tf_lo=(5, 5, 5), tf_hi=(20, 20, 20))
    assert _check(aux) == 0
    aux.k_pw[:] = (0.6, 0.8, 0.0)
    aux.e_pw[:] = (0.0, 0.0, 1.0)
    assert _check(aux) == _lib.LX_EINVAL                                                     # oblique aux line
    ps = default_params(use_ps=1, src_pos=(40, 3, 3))
    assert _check(ps) == _lib.LX_EINVAL
    assert b"point source" in _lib.lib().lx_last_error()
    assert _check(default_params(precision=7)) == _lib.LX_EINVAL


def test_load_reference_yaml_fixtures():
    from luxum_amd.sim import load_yaml
    p = load_yaml(os.path.join(ROOT, "tests", "golden", "parameters.yaml"))
    assert (p.is_planar, p.nx, p.ny, p.nz) == (1, 1002, 1, 1002)
    assert p.dx == 0.001 and p.dy == 0.0 and p.courant == 0.70710678118 and p.num_timesteps == 2000
    assert (p.npx, p.npy, p.n_threads) == (2, 1, 1)
    assert p.use_pw == 1 and p.use_ps == 0 and p.off_pw == 50.0
    assert tuple(p.k_pw) == (0.70710678118, 0.0, 0.70710678118) and tuple(p.e_pw) == (0.0, 1.0, 0.0)
    assert tuple(p.tf_lo) == (20, 20, 20) and tuple(p.tf_hi) == (980, 980, 980)
    assert p.use_pml == 0 and p.pml_thickness == 0.01 and p.sigma_ratio == 0.65 and p.kappa_max == 15
    assert p.src_comp == LX_EY
    q = load_yaml(os.path.join(ROOT, "tests", "golden", "parameters_3d.yaml"))
    assert (q.is_planar, q.nx, q.ny, q.nz) == (0, 102, 102, 102)
    assert q.courant == 0.5773502691896258 and q.num_timesteps == 600
    assert q.use_pml == 1 and q.a_max == 0.0 and q.a_m == 1.0 and q.sigma_m == 3.0
    assert tuple(q.e_pw) == (0.70710678118, 0.0, -0.70710678118)


def test_yaml_extension_document_and_errors(tmp_path):
    from luxum_amd.sim import load_yaml
    base = open(os.path.join(ROOT, "tests", "golden", "parameters_3d.yaml"), encoding="utf-8-sig").read()
    f = tmp_path / "p6.yaml"
    f.write_text(base + '\n{ # extensions\n  "precision": "fp16",\n  "snapshot every": 50,  # comment\n'
                        '  "pw sigma": 1e-12,\n}\n')
    p = load_yaml(str(f))
    assert p.precision == LX_FP16 and p.snapshot_every == 50 and p.pw_sigma == 1e-12
    g = tmp_path / "missing.yaml"
    g.write_text(base.replace('"Courant factor": 0.5773502691896258,', ""))
    with pytest.raises(_lib.LuxumError) as e:
        load_yaml(str(g))
    assert "Courant factor" in str(e.value)
    h = tmp_path / "bad.yaml"
    h.write_text(base + '\n{ "precision": "fp8" }\n')
    with pytest.raises(_lib.LuxumError):
        load_yaml(str(h))
    with pytest.raises(_lib.LuxumError) as e:
        load_yaml(str(tmp_path / "nope.yaml"))
    assert e.value.status == _lib.LX_EIO


def test_footprint_estimate():
    """Admission footprints: grow with the grid, count the second buffers of a
    tiled batch member (2D only), cover the fields and ids, reject invalid
    parameters."""
    from luxum_amd import configs
    from luxum_amd.server import footprint
    from luxum_amd._lib import LuxumError
    w1 = configs.config0(n=256)
    w2 = configs.config0(n=512)
    a, b = footprint(w1.p, False, 1000, 4), footprint(w2.p, False, 1000, 4)
    assert b > a > 256 * 256 * (6 * 4 + 1)
    assert footprint(w1.p, True, 1000, 4) > a + 6 * 4 * 256 * 256
    w3 = configs.config1(n=64)
    assert footprint(w3.p, True, 100, 8) == footprint(w3.p, False, 100, 8) >= 64 ** 3 * (6 * 4 + 1)
    w3.p.num_timesteps = 10   # probe buffers hold max(steps, num_timesteps) records
    assert footprint(w3.p, False, 1000, 8) - footprint(w3.p, False, 100, 8) >= 900 * 8 * 8
    bad = configs.config0(n=256).p
    bad.nx = 0
    with pytest.raises(LuxumError):
        footprint(bad, True, 10, 0)


def test_storage_scale_validation():
    """storage_scale: 0 (automatic) or a power of two; anything else is
    rejected before a simulation is created."""
    from luxum_amd.params import default_params
    from luxum_amd.sim import params_check
    for ok in (0.0, 1.0, 0.25, 2.0 ** 20):
        params_check(default_params(storage_scale=ok))
    for bad in (3.0, -2.0, float("inf"), 2.0 ** 70):
        with pytest.raises(Exception):
            params_check(default_params(storage_scale=bad))
