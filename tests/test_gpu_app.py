"""The luxum executable (the reference's main.cc on include/luxum.hpp) on the
GPU, one case per solve* loop: the HDF5 file, dataset names and cadence of
that loop, every dataset equal to the state the Python API reaches after the
same number of steps, and the printed "t: <t>, max: <max>" lines equal to the
monitor at the step the reference prints."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, reference_yaml
from luxum_amd.params import LX_EX, LX_EY, LX_EZ

pytestmark = pytest.mark.gpu

APP = os.path.join(ROOT, "luxum_amd", "luxum")

# base yaml, overrides, output file, writes every, from t, print after step, time line
CASES = {
    "solve3d_pml": ("parameters_3d.yaml", {"Number of timesteps": 30}, "Ez_data.h5", 3, 0, False, False),
    "solve3d": ("parameters_3d.yaml", {"Number of timesteps": 21, "usePML": False}, "Ez_data2.h5", 10, 0, False, True),
    "solve_planar_pml": ("parameters.yaml", {"X Length": 202, "Z Length": 202, "x2_pw": 180, "y2_pw": 180,
                                             "z2_pw": 180, "Number of timesteps": 40, "usePML": True},
                         "Ey_data_2d.h5", 10, 0, False, False),
    "solve_planar": ("parameters.yaml", {"X Length": 202, "Z Length": 202, "x2_pw": 180, "y2_pw": 180,
                                         "z2_pw": 180, "Number of timesteps": 141}, "Ey_data_2d.h5", 20, 100, True, True),
}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def e2(s, p):
    nx, ny, nz = p.dims
    ex, ey, ez = s.field(LX_EX), s.field(LX_EY), s.field(LX_EZ)
    out = np.zeros((nx, ny, nz - 1))
    c = (slice(0, nx - 1), slice(0, ny - 1), slice(0, nz - 1))
    out[c] = ez[c] ** 2 + ey[c] ** 2 + ex[c] ** 2
    return out


@pytest.mark.parametrize("name", list(CASES))
def test_luxum_app_matches_api(gpu_ctx, tmp_path, name):
    from luxum_amd.sim import Simulation, h5_read, load_yaml
    base, over, fname, every, start, after, timed = CASES[name]
    say(f"luxum app case {name}")
    path = reference_yaml(base, tmp_path / "parameters.yaml", **over)
    p = load_yaml(path)
    mats = None
    if name == "solve3d_pml":
        # Cb inputs of maxwell_three_dim_pml.cc:109-128: dataset "data" of each
        # component's shape, eps_r = 1 + 1e6*data, three different files: a
        # block in the TF box with eps_r 4 for Ex, 2 for Ey and 8 for Ez
        from luxum_amd.sim import h5_write
        nx, ny, nz = p.dims
        for c, ax in enumerate("xyz"):
            shape = [nx, ny, nz]
            shape[c] -= 1
            data = np.zeros(shape)
            data[40:60, 45:55, 38:62] = (3e-6, 1e-6, 7e-6)[c]
            h5_write(str(tmp_path / f"airplane_100_{ax}.h5"), "data", data)
        ids = np.zeros((nx, ny, nz), np.uint8)
        ids[40:60, 45:55, 38:62] = 1
        mats = (ids, np.array([[1.0, 1.0, 1.0], [4.0, 2.0, 8.0]]))
    r = subprocess.run([APP], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    say(r.stdout[-300:] + r.stderr[-300:])
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().split("\n")
    prints = [ln for ln in lines if ln.startswith("t: ")]
    assert len(prints) == p.num_timesteps
    assert ("Time taken" in lines[-1]) == timed

    s = Simulation(p, gpu_ctx)
    if mats is not None:
        s.set_materials(*mats)
    planar = p.is_planar
    steps_done = 0
    mx_after = [0.0]                      # max after k steps
    expect = {}
    for t in range(p.num_timesteps):
        s.step(1)
        steps_done += 1
        mx_after.append(s.monitor()[0])
        if t % every == 0 and t >= start:
            expect[t] = s.field(LX_EY) if planar else e2(s, p)
    for t, ln in enumerate(prints):
        m = re.fullmatch(r"t: (\d+), max: (\S+)", ln)
        assert m and int(m.group(1)) == t
        ref = mx_after[t + 1] if after else mx_after[t]
        assert float(m.group(2)) == pytest.approx(ref, rel=1e-5, abs=1e-300)
    from luxum_amd.sim import h5_dataset_name
    for t, want in expect.items():
        got = h5_read(str(tmp_path / fname), h5_dataset_name(t))
        assert got.shape == want.shape
        assert np.array_equal(got, want), (name, t)
    with pytest.raises(Exception):
        h5_read(str(tmp_path / fname), h5_dataset_name(max(expect) + every))
    assert max(mx_after) > 0


def test_luxum_app_three_material_files_vs_oracle(gpu_ctx, tmp_path):
    """solve3d_pml with three different airplane_100_{x,y,z}.h5 files: the
    last |E|^2 dataset the app writes against the FP64 oracle run with the
    same per-component permittivities (|E|^2 relative L2 <= 2e-5, twice the
    field tolerance)."""
    from luxum_amd.sim import h5_dataset_name, h5_read, h5_write, load_yaml
    from oracle.lxoracle import OracleSim
    path = reference_yaml("parameters_3d.yaml", tmp_path / "parameters.yaml", **{"Number of timesteps": 100})
    p = load_yaml(path)
    nx, ny, nz = p.dims
    say("luxum app: three different Cb files vs oracle, 100 steps")
    for c, ax in enumerate("xyz"):
        shape = [nx, ny, nz]
        shape[c] -= 1
        data = np.zeros(shape)
        data[30:70, 35:65, 28:72] = (3e-6, 1e-6, 7e-6)[c]
        h5_write(str(tmp_path / f"airplane_100_{ax}.h5"), "data", data)
    r = subprocess.run([APP, "--print-every", "50"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = h5_read(str(tmp_path / "Ez_data.h5"), h5_dataset_name(99))      # state after 100 steps
    ids = np.zeros((nx, ny, nz), np.uint8)
    ids[30:70, 35:65, 28:72] = 1
    o = OracleSim(p)
    o.set_materials(ids, np.array([[1.0, 1.0, 1.0], [4.0, 2.0, 8.0]]))
    o.step(100)

    class _O:
        def field(self, c):
            return o.field(c)
    want = e2(_O(), p)
    e = float(np.linalg.norm(got - want) / np.linalg.norm(want))
    iso = OracleSim(p)
    iso.set_materials(ids, [1.0, 4.0])
    iso.step(100)

    class _I:
        def field(self, c):
            return iso.field(c)
    d = float(np.linalg.norm(e2(_I(), p) - want) / np.linalg.norm(want))
    say(f"app |E|^2 vs oracle {e:.2e}; isotropic eps 4 vs per-component {d:.2e}")
    assert e <= 2e-5
    assert d > 1e-3


def test_luxum_app_partial_material_files_fail(tmp_path):
    """Two of the three Cb files: the app reports an error instead of
    silently using vacuum for the missing component (the reference aborts
    when it cannot read its input)."""
    from luxum_amd.sim import h5_write, load_yaml
    path = reference_yaml("parameters_3d.yaml", tmp_path / "parameters.yaml", **{"Number of timesteps": 3})
    p = load_yaml(path)
    nx, ny, nz = p.dims
    h5_write(str(tmp_path / "airplane_100_x.h5"), "data", np.zeros((nx - 1, ny, nz)))
    h5_write(str(tmp_path / "airplane_100_y.h5"), "data", np.zeros((nx, ny - 1, nz)))
    r = subprocess.run([APP], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    say(r.stdout[-200:] + r.stderr[-200:])
    assert r.returncode != 0
    assert "airplane_100_z.h5" in r.stderr
