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
        # Cb inputs of maxwell_three_dim_pml.cc:109-125: dataset "data" per
        # component, eps_r = 1 + 1e6*data; a block of eps_r = 4 in the TF box
        from luxum_amd.sim import h5_write
        nx, ny, nz = p.dims
        data = np.zeros((nx - 1, ny, nz))
        data[40:60, 45:55, 38:62] = 3e-6
        for c in "xyz":
            h5_write(str(tmp_path / f"airplane_100_{c}.h5"), "data", data)
        mats = data
    r = subprocess.run([APP], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    say(r.stdout[-300:] + r.stderr[-300:])
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().split("\n")
    prints = [ln for ln in lines if ln.startswith("t: ")]
    assert len(prints) == p.num_timesteps
    assert ("Time taken" in lines[-1]) == timed

    s = Simulation(p, gpu_ctx)
    if mats is not None:
        nx, ny, nz = p.dims
        ids = np.zeros((nx, ny, nz), np.uint8)
        ids[: nx - 1] = (mats != 0)
        s.set_materials(ids, [1.0, 1.0 + mats.max() * 1e6])
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
