"""The started monitor (lx_sim_max_e2_start / _finish): the reference's
per-step max|E|^2 (maxwell_three_dim_pml.cc:266-276, printed at the top of
the next iteration, :166-173) computed beside the next step.  Every value
must equal the synchronous lx_sim_max_e2 of the same state bit for bit, and
the fields must be those of a run without it, on the paths that overlap (one
fused 3D step, FP32 and FP16) and on those that wait (point source, 2D,
several steps, a batch step).  Also: an LX_TUNE pin applies over the
autotuner's cached result for the same grid."""
import sys

import numpy as np
import pytest

from luxum_amd import configs
from luxum_amd.params import LX_EY, LX_EZ, LX_FP16, LX_SRC_SOFT, default_params

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make(w, ctx):
    from luxum_amd.sim import Simulation
    s = Simulation(w.p, ctx)
    s.set_materials(w.ids, w.eps_r)
    if w.probes is not None and len(w.probes):
        s.set_probes(w.probes)
    return s


def point_source_3d():
    p = default_params(nx=40, ny=36, nz=32, use_pml=1, pml_thickness=6 * 5e-8, use_ps=1, src_kind=LX_SRC_SOFT,
                       src_comp=LX_EZ, src_amp=1.0, src_lambda0=20 * 5e-8, src_fwhm=3.0, src_delay=2.0,
                       src_pos=(20, 18, 16), num_timesteps=120)
    return configs.Workload("3D soft dipole", p, np.zeros((40, 36, 32), np.uint8), np.array([1.0]), None)


CASES = {
    "3d_fused_fp32": lambda: configs.config1(n=48, steps=120, pml=6, margin=4),
    "3d_fused_fp16": lambda: configs.config1(n=48, steps=120, pml=6, margin=4, precision=LX_FP16),
    "3d_point_source": point_source_3d,
    "2d_cpml": lambda: configs.config0(n=96, steps=120),
}


@pytest.mark.parametrize("name", list(CASES))
def test_started_monitor_matches_synchronous(gpu_ctx, name):
    w = CASES[name]()
    w.p.pw_delay = 1.0
    steps = 60
    a, b = make(w, gpu_ctx), make(w, gpu_ctx)
    say(f"{name}: {steps} steps, started monitor beside the next step vs lx_sim_max_e2 after each step")
    got, ref = [], []
    a.step(1)
    for t in range(1, steps):
        a.max_e2_start()          # state after step t-1 ... t
        a.step(1)                 # runs beside it where the path allows
        got.append(a.max_e2_finish())
    for t in range(1, steps):
        b.step(1)
        ref.append(b.max_e2())
    b.step(1)
    say(f"{name}: last max {got[-1]:.6e}, largest {max(got):.6e}")
    assert got == ref
    assert max(got) > 0
    comps = [LX_EY] if w.p.is_planar else list(w.p.components())
    for c in comps:
        assert np.array_equal(a.field(c), b.field(c)), c
    a.close()
    b.close()


def test_started_monitor_across_longer_steps_and_a_batch(gpu_ctx):
    """A started monitor followed by a several-step call, then by a batch
    step of the simulation as a member: each value is that of the state at
    its start."""
    from luxum_amd.sim import Batch
    w = configs.config1(n=48, steps=120, pml=6, margin=4)
    w.p.pw_delay = 1.0
    a, b = make(w, gpu_ctx), make(w, gpu_ctx)
    a.step(20)
    b.step(20)
    a.max_e2_start()
    a.step(7)
    m1 = a.max_e2_finish()
    r1 = b.max_e2()
    b.step(7)
    assert m1 == r1
    bt = Batch(gpu_ctx)
    bt.add(a)
    a.max_e2_start()
    bt.step(9)
    bt.sync()
    m2 = a.max_e2_finish()
    r2 = b.max_e2()
    b.step(9)
    say(f"after 20 steps {m1:.6e}, after 27 {m2:.6e}")
    assert m2 == r2 and m2 > 0
    # the batched step rounds as the single-simulation step does (3D members)
    for c in w.p.components():
        assert np.array_equal(a.field(c), b.field(c)), c
    with pytest.raises(Exception):
        a.max_e2_finish()         # nothing started
    bt.close()
    a.close()
    b.close()


def test_autotune_pin_over_cached_result(gpu_ctx, monkeypatch):
    """LX_TUNE pins the settings even after the same grid was autotuned in
    this process (the cached result does not override the pin), and the
    pinned run's fields equal the tuned run's (every variant is bit-identical)."""
    w = configs.config1(n=48, steps=80, pml=6, margin=4)
    monkeypatch.delenv("LX_TUNE", raising=False)
    a = make(w, gpu_ctx)
    a.autotune()
    ia = a.info()
    pin = (1, 8, 1, 5, 0, 3) if (ia.kz, ia.interior_rows) != (5, 3) else (1, 4, 1, 6, 1, 0)
    monkeypatch.setenv("LX_TUNE", ",".join(map(str, pin)))
    b = make(w, gpu_ctx)
    b.autotune()
    ib = b.info()
    say(f"tuned kz {ia.kz} rows {ia.interior_rows} vec {ia.cells_per_lane}; pinned kz {ib.kz} rows {ib.interior_rows} "
        f"vec {ib.cells_per_lane}")
    assert (ib.kz, ib.waves, ib.cells_per_lane, ib.interior_rows) == (pin[3], pin[1], pin[2], pin[5])
    a.step(40)
    b.step(40)
    for c in w.p.components():
        assert np.array_equal(a.field(c), b.field(c)), c
    a.close()
    b.close()
