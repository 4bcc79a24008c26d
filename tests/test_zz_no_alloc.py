"""lx_sim_read_field_noalloc (the allocation-free readback) against
lx_sim_read_field through the C ABI: the same gather kernel, so the values
must be identical (bit-exact, as for any FP64 copy), on two calls in a row,
the second on a smaller simulation; a component whose x slice exceeds the
fixed scratch is refused with LX_EINVAL."""
import ctypes as C
import sys

import numpy as np
import pytest

from luxum_amd.params import LX_EY, LX_EZ, LX_FP16, LX_SRC_SOFT, default_params

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def noalloc(sim, comp):
    from luxum_amd._lib import lib
    out = np.empty(sim.p.field_shape(comp), dtype=np.float64)
    st = lib().lx_sim_read_field_noalloc(sim._h, comp, out.ctypes.data, C.c_size_t(out.size))
    return st, out


def dipole(nx, ny, nz, planar=False, precision=0):
    p = default_params(is_planar=int(planar), nx=nx, ny=1 if planar else ny, nz=nz, dy=0.0 if planar else 5e-8,
                       use_ps=1, src_kind=LX_SRC_SOFT, src_comp=LX_EY if planar else LX_EZ, src_amp=1.0,
                       src_lambda0=20 * 5e-8, src_fwhm=1.0, src_delay=1.0, precision=precision,
                       src_pos=(nx // 2, 0 if planar else ny // 2, nz // 2))
    return p


def test_noalloc_readback_matches_read_field(gpu_ctx):
    from luxum_amd.sim import Simulation
    cases = [("3D 12x1100x1050 (four slices per component)", dipole(12, 1100, 1050), 40),
             ("2D 300x1x260, FP16 storage (second call, smaller)", dipole(300, 1, 260, True, LX_FP16), 60)]
    for name, p, steps in cases:
        say(f"case {name}: {steps} steps")
        s = Simulation(p, gpu_ctx)
        s.step(steps)
        s.sync()
        for c in p.components():
            st, got = noalloc(s, c)
            assert st == 0, (name, c, st)
            want = s.field(c)
            assert got.shape == want.shape
            assert np.array_equal(got, want), (name, c)
            if c == p.src_comp:
                assert np.count_nonzero(want) > 0, (name, c)
        say(f"  {name}: every component identical to lx_sim_read_field")
        s.close()


def test_noalloc_readback_refuses_oversized_slice(gpu_ctx):
    from luxum_amd.sim import Simulation
    p = dipole(4, 2100, 2100)
    say("case 3D 4x2100x2100: an x slice of 2100*2100 values exceeds the scratch")
    s = Simulation(p, gpu_ctx)
    st, _ = noalloc(s, LX_EZ)
    assert st == 1                       # LX_EINVAL
    s.close()
