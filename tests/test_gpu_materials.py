"""Per-component materials and the FP64 monitor against the oracle.

The reference builds one Cb array per E component from three files
(airplane_100_{x,y,z}.h5, maxwell_three_dim_pml.cc:109-128), so Ex, Ey and
Ez at one storage index may see different permittivities.  The device keeps
one uint8 id per node and one coefficient table per component
(lx_sim_set_materials3).  Tolerance: relative L2 <= 1e-5 for FP32 storage
against the FP64 oracle on the E and H fields and the probe series; the
batched fused step equals the single-simulation one to 1e-6; FP16 storage
within its measured bound (5e-2).  The monitor's energy is
sum(eps0*eps_r*E^2 + mu0*H^2)/2 * cell volume, compared with the oracle's
to 1e-5, and its max|E|^2 likewise.
"""
import sys

import numpy as np
import pytest

from conftest import rel_l2
from luxum_amd import configs
from luxum_amd.params import LX_EY, LX_FP16
from oracle.lxoracle import OracleSim

pytestmark = pytest.mark.gpu
TOL32 = 1e-5

# eps_r of (Ex, Ey, Ez) per id: vacuum, the cube, a slab, scattered voxels
EPS3 = np.array([[1.0, 1.0, 1.0], [6.0, 2.5, 9.0], [3.0, 8.0, 1.5], [12.0, 1.0, 4.0]])


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def aniso(n=40, steps=260, precision=None):
    w = configs.config1(n=n, steps=steps, pml=6, margin=4) if precision is None else \
        configs.config1(n=n, steps=steps, pml=6, margin=4, precision=precision)
    ids = w.ids.copy()
    ids[n // 6:n // 6 + 3, :, :] = 2
    rng = np.random.default_rng(5)
    vox = rng.random(ids.shape) < 0.02
    ids[vox] = 3
    w.ids = ids
    w.eps_r = EPS3
    return w


def field_err(a, b):
    num = {"E": 0.0, "H": 0.0}
    den = {"E": 0.0, "H": 0.0}
    for c in a.p.components():
        x, y = a.field(c), b.field(c)
        assert x.shape == y.shape and np.all(np.isfinite(x))
        k = "E" if c < 3 else "H"
        num[k] += float(np.sum((x - y) ** 2))
        den[k] += float(np.sum(y ** 2))
    return max((num[k] / den[k]) ** 0.5 for k in num)


def make(w, ctx):
    from luxum_amd.sim import Simulation
    s = Simulation(w.p, ctx)
    s.set_materials(w.ids, w.eps_r)
    s.set_probes(w.probes)
    return s


def oracle(w, steps):
    o = OracleSim(w.p)
    o.set_materials(w.ids, w.eps_r)
    o.set_probes(w.probes)
    o.step(steps)
    return o


@pytest.mark.parametrize("mode", ["fused", "half"])
def test_per_component_materials_vs_oracle(gpu_ctx, monkeypatch, mode):
    """Three different permittivities per id on the fused step and on the
    half-step pair; the same run with one permittivity per id (the Ex column)
    differs from the oracle by O(1e-1), so the columns are in use."""
    monkeypatch.setenv("LX_FUSED", "1" if mode == "fused" else "0")
    w = aniso()
    say(f"per-component materials, {mode}: 40^3 CPML aux TF/SF, 4 ids, 260 steps")
    g = make(w, gpu_ctx)
    assert bool(g.info().fused_step) == (mode == "fused")
    g.step(260)
    g.sync()
    o = oracle(w, 260)
    e, pe = field_err(g, o), rel_l2(g.probes(), o.probes())
    say(f"{mode}: fields {e:.2e} probes {pe:.2e}")
    assert e <= TOL32 and pe <= TOL32
    w1 = aniso()
    w1.eps_r = EPS3[:, 0].copy()
    iso = oracle(w1, 260)
    d = field_err(iso, o)
    say(f"isotropic (Ex column) vs per-component oracle: {d:.2e}")
    assert d > 1e-2
    g.close()


def test_per_component_materials_fp16(gpu_ctx):
    w = aniso(precision=LX_FP16)
    say("per-component materials, FP16 storage")
    g = make(w, gpu_ctx)
    g.step(260)
    g.sync()
    o = oracle(w, 260)
    e, pe = field_err(g, o), rel_l2(g.probes(), o.probes())
    say(f"fp16: fields {e:.2e} probes {pe:.2e} (cells per lane {g.info().cells_per_lane})")
    assert e <= 5e-2 and pe <= 1e-2
    g.close()


def test_per_component_materials_batch_member(gpu_ctx):
    """A member with per-component materials in the batched fused step
    equals its single-simulation run."""
    from luxum_amd.sim import Batch
    ws = [aniso(), aniso(n=48)]
    members = [make(w, gpu_ctx) for w in ws]
    b = Batch(gpu_ctx)
    for s in members:
        b.add(s)
    say("per-component materials: batch of 2 (40^3, 48^3), 77 steps")
    b.step(77)
    b.sync()
    for w, s in zip(ws, members):
        ref = make(w, gpu_ctx)
        ref.step(77)
        ref.sync()
        e, pe = field_err(s, ref), rel_l2(s.probes(), ref.probes())
        say(f"{w.name}: batch vs single {e:.2e} probes {pe:.2e}")
        assert e <= 1e-6 and pe <= 1e-6
        ref.close()
    b.close()


@pytest.mark.parametrize("planar", [False, True])
def test_monitor_energy_vs_oracle(gpu_ctx, planar):
    """lx_sim_monitor's FP64 max|E|^2 and energy (eps_r-weighted) against
    lxo_max_e2 / lxo_energy after the pulse has entered the dielectric."""
    if planar:
        w = configs.config0(n=160, steps=300)
        steps = 300
    else:
        w = aniso()
        steps = 200
    say(f"monitor energy vs oracle: {'2D' if planar else '3D'}, {steps} steps")
    g = make(w, gpu_ctx)
    g.step(steps)
    g.sync()
    mx, en = g.monitor()
    o = oracle(w, steps)
    omx, oen = o.max_e2(), o.energy()
    say(f"max|E|^2 {mx:.6e} vs {omx:.6e}; energy {en:.6e} vs {oen:.6e}")
    assert oen > 0 and omx > 0
    assert abs(en - oen) <= TOL32 * oen
    assert abs(mx - omx) <= 2 * TOL32 * omx
    g.close()


@pytest.mark.parametrize("planar", [False, True], ids=["3d", "2d"])
def test_monitor_reduction_exact(gpu_ctx, planar):
    """The monitor's reduction against numpy on the same run's read-back
    fields (no oracle tolerance in between): max|E|^2 over i < nx-1, j < ny-1,
    k < nz-1 at equal indices (3D; every Ey cell in 2D) equal to 1e-12, the
    energy sum(eps_r*E^2 + (eta0*H)^2)/2 * eps0 * cell volume to 1e-12, and
    a second call bit-identical to the first (fixed partition and reduction
    order); lx_sim_max_e2 (the E components alone) returns the same max."""
    from luxum_amd.params import LX_EX, LX_EZ, LX_HX, LX_HY, LX_HZ
    w = configs.config0(n=160, steps=300) if planar else aniso(n=40, steps=200)
    steps = 300 if planar else 200
    say(f"monitor reduction: {'2D' if planar else '3D'}, {steps} steps")
    g = make(w, gpu_ctx)
    info = g.info()
    g.step(steps)
    g.sync()
    mx, en = g.monitor()
    mx2, en2 = g.monitor()
    assert (mx, en) == (mx2, en2)
    assert g.max_e2() == mx   # the E-only pass: the same maximum
    p = w.p
    nx, ny, nz = p.dims
    eps = np.asarray(w.eps_r, dtype=np.float64)
    eps3 = eps if eps.ndim == 2 else np.repeat(eps[:, None], 3, axis=1)
    # the device weighs E by cb0/cb of its FP32 coefficient tables
    cb = (info.dt / (info.eps0 * eps3) / (info.eta0 * p.dx)).astype(np.float32).astype(np.float64)
    eps3 = (info.c0 * info.dt / p.dx) / cb
    ids = w.ids
    if planar:
        ey = g.field(LX_EY)
        hx, hz = g.field(LX_HX), g.field(LX_HZ)
        ref_mx = float(np.max(ey ** 2))
        e_sum = float(np.sum(ey ** 2 * eps3[ids.reshape(ey.shape), 1]))
        h_sum = float(np.sum((info.eta0 * hx) ** 2) + np.sum((info.eta0 * hz) ** 2))
        vol = p.dx * p.dz
    else:
        es = [g.field(c) for c in (LX_EX, LX_EY, LX_EZ)]
        hs = [g.field(c) for c in (LX_HX, LX_HY, LX_HZ)]
        ref_mx = float(np.max(sum(e[:nx - 1, :ny - 1, :nz - 1] ** 2 for e in es)))
        e_sum = 0.0
        for c, e in enumerate(es):
            sx, sy, sz = e.shape
            e_sum += float(np.sum(e ** 2 * eps3[ids[:sx, :sy, :sz], c]))
        h_sum = float(sum(np.sum((info.eta0 * h) ** 2) for h in hs))
        vol = p.dx * p.dy * p.dz
    ref_en = 0.5 * info.eps0 * (e_sum + h_sum) * vol
    say(f"max|E|^2 {mx:.15e} vs {ref_mx:.15e}; energy {en:.15e} vs {ref_en:.15e}")
    assert ref_mx > 0 and ref_en > 0
    assert abs(mx - ref_mx) <= 1e-12 * ref_mx
    assert abs(en - ref_en) <= 1e-12 * ref_en
    g.close()


def test_monitor_unstable_guard(gpu_ctx):
    """The non-finite guard: a 3D grid stepped above its Courant limit (0.9
    against 1/sqrt(3)) grows until its fields overflow (FP32 inf, then NaN;
    no trap).  With monitor_every = 25, lx_sim_step returns LX_EUNSTABLE at
    the first check after that, and lx_sim_max_e2 and lx_sim_monitor report
    it too; a stable run with the same guard steps through."""
    from luxum_amd._lib import LX_EUNSTABLE, LuxumError
    from luxum_amd.params import LX_SRC_SOFT, default_params
    from luxum_amd.sim import Simulation
    p = default_params(nx=40, ny=36, nz=32, use_pml=1, pml_thickness=6 * 5e-8, use_ps=1, src_kind=LX_SRC_SOFT,
                       src_amp=1.0, src_lambda0=20 * 5e-8, src_fwhm=3.0, src_delay=4.0, src_pos=(20, 18, 16),
                       courant=0.9)
    p.monitor_every = 25
    s = Simulation(p, gpu_ctx)
    steps = 0
    with pytest.raises(LuxumError) as e:
        for _ in range(200):       # 5000 steps at most; growth overflows FP32 long before
            s.step(25)
            steps += 25
    say(f"unstable run stopped after {steps + 25} steps: {e.value}")
    assert e.value.status == LX_EUNSTABLE
    for call in (s.max_e2, s.monitor):
        with pytest.raises(LuxumError) as e2:
            call()
        assert e2.value.status == LX_EUNSTABLE
    s.close()
    p.courant = 0.5
    s = Simulation(p, gpu_ctx)
    s.step(500)
    s.sync()
    mx = s.max_e2()
    assert np.isfinite(mx) and mx > 0
    s.close()
