"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Tolerance: relative L2 <= 1e-5 for FP32 storage on every field component
and on the probe time series (the FP64 oracle is the reference).  FP16
storage has its own measured bound (test_fp16_bound).
"""
import sys

import numpy as np
import pytest

from conftest import rel_l2
from luxum_amd import configs
from luxum_amd.params import (LX_EX, LX_EY, LX_EZ, LX_FP16, LX_HX, LX_HY, LX_HZ, LX_INC_ANALYTIC,
                              LX_INC_AUX1D, LX_SRC_HARD_CW, LX_SRC_SOFT, default_params)
from oracle.lxoracle import OracleSim

pytestmark = pytest.mark.gpu
TOL32 = 1e-5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_pair(ctx, p, steps, ids=None, eps=None, probes=None):
    from luxum_amd.sim import Simulation
    say(f"case nx={p.nx} ny={p.ny} nz={p.nz} planar={p.is_planar} pml={p.use_pml} pw={p.use_pw} "
        f"inc={p.incident} ps={p.use_ps} prec={p.precision} steps={steps}")
    g = Simulation(p, ctx)
    o = OracleSim(p)
    if ids is not None:
        g.set_materials(ids, eps)
        o.set_materials(ids, eps)
    if probes is not None:
        g.set_probes(probes)
        o.set_probes(probes)
    g.step(steps)
    o.step(steps)
    g.sync()
    return g, o


def field_errors(g, o):
    """Relative L2 of the E and H vector fields (all components jointly)."""
    out = {}
    comps = g.p.components()
    for name, group in (("E", [c for c in comps if c < 3]), ("H", [c for c in comps if c >= 3])):
        num = den = 0.0
        for c in group:
            a, b = g.field(c), o.field(c)
            assert a.shape == b.shape
            assert np.all(np.isfinite(a))
            num += float(np.sum((a - b) ** 2))
            den += float(np.sum(b ** 2))
        out[name] = (num / den) ** 0.5 if den > 0 else num ** 0.5
    return out


def check_fields(g, o, tol, label=""):
    err = field_errors(g, o)
    say(f"{label}: rel L2 E {err['E']:.3e} H {err['H']:.3e}")
    for k, e in err.items():
        assert e <= tol, f"{label} {k}: rel L2 {e:.3e} > {tol:.1e}"
    return max(err.values())


def test_3d_cpml_aux_cube(gpu_ctx):
    w = configs.config1(n=48, steps=300, pml=8, margin=6)
    g, o = run_pair(gpu_ctx, w.p, 300, w.ids, w.eps_r, w.probes)
    worst = check_fields(g, o, TOL32, "3d cpml aux")
    pg, po = g.probes(), o.probes()
    assert pg.shape == po.shape == (300, 8)
    e = rel_l2(pg, po)
    say(f"3d cpml aux: fields {worst:.2e} probes {e:.2e}")
    assert e <= TOL32


def test_probes_on_tfsf_faces(gpu_ctx):
    """Probes of every component on and beside the TF/SF faces (where the
    fused step applies the H corrections before the step and the E ones after
    it) against the oracle's end-of-step records."""
    w = configs.config1(n=48, steps=260, pml=8, margin=6)
    lo, hi, m = w.p.tf_lo[0], w.p.tf_hi[0], 24
    probes = np.array([[LX_EZ, lo, m, m], [LX_EZ, hi, m, m], [LX_HY, lo - 1, m, m], [LX_HY, hi, m, m],
                       [LX_HY, lo, m, m], [LX_EZ, lo - 1, m, m], [LX_EX, m, m, lo], [LX_EX, m, m, hi],
                       [LX_HX, m, lo - 1, m], [LX_HX, m, hi, m], [LX_EY, lo, m, lo], [LX_HZ, lo - 1, lo - 1, m]],
                      dtype=np.int32)
    g, o = run_pair(gpu_ctx, w.p, 260, w.ids, w.eps_r, probes)
    assert g.info().fused_step
    pg, po = g.probes(), o.probes()
    scale = np.abs(po).max()
    for q in range(len(probes)):
        d = np.abs(pg[:, q] - po[:, q]).max() / scale
        say(f"probe {probes[q].tolist()}: max |diff| / max |probe| = {d:.2e}")
        assert d <= 1e-5
    e = rel_l2(pg, po)
    say(f"tfsf-face probes: rel L2 {e:.2e}")
    assert e <= TOL32


def test_3d_plain_analytic_oblique(gpu_ctx):
    p = default_params(nx=40, ny=36, nz=34, use_pw=1, incident=LX_INC_ANALYTIC, off_pw=90.0,
                       pw_sigma=(20 * 5e-8) ** 2, tf_lo=(8, 7, 6), tf_hi=(30, 27, 26))
    p.k_pw[:] = (0.5773502691896258,) * 3
    p.e_pw[:] = (0.70710678118, 0.0, -0.70710678118)
    g, o = run_pair(gpu_ctx, p, 160)
    check_fields(g, o, TOL32, "3d plain analytic")


def test_3d_cpml_analytic_anisotropic(gpu_ctx):
    p = default_params(nx=36, ny=30, nz=40, dx=5e-8, dy=4e-8, dz=6e-8, courant=0.45, use_pml=1,
                       pml_thickness=6 * 5e-8, use_pw=1, incident=LX_INC_ANALYTIC, off_pw=80.0,
                       pw_sigma=(20 * 5e-8) ** 2, tf_lo=(12, 10, 12), tf_hi=(24, 20, 26))
    p.k_pw[:] = (0.0, 0.6, 0.8)
    p.e_pw[:] = (1.0, 0.0, 0.0)
    g, o = run_pair(gpu_ctx, p, 150)
    check_fields(g, o, TOL32, "3d cpml analytic aniso")


def test_2d_cpml_aux(gpu_ctx):
    w = configs.config0(n=96, steps=400)
    g, o = run_pair(gpu_ctx, w.p, 400, w.ids, w.eps_r, w.probes)
    worst = check_fields(g, o, TOL32, "2d cpml aux")
    e = rel_l2(g.probes(), o.probes())
    say(f"2d cpml aux: fields {worst:.2e} probes {e:.2e}")
    assert e <= TOL32


def test_2d_plain_analytic(gpu_ctx):
    p = default_params(is_planar=1, nx=80, ny=1, nz=70, dy=0.0, use_pw=1, incident=LX_INC_ANALYTIC,
                       off_pw=90.0, pw_sigma=(20 * 5e-8) ** 2, tf_lo=(10, 0, 10), tf_hi=(60, 0, 55),
                       courant=0.5, src_comp=LX_EY)
    p.k_pw[:] = (0.70710678118, 0.0, 0.70710678118)
    p.e_pw[:] = (0.0, 1.0, 0.0)
    g, o = run_pair(gpu_ctx, p, 200)
    check_fields(g, o, TOL32, "2d plain analytic")


def test_point_sources(gpu_ctx):
    # reference hard CW source (int amplitude and omega), 3D CPML
    p = default_params(nx=32, ny=30, nz=28, use_pml=1, pml_thickness=6 * 5e-8, use_ps=1,
                       src_kind=LX_SRC_HARD_CW, src_amp=2.7, src_omega=3.0e15, src_pos=(16, 15, 14))
    g, o = run_pair(gpu_ctx, p, 120)
    check_fields(g, o, TOL32, "hard cw")
    # soft Gaussian-modulated dipole
    p2 = default_params(nx=32, ny=30, nz=28, use_pml=1, pml_thickness=6 * 5e-8, use_ps=1,
                        src_kind=LX_SRC_SOFT, src_amp=1.0, src_lambda0=20 * 5e-8, src_fwhm=3.0,
                        src_delay=4.0, src_pos=(15, 14, 13))
    g, o = run_pair(gpu_ctx, p2, 150)
    check_fields(g, o, TOL32, "soft dipole")


def test_2d_soft_source(gpu_ctx):
    p = default_params(is_planar=1, nx=64, ny=1, nz=60, dy=0.0, use_pml=1, pml_thickness=8 * 5e-8, use_ps=1,
                       src_kind=LX_SRC_SOFT, src_comp=LX_EY, src_amp=1.0, src_lambda0=20 * 5e-8,
                       src_fwhm=3.0, src_delay=4.0, src_pos=(30, 0, 28))
    g, o = run_pair(gpu_ctx, p, 200)
    check_fields(g, o, TOL32, "2d soft")


def test_monitor_and_snapshot(gpu_ctx):
    w = configs.config1(n=32, steps=120, pml=6, margin=4)
    g, o = run_pair(gpu_ctx, w.p, 120, w.ids, w.eps_r)
    mx, en = g.monitor()
    ref = o.max_e2()
    assert abs(mx - ref) <= 1e-4 * ref
    snap = np.zeros(w.p.field_shape(LX_EZ))
    g.snapshot_async(LX_EZ, snap)
    g.sync()
    assert np.array_equal(snap, g.field(LX_EZ))


def test_snapshot_pinned_dma(gpu_ctx):
    """Snapshots into pinned host memory (hipHostMalloc) go to the DMA engine
    outside the streams: four in flight over two staging slots, with step
    graphs (CPML shell beside the interior) between them, equal a second
    run's synchronous readbacks at the same steps bit for bit."""
    import ctypes

    from luxum_amd.sim import Simulation
    w = configs.config1(n=48, steps=200, pml=6, margin=4)
    shape = w.p.field_shape(LX_EZ)
    m = int(np.prod(shape))
    hip = ctypes.CDLL("libamdhip64.so")
    ptrs, bufs = [], []
    try:
        for _ in range(4):
            q = ctypes.c_void_p()
            assert hip.hipHostMalloc(ctypes.byref(q), ctypes.c_size_t(8 * m), 0) == 0
            ptrs.append(q)
            bufs.append(np.ctypeslib.as_array(ctypes.cast(q, ctypes.POINTER(ctypes.c_double)), shape=shape))
            bufs[-1][...] = -1.0
        g = Simulation(w.p, gpu_ctx)
        g.set_materials(w.ids, w.eps_r)
        for q in range(4):
            g.step(37)
            g.snapshot_async(LX_EZ, bufs[q])
        g.sync()
        r = Simulation(w.p, gpu_ctx)
        r.set_materials(w.ids, w.eps_r)
        for q in range(4):
            r.step(37)
            ref = r.field(LX_EZ)
            assert np.array_equal(bufs[q], ref), f"snapshot {q}"
        assert float(np.abs(bufs[3]).max()) > 0
        say(f"pinned snapshots: 4 x {m} values equal the readbacks")
        g.close()
        r.close()
    finally:
        for q in ptrs:
            hip.hipHostFree(q)


def test_3d_cpml_aux_2000_steps(gpu_ctx):
    """Config 1's set-up at 128^3 over config 1's 2000 steps, FP32 and FP16
    storage side by side against one FP64 oracle run.  The error grows as the
    field decays (the pulse has left the grid; the cube rings down): FP32
    stays <= 1e-5, FP16 <= 5e-2 on the fields and 1e-2 on the probes."""
    from luxum_amd.sim import Simulation
    w = configs.config1(n=128, steps=2000, pml=12, margin=10)
    w16 = configs.config1(n=128, steps=2000, pml=12, margin=10, precision=LX_FP16)
    gs = {}
    for tag, ww in (("fp32", w), ("fp16", w16)):
        g = Simulation(ww.p, gpu_ctx)
        g.set_materials(ww.ids, ww.eps_r)
        g.set_probes(ww.probes)
        gs[tag] = g
    o = OracleSim(w.p)
    o.set_materials(w.ids, w.eps_r)
    o.set_probes(w.probes)
    for chunk in range(4):
        for g in gs.values():
            g.step(500)
            g.sync()
        o.step(500)
        say(f"128^3 step {(chunk + 1) * 500}: " + " ".join(
            f"{tag} fields {max(field_errors(g, o).values()):.2e} probes {rel_l2(g.probes(), o.probes()):.2e}"
            for tag, g in gs.items()))
    e32, p32 = max(field_errors(gs["fp32"], o).values()), rel_l2(gs["fp32"].probes(), o.probes())
    e16, p16 = max(field_errors(gs["fp16"], o).values()), rel_l2(gs["fp16"].probes(), o.probes())
    assert e32 <= TOL32 and p32 <= TOL32
    assert e16 <= 5e-2 and p16 <= 1e-2


def test_phc_slab_scaled(gpu_ctx):
    """Config 2's structure (square eps 12 rods in a slab, a removed rod row,
    soft Ez dipole at a/lambda 0.3, CPML 16) at half the pitch (16, side 6)
    on 128 x 128 x 80 with a 32-cell slab, FP32 and FP16 storage against one
    FP64 oracle run; one probe on the source cell (its E series is recorded
    before each step's source term).  FP32 <= 1e-5 (measured 8.4e-6 fields,
    2.1e-7 probes); FP16 storage <= 6e-2 on the fields and 1e-2 on the probes
    (measured 4.2e-2, 2.8e-3)."""
    from luxum_amd.sim import Simulation
    steps = 700
    runs = {}
    c = 64
    probes = np.array([[LX_EZ, c + d, c, 40] for d in (0, 8, 16, 32)], dtype=np.int32)
    for tag, prec in (("fp32", 0), ("fp16", LX_FP16)):
        w = configs.config2(nxy=128, nz=80, steps=steps, precision=prec, pitch=16, side=6, slab=32)
        w.p.src_delay = 2.5                       # smooth start, pulse peak near step 480
        g = Simulation(w.p, gpu_ctx)
        g.set_materials(w.ids, w.eps_r)
        g.set_probes(probes)
        g.step(steps)
        g.sync()
        runs[tag] = (w, g)
    w, _ = runs["fp32"]
    o = OracleSim(w.p)
    o.set_materials(w.ids, w.eps_r)
    o.set_probes(probes)
    o.step(steps)
    errs = {}
    for tag, (_, g) in runs.items():
        errs[tag] = (max(field_errors(g, o).values()), rel_l2(g.probes(), o.probes()))
        say(f"phc slab 128x128x80 {tag}: fields {errs[tag][0]:.2e} probes {errs[tag][1]:.2e}")
    assert errs["fp32"][0] <= TOL32 and errs["fp32"][1] <= TOL32
    assert errs["fp16"][0] <= 6e-2 and errs["fp16"][1] <= 1e-2


def test_fp16_bound(gpu_ctx):
    w = configs.config1(n=48, steps=300, pml=8, margin=6, precision=LX_FP16)
    g, o = run_pair(gpu_ctx, w.p, 300, w.ids, w.eps_r, w.probes)
    worst = max(field_errors(g, o).values())
    e = rel_l2(g.probes(), o.probes())
    say(f"fp16 storage: fields {worst:.2e} probes {e:.2e}")
    assert worst <= 2e-2 and e <= 2e-2


def test_chunked_stepping_matches(gpu_ctx):
    """Graph replay (chunks of 32 steps) and eager steps give the same bits."""
    from luxum_amd.sim import Simulation
    w = configs.config1(n=32, steps=100, pml=6, margin=4)
    a = Simulation(w.p, gpu_ctx)
    b = Simulation(w.p, gpu_ctx)
    for s in (a, b):
        s.set_materials(w.ids, w.eps_r)
    a.step(100)
    for _ in range(100):
        b.step(1)
    for c in w.p.components():
        assert np.array_equal(a.field(c), b.field(c))


def test_fused_step_matches_half_steps(gpu_ctx, monkeypatch):
    """The fused 3D timestep (one launch, ping-pong buffers) against the two
    half-step kernels on the same inputs: CPML + aux TF/SF + dielectric, an
    odd number of steps (state in the second buffer), probes; and the oblique
    analytic plane wave.  Same arithmetic, so the tolerance is 1e-6."""
    from luxum_amd.sim import Simulation

    def run(p, steps, w=None, fused=True):
        monkeypatch.setenv("LX_FUSED", "1" if fused else "0")
        s = Simulation(p, gpu_ctx)
        if w is not None:
            s.set_materials(w.ids, w.eps_r)
            s.set_probes(w.probes)
        assert bool(s.info().fused_step) == fused
        s.step(steps)
        s.sync()
        return s

    w = configs.config1(n=48, steps=201, pml=8, margin=6)
    a, b = run(w.p, 201, w, True), run(w.p, 201, w, False)
    num = den = 0.0
    for c in w.p.components():
        x, y = a.field(c), b.field(c)
        num += float(np.sum((x - y) ** 2))
        den += float(np.sum(y ** 2))
    e = (num / den) ** 0.5
    pe = rel_l2(a.probes(), b.probes())
    say(f"fused vs half-step (cpml aux, 201 steps): fields {e:.2e} probes {pe:.2e}")
    assert e <= 1e-6 and pe <= 1e-6
    p = default_params(nx=40, ny=36, nz=44, use_pw=1, incident=LX_INC_ANALYTIC, off_pw=90.0,
                       pw_sigma=(20 * 5e-8) ** 2, tf_lo=(8, 7, 9), tf_hi=(30, 27, 33), courant=0.5)
    p.k_pw[:] = (0.6, 0.48, 0.64)
    p.e_pw[:] = (0.6246950475544243, -0.7808688094430304, 0.0)
    a, b = run(p, 150), run(p, 150, fused=False)
    num = den = 0.0
    for c in p.components():
        x, y = a.field(c), b.field(c)
        num += float(np.sum((x - y) ** 2))
        den += float(np.sum(y ** 2))
    e = (num / den) ** 0.5
    say(f"fused vs half-step (plain, oblique analytic): fields {e:.2e}")
    assert e <= 1e-6


def test_fused_odd_and_thin_grids(gpu_ctx):
    """The default 3D path (the fused step where it applies) on uneven, thin and small grids
    against the FP64 oracle: x extents off the 64-cell chunk, few rows or
    planes (z blocks shorter than kz, slabs covering most of an axis), CPML
    thinner than the slabs' 4-cell x alignment, aux TF/SF and probes."""
    cases = [(33, 13, 12, 2, 3), (70, 12, 23, 2, 3), (129, 14, 12, 2, 3), (20, 30, 44, 3, 2)]
    for nx, ny, nz, pml, margin in cases:
        w = configs.box3d(nx, ny, nz, pml=pml, margin=margin, steps=160)
        w.p.pw_delay = 1.5                       # pulse peak at the TF face by step ~110
        yc, zc = (w.p.tf_lo[1] + w.p.tf_hi[1]) // 2, (w.p.tf_lo[2] + w.p.tf_hi[2]) // 2
        probes = np.array([[LX_EZ, w.p.tf_lo[0] + 2, yc, zc], [LX_EZ, nx // 2, yc, zc],
                           [LX_EY, nx // 2, yc, zc]], dtype=np.int32)   # inside the TF box
        g, o = run_pair(gpu_ctx, w.p, 160, w.ids, w.eps_r, probes)
        path = "fused" if g.info().fused_step else "half-step"   # fused unless an H face lies in a slab
        worst = check_fields(g, o, TOL32, f"{path} {nx}x{ny}x{nz} pml {pml}")
        e = rel_l2(g.probes(), o.probes())
        say(f"{path} {nx}x{ny}x{nz}: fields {worst:.2e} probes {e:.2e}")
        assert e <= TOL32


def test_fused_shell_block_depth(gpu_ctx, monkeypatch):
    """The CPML shell launch in z blocks of its own depth (5, 11, 3 planes
    against the interior's 8) tiles the grid with the interior box exactly:
    fields and probes bit-identical to the shell at the interior's depth."""
    from luxum_amd.sim import Simulation
    w = configs.box3d(72, 44, 50, pml=9, margin=4)

    def run(shell_kz):
        monkeypatch.setenv("LX_FUSED", "1")
        monkeypatch.setenv("LX_SHELL_KZ", str(shell_kz))
        s = Simulation(w.p, gpu_ctx)
        s.set_materials(w.ids, w.eps_r)
        s.set_probes(w.probes)
        assert s.info().shell_kz == (shell_kz or 8)
        s.step(57)
        s.sync()
        return s

    ref = run(0)
    for d in (5, 11, 3):
        say(f"case shell kz {d}")
        s = run(d)
        for c in w.p.components():
            assert np.array_equal(s.field(c), ref.field(c)), f"shell kz {d} component {c}"
        assert np.array_equal(s.probes(), ref.probes())


def test_tfsf_between_steps(gpu_ctx, monkeypatch):
    """One TF/SF launch between fused steps (E corrections of step t, H
    corrections of step t+1, aux line through its second set) against two
    launches per step: fields and probes bit-identical.  Aux and analytic
    incidence, FP32 and FP16, step counts that end a sequence on either aux
    set, graph replay and eager launches, E probes on the TF/SF faces, H
    probes beside them, a point source with probes, and H probes on H face
    cells (the host keeps two launches there)."""
    from luxum_amd.sim import Simulation

    def run(p, steps, merged, graphs, w=None, probes=None):
        monkeypatch.setenv("LX_FUSED", "1")
        monkeypatch.setenv("LX_TFSF_MERGE", "1" if merged else "0")
        monkeypatch.setenv("LX_NO_GRAPH", "0" if graphs else "1")
        s = Simulation(p, gpu_ctx)
        if w is not None:
            s.set_materials(w.ids, w.eps_r)
        if probes is not None:
            s.set_probes(probes)
        assert s.info().fused_step
        s.step(steps)
        s.sync()
        return s

    def same(label, p, steps, w=None, probes=None):
        for graphs in (True, False):
            say(f"case {label}, {steps} steps, graphs {graphs}")
            a = run(p, steps, True, graphs, w, probes)
            b = run(p, steps, False, graphs, w, probes)
            for c in p.components():
                assert np.array_equal(a.field(c), b.field(c)), f"{label} graphs {graphs} component {c}"
            if probes is not None:
                assert np.array_equal(a.probes(), b.probes()), f"{label} graphs {graphs} probes"
            a.close()
            b.close()

    w = configs.config1(n=48, steps=200, pml=8, margin=6)
    lo, hi, m = w.p.tf_lo[0], w.p.tf_hi[0], 24
    pr = np.concatenate([w.probes, np.array([[LX_EZ, lo, m, m], [LX_EX, m, m, hi], [LX_HY, lo - 2, m, m],
                                             [LX_HZ, m, m + 3, m]], dtype=np.int32)])
    for steps in (37, 70, 1):
        same("aux fp32", w.p, steps, w, pr)
    w16 = configs.config1(n=48, steps=200, pml=8, margin=6, precision=LX_FP16)
    same("aux fp16", w16.p, 45, w16, pr)
    hface = np.concatenate([pr, np.array([[LX_HY, lo - 1, m, m]], dtype=np.int32)])
    same("aux, H probe on an H face", w.p, 40, w, hface)
    p = default_params(nx=40, ny=36, nz=44, use_pw=1, incident=LX_INC_ANALYTIC, off_pw=90.0,
                       pw_sigma=(20 * 5e-8) ** 2, tf_lo=(8, 7, 9), tf_hi=(30, 27, 33), courant=0.5)
    p.k_pw[:] = (0.6, 0.48, 0.64)
    p.e_pw[:] = (0.6246950475544243, -0.7808688094430304, 0.0)
    same("analytic", p, 41, None, np.array([[LX_EX, 8, 18, 20], [LX_HY, 20, 18, 20]], dtype=np.int32))
    ps = default_params(nx=36, ny=34, nz=30, use_pml=1, pml_thickness=6 * 5e-8, use_ps=1, src_kind=LX_SRC_SOFT,
                        src_amp=1.0, src_lambda0=20 * 5e-8, src_fwhm=3.0, src_delay=4.0, src_pos=(17, 16, 15))
    same("point source", ps, 35, None, np.array([[LX_EZ, 17, 16, 15], [LX_HX, 17, 18, 15]], dtype=np.int32))


def test_fused_shell_beside_interior(gpu_ctx, monkeypatch):
    """The CPML shell launch on a second stream, running at the same time as
    the interior launch, writes the same fields and probes as the shell
    behind it on one stream (the two read one buffer set and write disjoint
    cells of the other): FP32 one cell per lane and FP16 four cells per
    lane, an odd step count, graph replay and eager launches."""
    from luxum_amd.sim import Simulation

    def run(w, mode, beside, graphs):
        monkeypatch.setenv("LX_FUSED", mode)
        monkeypatch.setenv("LX_SHELL_STREAM", "1" if beside else "0")
        monkeypatch.setenv("LX_NO_GRAPH", "0" if graphs else "1")
        s = Simulation(w.p, gpu_ctx)
        s.set_materials(w.ids, w.eps_r)
        s.set_probes(w.probes)
        assert s.info().shell_concurrent == int(beside)
        assert s.info().cells_per_lane == int(mode)
        s.step(45)
        s.sync()
        return s

    for prec, mode in ((configs.LX_FP32, "1"), (LX_FP16, "4")):
        w = configs.box3d(200, 70, 60, pml=10, margin=5, precision=prec)
        ref = run(w, mode, False, True)
        for graphs in (True, False):
            say(f"case {prec} shell beside, graphs {graphs}")
            s = run(w, mode, True, graphs)
            for c in w.p.components():
                assert np.array_equal(s.field(c), ref.field(c)), f"{prec} graphs {graphs} component {c}"
            assert np.array_equal(s.probes(), ref.probes())
            s.close()
        ref.close()


def test_fused_two_cells_per_lane(gpu_ctx, monkeypatch):
    """The fused step with two x cells per lane (128-cell chunks) against
    the half-step kernels (1e-6) and the FP64 oracle (1e-5): two chunks per
    row with a partial second one (nx = 200, padded pitch 256), and one
    partial chunk on a plain grid (nx = 100); FP16 storage against the
    one-cell-per-lane fused step, with two and with four cells per lane."""
    from luxum_amd.sim import Simulation

    def run(p, steps, w, mode):
        monkeypatch.setenv("LX_FUSED", mode)
        s = Simulation(p, gpu_ctx)
        if w is not None:
            s.set_materials(w[0], w[1])
            s.set_probes(w[2])
        inf = s.info()
        assert bool(inf.fused_step) == (mode != "0")
        assert inf.cells_per_lane == (int(mode) if mode in ("2", "4") else 1)
        s.step(steps)
        s.sync()
        return s

    def rel(a, b, comps):
        num = den = 0.0
        for c in comps:
            x, y = a.field(c), b.field(c)
            assert np.all(np.isfinite(x))
            num += float(np.sum((x - y) ** 2))
            den += float(np.sum(y ** 2))
        return (num / den) ** 0.5

    w = configs.box3d(200, 40, 44)
    p, ids, eps, probes = w.p, w.ids, w.eps_r, w.probes
    say(f"case vec2 cpml aux nx={p.nx} ny={p.ny} nz={p.nz}")
    a, b = run(p, 161, (ids, eps, probes), "2"), run(p, 161, (ids, eps, probes), "0")
    e, pe = rel(a, b, p.components()), rel_l2(a.probes(), b.probes())
    say(f"vec2 vs half-step (cpml aux, 161 steps): fields {e:.2e} probes {pe:.2e}")
    assert e <= 1e-6 and pe <= 1e-6
    o = OracleSim(p)
    o.set_materials(ids, eps)
    o.set_probes(probes)
    o.step(161)
    check_fields(a, o, TOL32, "vec2 vs oracle")

    q = default_params(nx=100, ny=36, nz=44, use_pw=1, incident=LX_INC_ANALYTIC, off_pw=90.0,
                       pw_sigma=(20 * 5e-8) ** 2, tf_lo=(8, 7, 9), tf_hi=(90, 27, 33), courant=0.5)
    q.k_pw[:] = (0.6, 0.48, 0.64)
    q.e_pw[:] = (0.6246950475544243, -0.7808688094430304, 0.0)
    say(f"case vec2 plain oblique nx={q.nx} ny={q.ny} nz={q.nz}")
    a, b = run(q, 120, None, "2"), run(q, 120, None, "0")
    e = rel(a, b, q.components())
    say(f"vec2 vs half-step (plain, oblique analytic): fields {e:.2e}")
    assert e <= 1e-6

    w = configs.box3d(120, 36, 40, precision=LX_FP16)
    p, ids, eps, probes = w.p, w.ids, w.eps_r, w.probes
    say(f"case vec2 fp16 cpml aux nx={p.nx} ny={p.ny} nz={p.nz}")
    a, b = run(p, 150, (ids, eps, probes), "2"), run(p, 150, (ids, eps, probes), "1")
    e = rel(a, b, p.components())
    say(f"fp16 vec2 vs vec1 fused (cpml aux, 150 steps): fields {e:.2e}")
    assert e <= 1e-6

    w = configs.box3d(200, 36, 40, precision=LX_FP16)
    p, ids, eps, probes = w.p, w.ids, w.eps_r, w.probes
    say(f"case vec4 fp16 cpml aux nx={p.nx} ny={p.ny} nz={p.nz}")
    a, b = run(p, 133, (ids, eps, probes), "4"), run(p, 133, (ids, eps, probes), "1")
    e, pe = rel(a, b, p.components()), rel_l2(a.probes(), b.probes())
    say(f"fp16 vec4 vs vec1 fused (cpml aux, 133 steps): fields {e:.2e} probes {pe:.2e}")
    assert e <= 1e-6 and pe <= 1e-6


def test_storage_scale(gpu_ctx):
    """Fields stored times a power of two give the unscaled run's results
    (TF/SF, probes, monitor: the same arithmetic, scaled exactly), and the
    automatic FP16 scale keeps a source of amplitude 1e-6 resolvable, which
    unscaled binary16 (subnormal below 6.1e-5) is not."""
    from luxum_amd.sim import Simulation

    def run(p, w, steps):
        s = Simulation(p, gpu_ctx)
        s.set_materials(w.ids, w.eps_r)
        s.set_probes(w.probes)
        s.step(steps)
        s.sync()
        return s

    w = configs.config1(n=48, steps=160, pml=8, margin=6)
    a = run(w.p, w, 160)
    w.p.storage_scale = 1024.0
    b = run(w.p, w, 160)
    for c in w.p.components():
        assert rel_l2(b.field(c), a.field(c)) <= 1e-12, c
    assert rel_l2(b.probes(), a.probes()) <= 1e-12
    ma, mb = a.monitor(), b.monitor()
    say(f"scaled FP32 run: monitor {mb} vs {ma}")
    assert mb[0] == pytest.approx(ma[0], rel=1e-12) and mb[1] == pytest.approx(ma[1], rel=1e-12)

    small = configs.config2(nxy=96, nz=96, steps=40, precision=LX_FP16)
    small.p.src_pos[:] = (48, 48, 48)
    small.p.src_delay = 0.5
    small.p.src_amp = 1e-6
    small.ids[:] = 0
    o = OracleSim(small.p)
    o.set_materials(small.ids, small.eps_r)
    o.step(40)
    ref = o.field(LX_EZ)
    o.close()
    errs = {}
    for scale in (0.0, 1.0):        # automatic (2^20 here), none
        small.p.storage_scale = scale
        s = Simulation(small.p, gpu_ctx)
        s.set_materials(small.ids, small.eps_r)
        s.step(40)
        s.sync()
        errs[scale] = rel_l2(s.field(LX_EZ), ref)
        s.close()
    say(f"FP16 dipole of amplitude 1e-6 vs FP64 oracle: automatic scale {errs[0.0]:.3e}, unscaled {errs[1.0]:.3e}")
    assert errs[0.0] <= 5e-2


def test_fused_shell_pipelined(gpu_ctx, monkeypatch):
    """The CPML shell launch of a step overlapping the TF/SF launch before it
    (only the interior launch before and after it order the shell) writes
    the same fields and probes as the shell between the interior launch's
    events: FP32 and FP16, an odd step count, graph replay and eager."""
    from luxum_amd.sim import Simulation

    def run(w, pipe, graphs):
        monkeypatch.setenv("LX_FUSED", "1" if w.p.precision == configs.LX_FP32 else "4")
        monkeypatch.setenv("LX_SHELL_STREAM", "1")
        monkeypatch.setenv("LX_SHELL_PIPE", "1" if pipe else "0")
        monkeypatch.setenv("LX_NO_GRAPH", "0" if graphs else "1")
        s = Simulation(w.p, gpu_ctx)
        s.set_materials(w.ids, w.eps_r)
        s.set_probes(w.probes)
        assert s.info().shell_concurrent == 1
        assert s.info().shell_pipelined == int(pipe)
        s.step(45)
        s.step(30)
        s.sync()
        return s

    for prec in (configs.LX_FP32, LX_FP16):
        w = configs.config1(n=96, steps=75, pml=8, margin=10, precision=prec)
        ref = run(w, False, True)
        for graphs in (True, False):
            say(f"case {prec} shell pipelined past the TF/SF launch, graphs {graphs}")
            s = run(w, True, graphs)
            for c in w.p.components():
                assert np.array_equal(s.field(c), ref.field(c)), f"{prec} graphs {graphs} component {c}"
            assert np.array_equal(s.probes(), ref.probes())
            s.close()
        ref.close()


def test_fused_rows_sharing_neighbours(gpu_ctx, monkeypatch):
    """The interior launch with 3 or 7 rows per block sharing their
    neighbour rows through LDS (rows j-1 and j+1 neither read nor recomputed
    inside a block; a halo wavefront computes row j0-1) writes the same
    fields and probes as one unit per wavefront, bit for bit: FP32 one and
    two cells per lane, FP16 four, a CPML grid whose interior rows are not a
    multiple of the group (a partial last group), a grid without CPML (first
    group at row 0), odd step counts, graph replay and eager launches."""
    from luxum_amd.sim import Simulation

    def run(w, mode, rows, graphs):
        monkeypatch.setenv("LX_FUSED", mode)
        monkeypatch.setenv("LX_ROWS", str(rows))
        monkeypatch.setenv("LX_NO_GRAPH", "0" if graphs else "1")
        s = Simulation(w.p, gpu_ctx)
        s.set_materials(w.ids, w.eps_r)
        if w.probes is not None:
            s.set_probes(w.probes)
        assert s.info().fused_step == 1
        assert s.info().interior_rows == rows
        s.step(37)
        s.step(8)
        s.sync()
        return s

    cases = [("fp32 v1", configs.box3d(200, 70, 60, pml=10, margin=5), "1"),
             ("fp32 v2", configs.box3d(200, 70, 60, pml=10, margin=5), "2"),
             ("fp16 v4", configs.box3d(200, 70, 60, pml=10, margin=5, precision=LX_FP16), "4"),
             ("fp32 v1 cube", configs.config1(n=64, steps=45, pml=8, margin=6), "1")]
    p = default_params(nx=90, ny=45, nz=50, use_pw=1, incident=LX_INC_ANALYTIC, off_pw=90.0,
                       pw_sigma=(20 * 5e-8) ** 2, tf_lo=(8, 7, 9), tf_hi=(80, 36, 40), courant=0.5)
    p.k_pw[:] = (0.6, 0.48, 0.64)
    p.e_pw[:] = (0.6246950475544243, -0.7808688094430304, 0.0)
    cases.append(("fp32 v1 no CPML", configs.Workload("plain", p, np.zeros((90, 45, 50), np.uint8), np.array([1.0]),
                                                       np.array([[LX_EX, 40, 20, 25], [LX_HY, 41, 20, 25]],
                                                                dtype=np.int32)), "1"))
    for name, w, mode in cases:
        ref = run(w, mode, 0, True)
        for rows in (3, 7):
            for graphs in ((True, False) if rows == 7 else (True,)):
                say(f"case {name}: rows {rows}, graphs {graphs}")
                s = run(w, mode, rows, graphs)
                for c in w.p.components():
                    assert np.array_equal(s.field(c), ref.field(c)), f"{name} rows {rows} graphs {graphs} comp {c}"
                if w.probes is not None:
                    assert np.array_equal(s.probes(), ref.probes())
                s.close()
        ref.close()


def test_snapshot_pinned_every_component(gpu_ctx):
    """Every component of a 3D FP32, a 3D FP16 and a 2D simulation through
    the DMA-engine snapshot path (pinned buffers, both staging slots reused)
    equals its synchronous readback."""
    import ctypes

    from luxum_amd.sim import Simulation
    hip = ctypes.CDLL("libamdhip64.so")
    cases = [configs.config1(n=40, steps=80, pml=6, margin=4),
             configs.config1(n=40, steps=80, pml=6, margin=4, precision=LX_FP16),
             configs.config0(n=96, steps=80)]
    for w in cases:
        w.p.pw_delay = 1.0
        g = Simulation(w.p, gpu_ctx)
        g.set_materials(w.ids, w.eps_r)
        g.step(50)
        comps = list(w.p.components())
        ptrs, bufs = [], []
        try:
            for c in comps:
                shape = w.p.field_shape(c)
                q = ctypes.c_void_p()
                assert hip.hipHostMalloc(ctypes.byref(q), ctypes.c_size_t(8 * int(np.prod(shape))), 0) == 0
                ptrs.append(q)
                bufs.append(np.ctypeslib.as_array(ctypes.cast(q, ctypes.POINTER(ctypes.c_double)), shape=shape))
                bufs[-1][...] = np.nan
            for c, b in zip(comps, bufs):
                g.snapshot_async(c, b)
            g.sync()
            for c, b in zip(comps, bufs):
                assert np.array_equal(b, g.field(c)), (w.name, c)
            say(f"{w.name}: {len(comps)} components through pinned snapshots equal the readbacks")
        finally:
            g.close()
            for q in ptrs:
                hip.hipHostFree(q)
