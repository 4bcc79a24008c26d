"""Batched simulations: the temporally tiled 2D kernel over a flattened tile
list (several planar simulations of uneven sizes in one launch) and the
batched fused 3D step over a flattened block list, against the
single-simulation kernels and the FP64 oracle.

Tolerance: relative L2 <= 1e-5 (FP32) of the E and H fields and of the probe
time series; FP16 storage <= 2e-2 against the oracle.
"""
import sys

import numpy as np
import pytest

from conftest import rel_l2
from luxum_amd import configs
from luxum_amd.params import LX_EY, LX_FP16, LX_INC_ANALYTIC, LX_SRC_SOFT, default_params
from oracle.lxoracle import OracleSim

pytestmark = pytest.mark.gpu
TOL32 = 1e-5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def field_err(a, b):
    """Relative L2 of the E and H vector fields of two simulations."""
    out = {}
    comps = a.p.components()
    for name, group in (("E", [c for c in comps if c < 3]), ("H", [c for c in comps if c >= 3])):
        num = den = 0.0
        for c in group:
            x, y = a.field(c), b.field(c)
            assert x.shape == y.shape and np.all(np.isfinite(x))
            num += float(np.sum((x - y) ** 2))
            den += float(np.sum(y ** 2))
        out[name] = (num / den) ** 0.5 if den > 0 else num ** 0.5
    return max(out.values())


def make(w, ctx):
    from luxum_amd.sim import Simulation
    s = Simulation(w.p, ctx)
    s.set_materials(w.ids, w.eps_r)
    if w.probes is not None and len(w.probes):
        s.set_probes(w.probes)
    return s


def phc(n, seed, precision=None, delay=2.0):
    w = configs.phc2d(n, seed, steps=400) if precision is None else configs.phc2d(n, seed, steps=400, precision=precision)
    w.p.pw_delay = delay   # pulse centre at step ~72*delay
    return w


def thick_pml(n):
    """TF/SF plane wave with a 30-cell CPML (a slab wider than the tile halo)."""
    w = phc(n, 11)
    w.p.pml_thickness = 30 * w.p.dx
    return w


def no_pml(n):
    """TF/SF plane wave in a PEC box without CPML."""
    w = phc(n, 12)
    w.p.use_pml = 0
    return w


def soft_2d():
    p = default_params(is_planar=1, nx=96, ny=1, nz=90, dy=0.0, use_pml=1, pml_thickness=8 * 5e-8, use_ps=1,
                       src_kind=LX_SRC_SOFT, src_comp=LX_EY, src_amp=1.0, src_lambda0=20 * 5e-8,
                       src_fwhm=3.0, src_delay=2.0, src_pos=(40, 0, 45), num_timesteps=400)
    return configs.Workload("2D soft source", p, np.zeros((96, 1, 90), np.uint8), np.array([1.0]),
                            np.array([[LX_EY, 50, 0, 45]], np.int32))


def test_batch_matches_single_and_oracle(gpu_ctx):
    from luxum_amd.sim import Batch
    ws = [phc(128, 1), phc(176, 2), phc(208, 3), phc(256, 4), thick_pml(200), no_pml(150), soft_2d(),
          configs.config1(n=32, steps=400, pml=6, margin=4)]
    b = Batch(gpu_ctx)
    members = [make(w, gpu_ctx) for w in ws]
    for s in members:
        b.add(s)
    say("batch: 8 members (6 tiled 2D, 1 point-source 2D, 1 3D), steps 37 + 200")
    b.step(37)
    b.step(200)
    b.sync()
    for w, s in zip(ws, members):
        ref = make(w, gpu_ctx)
        ref.step(237)
        ref.sync()
        e = field_err(s, ref)
        pe = rel_l2(s.probes(), ref.probes())
        say(f"{w.name}: batch vs single fields {e:.2e} probes {pe:.2e}")
        assert e <= TOL32 and pe <= TOL32
        assert s.info().step == 237
    o = OracleSim(ws[0].p)
    o.set_materials(ws[0].ids, ws[0].eps_r)
    o.set_probes(ws[0].probes)
    o.step(237)
    e = field_err(members[0], o)
    pe = rel_l2(members[0].probes(), o.probes())
    say(f"{ws[0].name}: batch vs oracle fields {e:.2e} probes {pe:.2e}")
    assert e <= TOL32 and pe <= TOL32
    b.close()


def test_batch_interleaved_with_single_steps(gpu_ctx):
    """Members keep stepping alone between batch steps and after removal; the
    ping-pong buffers stay consistent (odd chunk counts, graph replay from
    either parity)."""
    from luxum_amd.sim import Batch
    ws = [phc(160, 7, delay=1.0), phc(144, 8, delay=1.0)]
    members = [make(w, gpu_ctx) for w in ws]
    b = Batch(gpu_ctx)
    for s in members:
        b.add(s)
    say("interleave: batch 64, single 5, batch 100, remove, single 11, batch 43")
    b.step(64)
    members[0].step(5)
    b.step(100)
    b.remove(members[1])
    members[1].step(11)
    b.step(43)
    b.sync()
    counts = [64 + 5 + 100 + 43, 64 + 100 + 11]
    for w, s, n in zip(ws, members, counts):
        ref = make(w, gpu_ctx)
        ref.step(n)
        ref.sync()
        e = field_err(s, ref)
        pe = rel_l2(s.probes(), ref.probes())
        say(f"{w.name}: {n} steps, batch/single mix vs single fields {e:.2e} probes {pe:.2e}")
        assert e <= TOL32 and pe <= TOL32
    b.close()


def test_batch_fp16_bound(gpu_ctx):
    from luxum_amd.sim import Batch
    w = phc(192, 5, precision=LX_FP16)
    s = make(w, gpu_ctx)
    b = Batch(gpu_ctx)
    b.add(s)
    b.step(240)
    b.sync()
    o = OracleSim(w.p)
    o.set_materials(w.ids, w.eps_r)
    o.set_probes(w.probes)
    o.step(240)
    e = field_err(s, o)
    pe = rel_l2(s.probes(), o.probes())
    say(f"fp16 batch vs oracle fields {e:.2e} probes {pe:.2e}")
    assert e <= 2e-2 and pe <= 2e-2
    b.close()


def plain_oblique_3d():
    """3D grid without CPML, analytic oblique plane wave, vacuum."""
    p = default_params(nx=40, ny=36, nz=44, use_pw=1, incident=LX_INC_ANALYTIC, off_pw=90.0,
                       pw_sigma=(20 * 5e-8) ** 2, tf_lo=(8, 7, 9), tf_hi=(30, 27, 33), courant=0.5)
    p.k_pw[:] = (0.6, 0.48, 0.64)
    p.e_pw[:] = (0.6246950475544243, -0.7808688094430304, 0.0)
    return configs.Workload("3D plain oblique 40x36x44", p, np.zeros((40, 36, 44), np.uint8), np.array([1.0]),
                            None)


def test_batch_3d_members(gpu_ctx):
    """3D members of uneven sizes advance in the batched fused step (FP32 and
    FP16 groups, CPML + aux TF/SF + probes, a grid without CPML, a planar
    member beside them); odd step counts, a member removed and stepped alone,
    then the batch again.  Same arithmetic as the single-simulation fused
    step: tolerance 1e-6."""
    from luxum_amd.sim import Batch
    ws = [configs.box3d(72, 40, 36), configs.config1(n=40, steps=400, pml=6, margin=4), plain_oblique_3d(),
          configs.box3d(100, 36, 40, precision=LX_FP16), configs.box3d(48, 44, 40, precision=LX_FP16),
          phc(128, 1)]
    members = [make(w, gpu_ctx) for w in ws]
    b = Batch(gpu_ctx)
    for s in members:
        b.add(s)
    say("batch 3d: 6 members (3 FP32 3D, 2 FP16 3D, 1 tiled 2D): batch 37, batch 100, remove #1, "
        "single 5, batch 21")
    b.step(37)
    b.step(100)
    b.remove(members[1])
    members[1].step(5)
    b.step(21)
    b.sync()
    counts = [158, 142, 158, 158, 158, 158]
    for w, s, n in zip(ws, members, counts):
        ref = make(w, gpu_ctx)
        ref.step(n)
        ref.sync()
        e = field_err(s, ref)
        msg = f"{w.name}: {n} steps, batch vs single fields {e:.2e}"
        if w.probes is not None:
            pe = rel_l2(s.probes(), ref.probes())
            msg += f" probes {pe:.2e}"
            assert pe <= 1e-6
        say(msg)
        assert e <= 1e-6
        assert s.info().step == n
    o = OracleSim(ws[0].p)
    o.set_materials(ws[0].ids, ws[0].eps_r)
    o.set_probes(ws[0].probes)
    o.step(158)
    e = field_err(members[0], o)
    pe = rel_l2(members[0].probes(), o.probes())
    say(f"{ws[0].name}: batch vs oracle fields {e:.2e} probes {pe:.2e}")
    assert e <= TOL32 and pe <= TOL32
    b.close()


def test_batch_members_change_between_steps(gpu_ctx):
    """Members change their device buffers between batch steps while staying
    in the batch: a 3D member steps an odd count alone (its buffer pair swaps
    order) and a planar member outgrows its probe buffer by stepping alone
    and again inside a batch step.  The batch must pick up the new buffers."""
    from luxum_amd.sim import Batch
    w3 = configs.box3d(64, 40, 36, steps=30)
    w2 = phc(160, 7, delay=1.0)
    w2.p.num_timesteps = 40
    ws = [w3, w2]
    members = [make(w, gpu_ctx) for w in ws]
    b = Batch(gpu_ctx)
    for s in members:
        b.add(s)
    say("members change: batch 20, 3D alone 5, 2D alone 30, batch 41")
    b.step(20)
    members[0].step(5)
    members[1].step(30)
    b.step(41)
    b.sync()
    counts = [66, 91]
    for w, s, n in zip(ws, members, counts):
        ref = make(w, gpu_ctx)
        ref.step(n)
        ref.sync()
        e = field_err(s, ref)
        pe = rel_l2(s.probes(), ref.probes())
        say(f"{w.name}: {n} steps, fields {e:.2e} probes {pe:.2e}")
        assert s.probes().shape == ref.probes().shape
        assert e <= 1e-6 and pe <= 1e-6
        assert s.info().step == n
    b.close()


def test_batch_3d_fp16_four_cells(gpu_ctx):
    """FP16 3D members whose row pitch allows four cells per lane advance in
    the batched fused step at that width: bit-identical to their single
    runs (which take the same width by default)."""
    from luxum_amd.sim import Batch
    ws = [configs.box3d(200, 36, 40, precision=LX_FP16), configs.box3d(240, 32, 36, precision=LX_FP16)]
    members = [make(w, gpu_ctx) for w in ws]
    assert all(s.info().cells_per_lane == 4 for s in members)
    b = Batch(gpu_ctx)
    for s in members:
        b.add(s)
    say("batch fp16 four cells: 2 members, batch 45")
    b.step(45)
    b.sync()
    for w, s in zip(ws, members):
        ref = make(w, gpu_ctx)
        ref.step(45)
        ref.sync()
        e = field_err(s, ref)
        pe = rel_l2(s.probes(), ref.probes())
        say(f"{w.name}: fields {e:.2e} probes {pe:.2e}")
        assert e <= 1e-6 and pe <= 1e-6
    b.close()


@pytest.mark.parametrize("how", ["remove", "close"])
def test_batch_member_after_unsynced_remove_or_close(gpu_ctx, how):
    """A member leaves the batch right after a batch step, without a sync:
    it must not wait on the batch's (destroyed, or later unrelated) step
    event, and its own next steps continue from the batch's state."""
    from luxum_amd.sim import Batch
    ws = [configs.box3d(64, 40, 36, steps=60), configs.box3d(56, 48, 40, steps=60)]
    members = [make(w, gpu_ctx) for w in ws]
    b = Batch(gpu_ctx)
    for s in members:
        b.add(s)
    say(f"batch of 2 3D members: step 20, {how} member 0 without sync, member 0 steps 7 alone")
    b.step(20)
    if how == "remove":
        b.remove(members[0])
        b.step(9)                     # the batch goes on with member 1
    else:
        b.close()
    members[0].step(7)
    members[0].sync()
    ref = make(ws[0], gpu_ctx)
    ref.step(27)
    ref.sync()
    e = field_err(members[0], ref)
    say(f"member 0 after {how}: vs single {e:.2e}")
    assert e <= 1e-6
    if how == "remove":
        b.sync()
        ref1 = make(ws[1], gpu_ctx)
        ref1.step(29)
        ref1.sync()
        assert field_err(members[1], ref1) <= 1e-6
        b.close()


def test_batch_members_with_storage_scale(gpu_ctx):
    """Members stored times a power of two in the tiled 2D kernel and the
    batched fused 3D step give the results of the same batch unscaled (FP32:
    the scale is exact; the tiled kernel's own TF/SF and probe records divide
    it out), and an FP16 3D member with a scale equals its single-simulation
    run with the same scale."""
    from luxum_amd.sim import Batch

    def run_batch(ws, steps):
        ms = [make(w, gpu_ctx) for w in ws]
        b = Batch(gpu_ctx)
        for m in ms:
            b.add(m)
        b.step(steps)
        b.sync()
        return b, ms

    w2, w3 = phc(160, 3), configs.box3d(64, 40, 36, steps=120)
    say("batch unscaled, then scaled 1024 (2D tiled, 3D FP32) with a scaled FP16 3D member: 120 steps")
    b0, plain = run_batch([w2, w3], 120)
    for w in (w2, w3):
        w.p.storage_scale = 1024.0
    w16 = configs.box3d(72, 40, 36, precision=LX_FP16, steps=120)
    w16.p.storage_scale = 2.0 ** 12
    b1, scaled = run_batch([w2, w3, w16], 120)
    for w, s, r in zip((w2, w3), scaled, plain):
        e, pe = field_err(s, r), rel_l2(s.probes(), r.probes())
        say(f"{w.name}: scaled vs unscaled batch fields {e:.2e} probes {pe:.2e}")
        assert e <= 1e-12 and pe <= 1e-12
    r16 = make(w16, gpu_ctx)
    r16.step(120)
    r16.sync()
    e, pe = field_err(scaled[2], r16), rel_l2(scaled[2].probes(), r16.probes())
    say(f"{w16.name}: scaled FP16 batch vs scaled FP16 single fields {e:.2e} probes {pe:.2e}")
    assert e <= 1e-6 and pe <= 1e-6
    b0.close()
    b1.close()


def test_batch_member_pinned_snapshots(gpu_ctx):
    """Snapshots of batch members into pinned host memory (the DMA-engine
    path) between batch steps, one 2D tiled and one 3D member: complete after
    the batch's sync alone, and equal to each member's synchronous readback."""
    import ctypes

    from luxum_amd.params import LX_EZ
    from luxum_amd.sim import Batch
    ws = [phc(160, 5), configs.box3d(64, 40, 36, steps=80)]
    comps = [LX_EY, LX_EZ]
    members = [make(w, gpu_ctx) for w in ws]
    hip = ctypes.CDLL("libamdhip64.so")
    ptrs, bufs = [], []
    b = Batch(gpu_ctx)
    try:
        for w, c in zip(ws, comps):
            shape = w.p.field_shape(c)
            m = int(np.prod(shape))
            q = ctypes.c_void_p()
            assert hip.hipHostMalloc(ctypes.byref(q), ctypes.c_size_t(8 * m), 0) == 0
            ptrs.append(q)
            bufs.append(np.ctypeslib.as_array(ctypes.cast(q, ctypes.POINTER(ctypes.c_double)), shape=shape))
            bufs[-1][...] = -1.0
        for s in members:
            b.add(s)
        say("batch of a 2D and a 3D member: step 24, pinned snapshots and readbacks, step 16, batch sync")
        b.step(24)
        refs = []
        for s, c, buf in zip(members, comps, bufs):
            s.snapshot_async(c, buf)
            refs.append(s.field(c))      # the same state, read synchronously
        b.step(16)
        b.sync()
        for w, g, ref in zip(ws, bufs, refs):
            say(f"{w.name}: snapshot vs readback max diff {float(np.abs(g - ref).max()):.2e}")
            assert np.array_equal(g, ref)
            assert float(np.abs(ref).max()) > 0
    finally:
        b.close()
        for s in members:
            s.sync()
        for q in ptrs:
            hip.hipHostFree(q)
