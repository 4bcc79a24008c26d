"""Serving (continuous batching of simulation requests) on the GPU.

Every request's result (final fields and probe series) is the same served in
a mix as served alone: relative L2 <= 1e-5 for FP32 storage and <= 5e-3 for
FP16 storage.  The kernels are the same; only the tick and chunk boundaries
differ, and with FP16 storage a tiled batch rounds the fields to FP16 at
every chunk end, so the rounding points (not the arithmetic) move with the
mix.  A served FP32 request matches the FP64 oracle to <= 1e-5.
"""
import sys

import numpy as np
import pytest

from conftest import rel_l2
from luxum_amd import configs
from luxum_amd.params import LX_FP16, LX_FP32
from oracle.lxoracle import OracleSim

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def fields_err(sa, ia, sb, ib, p):
    num = den = 0.0
    for c in p.components():
        a, b = sa.field(ia, c), sb.field(ib, c)
        assert np.all(np.isfinite(a))
        num += float(np.sum((a - b) ** 2))
        den += float(np.sum(b ** 2))
    return (num / den) ** 0.5 if den > 0 else num ** 0.5


def mix():
    rng = np.random.default_rng(5)
    reqs = [
        (configs.scatterer2d(192, rng, LX_FP32), 300, 0.0),
        (configs.scatterer2d(256, rng, LX_FP16), 260, 0.0),
        (configs.scatterer2d(320, rng, LX_FP32), 200, 0.02),
        (configs.scatterer3d(64, rng, LX_FP32), 150, 0.0),
        (configs.scatterer2d(160, rng, LX_FP32), 180, 0.2),
        (configs.scatterer3d(72, rng, LX_FP16), 120, 0.05),
    ]
    for w, steps, _ in reqs:
        w.p.pw_delay = 1.5
        w.p.num_timesteps = steps
    return reqs


def serve(ctx, reqs, plain=False, cap=200 * 10 ** 9):
    from luxum_amd.server import Server
    srv = Server(ctx, hbm_cap=cap, plain=plain)
    ids = [srv.submit(w.p, w.ids, w.eps_r, steps, arrival, w.probes, keep_fields=True) for w, steps, arrival in reqs]
    srv.run()
    return srv, ids


def test_served_results_same_alone_and_in_mix(gpu_ctx):
    from luxum_amd.server import LX_REQ_DONE
    reqs = mix()
    say(f"serving a mix of {len(reqs)} requests")
    srv, ids = serve(gpu_ctx, reqs)
    st = srv.stats()
    say(f"mix: completed {st.completed} ticks {st.ticks} peak active {st.peak_active} makespan {st.makespan:.3f}s")
    assert st.completed == len(reqs) and st.failed == 0
    for (w, steps, arrival), rid in zip(reqs, ids):
        r = srv.result(rid)
        assert r.state == LX_REQ_DONE and r.steps_done == steps
        assert r.finished >= r.admitted >= arrival - 1e-6
        alone, aid = serve(gpu_ctx, [(w, steps, 0.0)])
        e = fields_err(srv, rid, alone, aid[0], w.p)
        pe = rel_l2(srv.probes(rid), alone.probes(aid[0]))
        tol = 5e-3 if w.p.precision == LX_FP16 else 1e-5
        say(f"{w.name} {steps} steps: mix vs alone fields {e:.2e} probes {pe:.2e} (tol {tol:g})")
        assert e <= tol and pe <= tol
        alone.close()
    # a served FP32 request against the FP64 oracle
    w, steps, _ = reqs[0]
    o = OracleSim(w.p)
    o.set_materials(w.ids, w.eps_r)
    o.set_probes(w.probes)
    o.step(steps)
    num = den = 0.0
    for c in w.p.components():
        a, b = srv.field(ids[0], c), o.field(c)
        num += float(np.sum((a - b) ** 2))
        den += float(np.sum(b ** 2))
    e = (num / den) ** 0.5
    pe = rel_l2(srv.probes(ids[0]), o.probes())
    say(f"{w.name}: served vs oracle fields {e:.2e} probes {pe:.2e}")
    assert e <= 1e-5 and pe <= 1e-5
    srv.close()


def test_plain_mode_matches_batched(gpu_ctx):
    reqs = mix()[:4]
    a, ia = serve(gpu_ctx, reqs)
    b, ib = serve(gpu_ctx, reqs, plain=True)
    assert b.stats().peak_active == 1
    for (w, steps, _), x, y in zip(reqs, ia, ib):
        e = fields_err(a, x, b, y, w.p)
        tol = 2e-2 if w.p.precision == LX_FP16 else 1e-5   # FP16: tiled vs half-step rounding points
        say(f"{w.name}: batched vs plain fields {e:.2e} (tol {tol:g})")
        assert e <= tol
    a.close()
    b.close()


def test_admission_cap_and_overtaking(gpu_ctx):
    """Admission keeps the reserved footprint under the cap; a short request
    arriving after a long one joins at a tick boundary and finishes first."""
    from luxum_amd.server import footprint
    rng = np.random.default_rng(9)
    long_w = configs.scatterer2d(1024, rng, LX_FP32)
    short_w = configs.scatterer2d(192, rng, LX_FP32)
    big = [configs.scatterer2d(384, rng, LX_FP32) for _ in range(4)]
    fl = footprint(long_w.p, True, 3000, len(long_w.probes))
    fs = footprint(short_w.p, True, 100, len(short_w.probes))
    fb = max(footprint(w.p, True, 400, len(w.probes)) for w in big)
    cap = fl + fs + int(1.5 * fb)   # the long one, the short one and one big one at a time
    reqs = [(long_w, 3000, 0.0)] + [(w, 400, 0.0) for w in big] + [(short_w, 100, 0.001)]
    srv, ids = serve(gpu_ctx, reqs, cap=cap)
    st = srv.stats()
    say(f"cap {cap / 1e6:.1f} MB: peak {st.peak_bytes / 1e6:.1f} MB active {st.peak_active} ticks {st.ticks}")
    assert st.completed == len(reqs) and st.failed == 0
    assert st.peak_bytes <= cap
    r_long, r_short = srv.result(ids[0]), srv.result(ids[-1])
    say(f"long finished {r_long.finished:.3f}s, short (arrived {r_short.arrival}s) finished {r_short.finished:.3f}s")
    assert r_short.finished < r_long.finished
    srv.close()
