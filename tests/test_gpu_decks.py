"""The reference's two shipped input decks at their full sizes, read by the
library's own parser and run through the C ABI, against the FP64 oracle on
identical inputs.

  parameters.yaml     2D TM, 1002 x 1002 nodes, plain PEC walls, analytic
                      plane wave along the diagonal (k = (1,0,1)/sqrt 2,
                      E along y, Gaussian sigma 2e-4 m^2 = (14 cells)^2,
                      maxwell.cc:71), S = 0.70710678118 (the 2D Courant
                      limit to 9e-12), 2000 steps.
  parameters_3d.yaml  3D, 102^3 nodes, CPML 10 cells (kappa 15, a 0),
                      analytic plane wave along (1,1,1)/sqrt 3 starting at
                      the grid origin (off_pw 0), S = 1/sqrt 3 (the 3D
                      limit), 600 steps, vacuum (the deck's
                      airplane_100_*.h5 inputs are not shipped).

Compared at four snapshot times: the E and H fields, four probes and the
monitor's max|E|^2 and energy.  Bounds (FP32 storage):
  - every snapshot's error, relative to the run's largest snapshot norm,
    <= 1e-5 (measured <= 5e-6);
  - snapshots whose norm is at least half the largest: relative L2 <= 1e-5;
  - every snapshot within 2.5x of what FP32 storage itself costs: the same
    oracle with every stored value rounded to float (FP64 arithmetic)
    deviates from the FP64 oracle by e_f32; the GPU's deviation is at most
    2.5 e_f32 + 1e-6.  Once a pulse has left the grid the field falls by up
    to 2e4 (3D deck) while the rounding noise that FP32 storage left behind
    stays (mostly zero-group-velocity checkerboard modes that never reach
    the CPML), so the relative L2 of late snapshots is large for any FP32
    storage: 7e-3..1.5e-2 for the FP32-storage oracle at steps 300-600 of
    the 3D deck, 1.7e-5 at step 2000 of the 2D deck, where the pulse is
    leaving the TF box;
  - probe series relative L2 <= 1e-5; monitor max|E|^2 and energy within
    1e-5 of the run's largest value.
"""
import sys

import numpy as np
import pytest

from conftest import ROOT, rel_l2
from luxum_amd.params import LX_EX, LX_EY, LX_EZ
from oracle.lxoracle import OracleSim

pytestmark = pytest.mark.gpu
TOL32 = 1e-5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def vec_err(a, b, comps):
    num = den = 0.0
    for c in comps:
        x, y = a.field(c), b.field(c)
        assert x.shape == y.shape and np.all(np.isfinite(x))
        num += float(np.sum((x - y) ** 2))
        den += float(np.sum(y ** 2))
    return (num / den) ** 0.5 if den > 0 else num ** 0.5


DECKS = {
    "parameters.yaml": (2000, 4, [[LX_EY, 250, 0, 250], [LX_EY, 501, 0, 501], [LX_EY, 750, 0, 300],
                                  [LX_EY, 100, 0, 900]]),
    "parameters_3d.yaml": (600, 4, [[LX_EX, 50, 50, 50], [LX_EZ, 30, 60, 45], [LX_EX, 70, 35, 60],
                                    [LX_EY, 45, 45, 75]]),
}


def snapshot(sim, comps):
    return {k: [sim.field(c) for c in comps if (c < 3) == (k == "E")] for k in ("E", "H")}


def norm(fs):
    return sum(float(np.sum(f ** 2)) for f in fs) ** 0.5


def diff(a, b):
    return sum(float(np.sum((x - y) ** 2)) for x, y in zip(a, b)) ** 0.5


@pytest.mark.parametrize("deck", list(DECKS))
def test_shipped_deck_vs_oracle(gpu_ctx, deck):
    from luxum_amd.sim import Simulation, load_yaml
    steps, nsnap, probes = DECKS[deck]
    p = load_yaml(f"{ROOT}/tests/golden/{deck}")
    assert p.num_timesteps == steps
    say(f"deck {deck}: {p.nx}x{p.ny}x{p.nz}, S={p.courant!r}, pml={p.use_pml}, {steps} steps")
    g = Simulation(p, gpu_ctx)
    o = OracleSim(p)
    f = OracleSim(p, storage="fp32")
    for s in (g, o, f):
        s.set_probes(probes)
    comps = p.components()
    rows = []
    for q in range(nsnap):
        n = steps // nsnap
        for s in (g, o, f):
            s.step(n)
        g.sync()
        sg, so, sf = snapshot(g, comps), snapshot(o, comps), snapshot(f, comps)
        row = {"t": (q + 1) * n, "mon": g.monitor(), "omon": (o.max_e2(), o.energy())}
        for k in ("E", "H"):
            row[k] = (norm(so[k]), diff(sg[k], so[k]), diff(sf[k], so[k]))
        rows.append(row)
    peak = {k: max(r[k][0] for r in rows) for k in ("E", "H")}
    pmax = max(r["omon"][0] for r in rows)
    pen = max(r["omon"][1] for r in rows)
    for r in rows:
        for k in ("E", "H"):
            nrm, dg, df = r[k]
            eg, ef = dg / nrm, df / nrm
            say(f"{deck} t={r['t']} {k}: |F|/peak {nrm / peak[k]:.2e}  GPU rel L2 {eg:.3e} (vs peak {dg / peak[k]:.2e})"
                f"  FP32-storage oracle {ef:.3e}")
            assert dg / peak[k] <= TOL32
            if nrm >= 0.5 * peak[k]:
                assert eg <= TOL32
            assert eg <= 2.5 * ef + 1e-6
        (mx, en), (omx, oen) = r["mon"], r["omon"]
        say(f"{deck} t={r['t']}: max|E|^2 {omx:.4e} diff/peak {abs(mx - omx) / pmax:.2e}, "
            f"energy diff/peak {abs(en - oen) / pen:.2e}")
        assert abs(mx - omx) <= TOL32 * pmax and abs(en - oen) <= TOL32 * pen
    pe = rel_l2(g.probes(), o.probes())
    say(f"{deck}: probes rel L2 {pe:.3e} (FP32-storage oracle {rel_l2(f.probes(), o.probes()):.3e})")
    assert pe <= TOL32
    g.close()
    o.close()
    f.close()
