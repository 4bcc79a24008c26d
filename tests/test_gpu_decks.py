"""The reference's two shipped input decks at their full sizes, read by the
library's own parser and run through the C ABI, against the FP64 oracle on
identical inputs.

  parameters.yaml     2D TM, 1002 x 1002 nodes, plain PEC walls, analytic
                      plane wave along the diagonal (k = (1,0,1)/sqrt 2,
                      E along y, Gaussian sigma 2e-4 m^2 = (14 cells)^2,
                      maxwell.cc:71), S = 0.70710678118 (the 2D Courant
                      limit to 6.5e-12), 2000 steps.
  parameters_3d.yaml  3D, 102^3 nodes, CPML 10 cells, analytic plane wave
                      along (1,1,1)/sqrt 3, S = 1/sqrt 3 (the 3D limit),
                      600 steps, vacuum (the deck's airplane_100_*.h5 inputs
                      are not shipped).

Compared: the E and H fields at four snapshot times, four probes and the
monitor's max|E|^2 and energy.  Tolerance: relative L2 <= 1e-5 (FP32) on
every snapshot and probe series, monitor values to 1e-5.
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


@pytest.mark.parametrize("deck", list(DECKS))
def test_shipped_deck_vs_oracle(gpu_ctx, deck):
    from luxum_amd.sim import Simulation, load_yaml
    steps, nsnap, probes = DECKS[deck]
    p = load_yaml(f"{ROOT}/tests/golden/{deck}")
    assert p.num_timesteps == steps
    say(f"deck {deck}: {p.nx}x{p.ny}x{p.nz}, S={p.courant!r}, pml={p.use_pml}, {steps} steps")
    g = Simulation(p, gpu_ctx)
    o = OracleSim(p)
    g.set_probes(probes)
    o.set_probes(probes)
    comps = p.components()
    worst = 0.0
    for q in range(nsnap):
        n = steps // nsnap
        g.step(n)
        o.step(n)
        g.sync()
        t = (q + 1) * n
        eE = vec_err(g, o, [c for c in comps if c < 3])
        eH = vec_err(g, o, [c for c in comps if c >= 3])
        mx, en = g.monitor()
        omx, oen = o.max_e2(), o.energy()
        dm, de = abs(mx - omx) / omx, abs(en - oen) / oen
        say(f"{deck} t={t}: rel L2 E {eE:.3e} H {eH:.3e}; max|E|^2 {omx:.4e} (rel diff {dm:.2e}), "
            f"energy rel diff {de:.2e}")
        worst = max(worst, eE, eH, dm, de)
    pe = rel_l2(g.probes(), o.probes())
    say(f"{deck}: probes rel L2 {pe:.3e}; worst snapshot {worst:.3e}")
    assert pe <= TOL32
    assert worst <= TOL32
    g.close()
    o.close()
