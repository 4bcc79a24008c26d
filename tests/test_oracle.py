"""Pins of the FP64 oracle (CPU).

The reference ships no tests and cannot be built here, so every function of
the oracle is pinned by what the mathematics fixes: closed forms, exact
invariants of the Yee scheme and physical bounds, chosen so that a plausible
mistake (dropped term, wrong sign, off-by-one index, transposed operand)
fails one of them.
"""
import math

import numpy as np
import pytest

from luxum_amd.params import (LX_EX, LX_EY, LX_EZ, LX_HX, LX_HY, LX_HZ, LX_INC_ANALYTIC, LX_INC_AUX1D,
                              LX_SRC_HARD_CW, LX_SRC_SOFT, LX_WAVE_GAUSSIAN, default_params)
from oracle import lxoracle as O

DX = 5e-8


def consts():
    eps0 = 8.85 * 10 ** -12
    mu0 = math.pi * 4 * 10 ** -7
    return eps0, mu0, 1 / math.sqrt(eps0 * mu0)


# ---------------------------------------------------------------------------
# constants and CPML profiles (maxwell.cc:60-86, 997-1280)
# ---------------------------------------------------------------------------

def test_constants_match_reference_definitions():
    p = default_params(courant=0.5, use_pml=1, pml_thickness=8 * DX, sigma_m=3.0, sigma_ratio=0.65)
    c = O.ref_constants(p)
    eps0, mu0, c0 = consts()
    assert c["eps0"] == pytest.approx(8.85e-12, rel=1e-15)      # not 8.8541878e-12
    assert c["c"] == pytest.approx(c0, rel=1e-15)
    assert c["dt"] == pytest.approx(0.5 * DX / c0, rel=1e-15)    # dt uses dx only
    eta0 = math.sqrt(mu0 / eps0)
    assert c["sigma_opt"] == pytest.approx(0.8 * 4 / (DX * eta0), rel=1e-14)
    assert c["sigma_max"] == pytest.approx(0.65 * c["sigma_opt"], rel=1e-14)


def _pml_params(n=40, cells=8, m=3.0, kappa=5.0, a_max=0.05, a_m=1.0):
    return default_params(nx=n, ny=n, nz=n, use_pml=1, pml_thickness=cells * DX, sigma_m=m, kappa_max=kappa,
                          a_max=a_max, a_m=a_m)


def test_sigma_profile_closed_form():
    p = _pml_params()
    N, d = 40, 8.0
    smax = O.ref_constants(p)["sigma_max"]
    sE = O.profile(p, 0, True, O.SIGMA)
    sH = O.profile(p, 0, False, O.SIGMA)
    assert sE.shape == (N,) and sH.shape == (N - 1,)
    # wall node carries sigma_max; grading ((d-1-x)/(d-1))^3 with x measured in cells
    assert sE[0] == pytest.approx(smax, rel=1e-14)
    assert sE[3] == pytest.approx(smax * (4 / 7) ** 3, rel=1e-13)
    assert sH[0] == pytest.approx(smax * (6.5 / 7) ** 3, rel=1e-13)
    # interior and the inner edge x = d-1 are lossless
    assert np.all(sE[7:33] == 0.0) and sE[6] > 0
    # mirror symmetry about the grid centre: E nodes i <-> N-1-i, H nodes i <-> N-2-i
    np.testing.assert_allclose(sE, sE[::-1], rtol=1e-12, atol=0)
    np.testing.assert_allclose(sH, sH[::-1], rtol=1e-12, atol=0)
    # the exponent: ratio of two graded values
    assert sE[1] / sE[2] == pytest.approx((6 / 5) ** 3, rel=1e-13)


def test_kappa_alpha_b_c_profiles():
    p = _pml_params()
    eps0, mu0, c0 = consts()
    dt = 0.5 * DX / c0
    k = O.profile(p, 0, True, O.KAPPA)
    a = O.profile(p, 0, True, O.ALPHA)
    b = O.profile(p, 0, True, O.B)
    c = O.profile(p, 0, True, O.CC)
    s = O.profile(p, 0, True, O.SIGMA)
    assert k[0] == pytest.approx(5.0) and np.all(k[7:33] == 1.0)
    # alpha grading is reversed: zero at the wall, a_max at the inner edge (maxwell.cc:1164)
    assert a[0] == 0.0 and a[7] == pytest.approx(0.05) and a[39] == 0.0 and a[32] == pytest.approx(0.05)
    # c == 0 exactly where sigma == 0 (maxwell.cc:1272)
    assert np.array_equal(c == 0.0, s == 0.0)
    # b = exp(-(sigma/kappa - a) dt/eps0): at the inner edge sigma = 0 and a = a_max, so the
    # reference's minus sign on alpha gives b > 1 (textbook CPML would give b < 1)
    assert b[7] == pytest.approx(math.exp(0.05 * dt / eps0), rel=1e-14) and b[7] > 1.0
    assert b[0] == pytest.approx(math.exp(-(s[0] / 5.0) * dt / eps0), rel=1e-14)
    # c closed form at the wall (a = 0 there): c = sigma/(sigma*kappa) * (b - 1)
    assert c[0] == pytest.approx((b[0] - 1) / 5.0, rel=1e-13)
    # interior: b = 1 exactly, no psi coupling
    assert np.all(b[8:32] == 1.0)


def test_profile_uses_each_axis_spacing():
    p = default_params(nx=40, ny=30, nz=20, dx=DX, dy=2 * DX, dz=DX, use_pml=1, pml_thickness=8 * DX)
    sy = O.profile(p, 1, True, O.SIGMA)
    # along y the slab is 8*dx / (2*dx) = 4 cells thick: sigma > 0 on nodes 0..2 only
    assert np.all(sy[:3] > 0) and np.all(sy[3:27] == 0) and np.all(sy[27:] > 0)


# ---------------------------------------------------------------------------
# curls and updates
# ---------------------------------------------------------------------------

def test_hand_computed_single_step_3d():
    """One plain 3D step from a known E: H = -Db*curl E with forward
    differences on the H extents, then E += Cb*curl H with backward
    differences on the interior only (PEC rows untouched)."""
    p = default_params(nx=4, ny=4, nz=4)
    s = O.OracleSim(p, threads=1)
    eps0, mu0, c0 = consts()
    dt = 0.5 * DX / c0
    Db, Cb = dt / mu0, dt / eps0
    rng = np.random.default_rng(1)
    E0 = [rng.standard_normal(p.field_shape(c)) for c in (LX_EX, LX_EY, LX_EZ)]
    for c in range(3):
        s.set_field(c, E0[c])
    s.step(1)
    Ex, Ey, Ez = E0
    # Hx(i,j,k) on (4,3,3): (Ez(j+1)-Ez(j))/dy - (Ey(k+1)-Ey(k))/dz
    cx = (Ez[:, 1:, :] - Ez[:, :-1, :]) / DX - (Ey[:, :, 1:] - Ey[:, :, :-1]) / DX
    cy = (Ex[:, :, 1:] - Ex[:, :, :-1]) / DX - (Ez[1:, :, :] - Ez[:-1, :, :]) / DX
    cz = (Ey[1:, :, :] - Ey[:-1, :, :]) / DX - (Ex[:, 1:, :] - Ex[:, :-1, :]) / DX
    H = [-Db * cx, -Db * cy, -Db * cz]
    for c in range(3):
        np.testing.assert_allclose(s.field(3 + c), H[c], rtol=1e-13, atol=1e-30)
    Hx, Hy, Hz = H
    ex = Ex.copy()
    ex[:, 1:-1, 1:-1] += Cb * ((Hz[:, 1:, 1:-1] - Hz[:, :-1, 1:-1]) / DX - (Hy[:, 1:-1, 1:] - Hy[:, 1:-1, :-1]) / DX)
    ez = Ez.copy()
    ez[1:-1, 1:-1, :] += Cb * ((Hy[1:, 1:-1, :] - Hy[:-1, 1:-1, :]) / DX - (Hx[1:-1, 1:, :] - Hx[1:-1, :-1, :]) / DX)
    np.testing.assert_allclose(s.field(LX_EX), ex, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(s.field(LX_EZ), ez, rtol=1e-12, atol=1e-12)


def test_per_component_cb_single_step():
    """Per-component materials (maxwell_three_dim_pml.cc:109-128: Cb->Vx,
    Cb->Vy, Cb->Vz from three files): after one plain step from a known H
    field, E_c changed by Cb_c(i,j,k)*curl_c H, where Cb_c takes column c of
    the permittivity triple of the node's id, and the isotropic entry point
    equals the per-component one with three equal columns bit for bit."""
    p = default_params(nx=5, ny=6, nz=7)
    eps0, mu0, c0 = consts()
    dt = 0.5 * DX / c0
    rng = np.random.default_rng(3)
    ids = rng.integers(0, 3, size=(5, 6, 7)).astype(np.uint8)
    eps3 = np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0], [7.5, 1.5, 2.5]])
    H0 = [rng.standard_normal(p.field_shape(c)) for c in (LX_HX, LX_HY, LX_HZ)]
    s = O.OracleSim(p, threads=1)
    s.set_materials(ids, eps3)
    for c in range(3):
        s.set_field(3 + c, H0[c])
    s.step(1)
    Hx, Hy, Hz = [s.field(3 + c) for c in range(3)]     # H after the H half step (E was zero: unchanged)
    for c in range(3):
        np.testing.assert_array_equal(s.field(3 + c), H0[c])
    ex = np.zeros(p.field_shape(LX_EX))
    ey = np.zeros(p.field_shape(LX_EY))
    ez = np.zeros(p.field_shape(LX_EZ))
    ex[:, 1:-1, 1:-1] = (Hz[:, 1:, 1:-1] - Hz[:, :-1, 1:-1]) / DX - (Hy[:, 1:-1, 1:] - Hy[:, 1:-1, :-1]) / DX
    ey[1:-1, :, 1:-1] = (Hx[1:-1, :, 1:] - Hx[1:-1, :, :-1]) / DX - (Hz[1:, :, 1:-1] - Hz[:-1, :, 1:-1]) / DX
    ez[1:-1, 1:-1, :] = (Hy[1:, 1:-1, :] - Hy[:-1, 1:-1, :]) / DX - (Hx[1:-1, 1:, :] - Hx[1:-1, :-1, :]) / DX
    for c, curl in enumerate((ex, ey, ez)):
        sh = curl.shape
        cb = dt / (eps0 * eps3[ids[: sh[0], : sh[1], : sh[2]], c])
        np.testing.assert_allclose(s.field(c), cb * curl, rtol=1e-12, atol=1e-20)
    a = O.OracleSim(p, threads=1)
    b = O.OracleSim(p, threads=1)
    a.set_materials(ids, eps3[:, 1])
    b.set_materials(ids, np.repeat(eps3[:, 1:2], 3, axis=1))
    for o in (a, b):
        for c in range(3):
            o.set_field(3 + c, H0[c])
        o.step(3)
    for c in range(6):
        np.testing.assert_array_equal(a.field(c), b.field(c))


@pytest.mark.parametrize("pml", [0, 1])
def test_cavity_eigenmode_follows_yee_dispersion(pml):
    """TM110 mode of the PEC box (PML flag on but zero thickness effect is not
    allowed, so pml=1 uses a lossless profile with sigma_ratio=0, kappa=1).
    The discrete mode is an exact eigenvector: Ez^{n+1} + Ez^{n-1} = 2cos(w dt) Ez^n
    with sin^2(w dt/2) = (c dt)^2 sum_i sin^2(k_i d_i/2)/d_i^2."""
    nx, ny, nz = 21, 17, 6
    p = default_params(nx=nx, ny=ny, nz=nz, courant=0.5)
    if pml:
        p.use_pml = 1
        p.pml_thickness = 3 * DX
        p.sigma_ratio = 0.0
        p.kappa_max = 1.0
        p.a_max = 0.0
    s = O.OracleSim(p, threads=1)
    i = np.arange(nx)[:, None, None]
    j = np.arange(ny)[None, :, None]
    Ez = np.sin(math.pi * i / (nx - 1)) * np.sin(math.pi * j / (ny - 1)) * np.ones((1, 1, nz - 1))
    s.set_field(LX_EZ, Ez)
    s.set_probes([[LX_EZ, 7, 5, 2]])
    s.step(120)
    v = s.probes()[:, 0]
    eps0, mu0, c0 = consts()
    dt = 0.5 * DX / c0
    kx, ky = math.pi / ((nx - 1) * DX), math.pi / ((ny - 1) * DX)
    s2 = (c0 * dt) ** 2 * (math.sin(kx * DX / 2) ** 2 + math.sin(ky * DX / 2) ** 2) / DX ** 2
    cosw = 1 - 2 * s2
    ratio = (v[2:] + v[:-2]) / v[1:-1]
    mask = np.abs(v[1:-1]) > 1e-3 * np.abs(v).max()
    np.testing.assert_allclose(ratio[mask], 2 * cosw, rtol=0, atol=1e-10)


def _pec_random_e(p, seed):
    """Random E with tangential components zero on the PEC walls."""
    rng = np.random.default_rng(seed)
    out = []
    for c in (LX_EX, LX_EY, LX_EZ):
        f = rng.standard_normal(p.field_shape(c))
        st = [(1, 0, 0), (0, 1, 0), (0, 0, 1)][c]
        for a in range(3):
            if not st[a]:
                sl = [slice(None)] * 3
                sl[a] = 0
                f[tuple(sl)] = 0
                sl[a] = -1
                f[tuple(sl)] = 0
        out.append(f)
    return out


def test_discrete_energy_is_conserved_in_pec_box():
    """eps|E^n|^2 + mu H^{n-1/2}.H^{n+1/2} is invariant for the lossless Yee
    scheme; any asymmetric curl (wrong sign or index) breaks it."""
    p = default_params(nx=12, ny=11, nz=10, courant=0.5)
    s = O.OracleSim(p, threads=1)
    for c, f in enumerate(_pec_random_e(p, 3)):
        s.set_field(c, f)
    E_hist, H_hist = [], []
    for _ in range(30):
        s.step(1)
        E_hist.append([s.field(c) for c in (LX_EX, LX_EY, LX_EZ)])
        H_hist.append([s.field(c) for c in (LX_HX, LX_HY, LX_HZ)])
    eps0, mu0, _ = consts()
    energies = []
    for n in range(len(E_hist) - 1):
        # after step n: E^{n+1} and H^{n+1/2}; after step n+1: H^{n+3/2}
        we = sum(float(np.sum(e ** 2)) for e in E_hist[n]) * eps0
        wh = sum(float(np.sum(a * b)) for a, b in zip(H_hist[n], H_hist[n + 1])) * mu0
        energies.append(we + wh)
    energies = np.array(energies)
    assert np.max(np.abs(energies / energies[0] - 1)) < 1e-12


def test_divergence_of_e_is_invariant():
    p = default_params(nx=10, ny=9, nz=8, courant=0.5)
    s = O.OracleSim(p, threads=1)
    rng = np.random.default_rng(5)
    for c in (LX_EX, LX_EY, LX_EZ):
        s.set_field(c, rng.standard_normal(p.field_shape(c)))

    def div():
        Ex, Ey, Ez = (s.field(c) for c in (LX_EX, LX_EY, LX_EZ))
        # at interior nodes (i,j,k), backward differences of the staggered components
        return ((Ex[1:, 1:-1, 1:-1] - Ex[:-1, 1:-1, 1:-1]) + (Ey[1:-1, 1:, 1:-1] - Ey[1:-1, :-1, 1:-1])
                + (Ez[1:-1, 1:-1, 1:] - Ez[1:-1, 1:-1, :-1])) / DX

    d0 = div()
    s.step(25)
    np.testing.assert_allclose(div(), d0, rtol=0, atol=1e-9 * np.abs(d0).max())


# ---------------------------------------------------------------------------
# CPML
# ---------------------------------------------------------------------------

def _pulse_run(n, pml_cells, steps, probe):
    p = default_params(is_planar=1, nx=n, ny=1, nz=n, dy=0.0, courant=0.5, use_ps=1, src_kind=LX_SRC_SOFT,
                       src_comp=LX_EY, src_amp=1.0, src_wave=LX_WAVE_GAUSSIAN, src_lambda0=20 * DX,
                       src_fwhm=1.0, src_delay=4.0, src_pos=(n // 2, 0, n // 2))
    if pml_cells:
        p.use_pml = 1
        p.pml_thickness = pml_cells * DX
    s = O.OracleSim(p, threads=4)
    s.set_probes([[LX_EY, n // 2 + probe[0], 0, n // 2 + probe[1]]])
    s.step(steps)
    return s.probes()[:, 0]


def test_cpml_reflection_is_small():
    """A pulse from a soft source hits the CPML 40 cells away; compare a probe
    next to the CPML with a run on a domain large enough that no reflection
    returns within the window."""
    steps = 400
    small = _pulse_run(120, 16, steps, (30, 0))
    big = _pulse_run(600, 0, steps, (30, 0))
    refl = np.abs(small - big).max() / np.abs(big).max()
    assert refl < 5e-3


# ---------------------------------------------------------------------------
# TF/SF and the aux line
# ---------------------------------------------------------------------------

def test_tfsf_aux_line_vacuum_is_exact():
    p = default_params(nx=40, ny=30, nz=30, use_pml=1, pml_thickness=6 * DX, use_pw=1, incident=LX_INC_AUX1D,
                       tf_lo=(12, 10, 10), tf_hi=(28, 20, 20), num_timesteps=200, pw_delay=3.0)
    s = O.OracleSim(p, threads=4)
    s.step(150)
    Ez = s.field(LX_EZ)
    e, h, x0 = s.aux_line()
    # TF region carries the line's E exactly, SF region stays at rounding level
    tf = Ez[12:29, 12, 13]
    np.testing.assert_allclose(tf, e[12 - x0:29 - x0], rtol=0, atol=1e-13)
    m = np.ones_like(Ez, bool)
    m[11:30, 9:22, 9:22] = False
    assert np.abs(Ez[m]).max() < 1e-13 * np.abs(Ez).max() + 1e-300
    Hy = s.field(LX_HY)
    # H_dir = normalize(k x E) = -y for k = +x, E = +z
    assert np.abs(Hy[14:26, 12, 13]).max() > 0
    assert np.sign(Hy[20, 15, 15] * Ez[20, 15, 15]) <= 0


def test_aux_line_propagates_at_grid_speed():
    """The line's pulse peak moves S cells per step (to within the numerical
    dispersion of a 20-cell carrier) and keeps its amplitude."""
    p = default_params(nx=40, ny=30, nz=30, use_pw=1, incident=LX_INC_AUX1D, tf_lo=(10, 10, 10),
                       tf_hi=(28, 20, 20), num_timesteps=900, pw_wave=LX_WAVE_GAUSSIAN, pw_fwhm=3.0,
                       pw_delay=3.0)
    s = O.OracleSim(p, threads=2)
    peaks = []
    for n in (400, 600):
        s.step(n - (s.probes().shape[0] if s.n_probes else 0) if not peaks else 200)
        e, h, x0 = s.aux_line()
        peaks.append(int(np.argmax(np.abs(e))))
    speed = (peaks[1] - peaks[0]) / 200.0
    assert abs(speed - 0.5) < 0.02
    assert np.abs(e).max() == pytest.approx(1.0, rel=0.05)


def _analytic_g(p, x, y, z, tp):
    """The reference's incident scalar exp(-((k.r) + off*dx - t'*c*dt)^2/sigma)
    at node coordinates (x, y, z) in cells and time level t' (maxwell.cc:206,
    289, 604, 712)."""
    _, _, c0 = consts()
    dt = p.courant * p.dx / c0
    pos = x * p.dx * p.k_pw[0] + y * p.dy * p.k_pw[1] + z * p.dz * p.k_pw[2]
    return np.exp(-(pos + p.off_pw * p.dx - tp * c0 * dt) ** 2 / p.pw_sigma)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_tfsf_analytic_time_level_3d():
    """Analytic injection along +x (E along z, H_dir = -y) with a resolved
    pulse in vacuum.  The reference samples E_inc at t-1/2 for the curl-E
    correction of step t (maxwell.cc:604) and H_inc at t for the curl-H
    correction (:712), a consistent half-step delay: after n steps the TF
    region holds Ez^n = g(x, n-1/2) and Hy^{n-1/2} = -g(x+1/2, n-1)/(c mu0).
    Measured 3.6e-5 and 3.4e-5 (grid dispersion of a 20-cell pulse), while
    a level off by half a step is 1.2e-2 away.  An E_inc level of t (no
    delay) leaks 4e-3 of the pulse into the SF region; a correct injection
    leaks 1.3e-5."""
    n = 200
    p = default_params(nx=80, ny=24, nz=24, use_pml=1, pml_thickness=6 * DX, use_pw=1, incident=LX_INC_ANALYTIC,
                       off_pw=60.0, pw_sigma=(20 * DX) ** 2, tf_lo=(10, 8, 8), tf_hi=(69, 15, 15))
    p.k_pw[:] = (1.0, 0.0, 0.0)
    p.e_pw[:] = (0.0, 0.0, 1.0)
    s = O.OracleSim(p, threads=4)
    s.step(n)
    _, mu0, c0 = consts()
    x = np.arange(12, 68)
    ez = s.field(LX_EZ)[12:68, 11, 11]
    hy = s.field(LX_HY)[12:68, 11, 11]
    assert _rel(ez, _analytic_g(p, x, 0, 0, n - 0.5)) < 2e-4
    assert _rel(hy, -_analytic_g(p, x + 0.5, 0, 0, n - 1.0) / (c0 * mu0)) < 2e-4
    for off in (-0.5, +0.5):
        assert _rel(ez, _analytic_g(p, x, 0, 0, n - 0.5 + off)) > 5e-3
        assert _rel(hy, -_analytic_g(p, x + 0.5, 0, 0, n - 1.0 + off) / (c0 * mu0)) > 5e-3
    E = [s.field(c) for c in (LX_EX, LX_EY, LX_EZ)]
    tfmax = max(np.abs(f[12:68, 9:15, 9:15]).max() for f in E)
    sfmax = max(np.abs(np.concatenate([f[:8].ravel(), f[72:].ravel()])).max() for f in E)
    assert tfmax > 0.9
    assert sfmax < 2e-4 * tfmax


def test_tfsf_analytic_oblique_leakage_bound():
    """Reference analytic injection with an oblique k: the SF region sees only
    the grid-dispersion mismatch of a resolved pulse, measured 1.3e-5 of the
    TF amplitude.  A wrong face set or sign leaks O(1), E_inc taken at t
    instead of t-1/2 leaks 4.1e-3 and H_inc on the face instead of half a
    cell outside it 2.9e-2."""
    p = default_params(nx=44, ny=40, nz=40, use_pml=1, pml_thickness=6 * DX, use_pw=1, incident=LX_INC_ANALYTIC,
                       off_pw=100.0, pw_sigma=(20 * DX) ** 2, tf_lo=(12, 11, 11), tf_hi=(31, 28, 28))
    p.k_pw[:] = (0.5773502691896258,) * 3
    p.e_pw[:] = (0.70710678118, 0.0, -0.70710678118)
    s = O.OracleSim(p, threads=4)
    s.step(260)
    E = [s.field(c) for c in (LX_EX, LX_EY, LX_EZ)]
    tfmax = max(np.abs(f[14:29, 13:26, 13:26]).max() for f in E)
    sfmax = max(np.abs(np.concatenate([f[:9].ravel(), f[35:].ravel()])).max() for f in E)
    assert tfmax > 0.3
    assert sfmax < 2e-4 * tfmax


def test_tfsf_analytic_2d_oblique():
    """2D analytic injection with an oblique k in the x-z plane: the TF
    region holds the negated incident wave (maxwell.cc:206 vs :289),
    Ey^n = -g(x, z, n-1/2), to 1.4e-4 (measured; half a step off is 1.1e-2
    away), and the SF region outside the CPML sees 6e-5 of it."""
    n = 260
    p = default_params(is_planar=1, nx=90, ny=1, nz=90, dy=0.0, use_pml=1, pml_thickness=8 * DX, use_pw=1,
                       incident=LX_INC_ANALYTIC, off_pw=80.0, pw_sigma=(20 * DX) ** 2, tf_lo=(14, 0, 14),
                       tf_hi=(75, 0, 75), src_comp=LX_EY)
    p.k_pw[:] = (0.6, 0.0, 0.8)
    p.e_pw[:] = (0.0, 1.0, 0.0)
    s = O.OracleSim(p, threads=4)
    s.step(n)
    ey = s.field(LX_EY)[:, 0, :]
    i = np.arange(18, 72)[:, None]
    k = np.arange(18, 72)[None, :]
    tf = ey[18:72, 18:72]
    assert _rel(tf, -_analytic_g(p, i, 0, k, n - 0.5)) < 1e-3
    for off in (-0.5, +0.5):
        assert _rel(tf, -_analytic_g(p, i, 0, k, n - 0.5 + off)) > 5e-3
    m = np.zeros_like(ey, bool)
    m[8:82, 8:82] = True
    m[12:78, 12:78] = False
    assert np.abs(ey[m]).max() < 5e-4 * np.abs(tf).max()


def test_tfsf_2d_negated_wave():
    """2D TF/SF injects the negated incident wave (maxwell.cc:523 vs :604)."""
    p = default_params(is_planar=1, nx=60, ny=1, nz=50, dy=0.0, use_pw=1, incident=LX_INC_AUX1D,
                       tf_lo=(10, 0, 10), tf_hi=(45, 0, 40), num_timesteps=300, pw_delay=3.0, src_comp=LX_EY)
    p.e_pw[:] = (0.0, 1.0, 0.0)
    s = O.OracleSim(p, threads=2)
    s.step(180)
    Ey = s.field(LX_EY)
    e, h, x0 = s.aux_line()
    np.testing.assert_allclose(Ey[15:40, 0, 25], -e[15 - x0:40 - x0], rtol=0, atol=1e-13)


def test_fresnel_reflection_of_dielectric_half_space():
    """Normal incidence on a thick, tall eps_r = 4 block (n = 2) that lies
    inside the TF box (a TF/SF face must sit in the vacuum background): the
    front-face reflection leaving the TF box has amplitude |r| = (n-1)/(n+1)
    = 1/3 of the incident pulse.  The window closes before the block's edge
    diffraction, its back-face echo and the PEC-wall echo reach the probe."""
    nx, nz = 400, 500
    p = default_params(is_planar=1, nx=nx, ny=1, nz=nz, dy=0.0, use_pw=1, incident=LX_INC_AUX1D,
                       tf_lo=(60, 0, 20), tf_hi=(340, 0, 479), num_timesteps=830, pw_delay=3.0,
                       pw_wave=LX_WAVE_GAUSSIAN, pw_fwhm=2.0, src_comp=LX_EY)
    p.e_pw[:] = (0.0, 1.0, 0.0)
    ids = np.zeros((nx, 1, nz), np.uint8)
    ids[200:300, 0, 30:470] = 1
    s = O.OracleSim(p, threads=4)
    s.set_materials(ids, [1.0, 4.0])
    s.set_probes([[LX_EY, 30, 0, 250], [LX_EY, 120, 0, 250]])
    s.step(830)
    pr = s.probes()
    inc = pr[:450, 1]                          # incident pulse passing x = 120
    inc_peak = inc[int(np.argmax(np.abs(inc)))]
    refl = pr[:, 0]                            # only the reflection reaches x = 30 (SF region)
    refl_peak = refl[int(np.argmax(np.abs(refl)))]
    assert refl_peak / inc_peak == pytest.approx(-1 / 3, rel=0.01)   # r = (1-n)/(1+n) < 0


def test_hard_cw_point_source_semantics():
    """E(src) = int(amp)*sin(int(omega)*t*dt) is assigned at the start of step t
    (omega*t is an int product that wraps, maxwell.h:203,206), then the E update
    adds Cb*curl H at that cell."""
    omega = 1_000_000_000
    p = default_params(nx=12, ny=12, nz=12, use_ps=1, src_kind=LX_SRC_HARD_CW, src_amp=3.9,
                       src_omega=float(omega), src_pos=(6, 6, 6))
    s = O.OracleSim(p, threads=1)
    eps0, mu0, c0 = consts()
    dt = 0.5 * DX / c0
    for t in range(5):
        s.step(1)
        Hx, Hy = s.field(LX_HX), s.field(LX_HY)
        curl = (Hy[6, 6, 6] - Hy[5, 6, 6]) / DX - (Hx[6, 6, 6] - Hx[6, 5, 6]) / DX
        prod = (omega * t + 2 ** 31) % 2 ** 32 - 2 ** 31       # 32-bit wrap at t >= 3
        hard = 3 * math.sin(prod * dt)
        assert s.field(LX_EZ)[6, 6, 6] == pytest.approx(hard + dt / eps0 * curl, rel=1e-12, abs=1e-22)
    assert abs(hard) > 1e-8


# ---------------------------------------------------------------------------
# mirror symmetry of the cube workload (used at full size on the GPU)
# ---------------------------------------------------------------------------

def test_slab_workload_mirror_parities():
    """The cube workload's grid, CPML and aux TF/SF plane wave (along +x, Ez
    polarised) with the eps 6 cube stretched over y and z into a slab: the
    discrete problem is symmetric about the y and z centre planes, and the
    FP64 oracle's fields are exact mirror images with the parities that the
    reflections of E (polar) and H (axial) predict.  (The cube itself is not
    symmetric on the grid: a node's material sets the coefficient of all
    three E components at that array index, maxwell_three_dim_pml.cc:138-148,
    whatever their stagger, so the cube is half a cell longer for the
    staggered component along each axis.)  The same check flags one planted
    wrong cell and a wrong parity."""
    from conftest import MIRROR_PARITY, mirror_asymmetry
    from luxum_amd import configs
    w = configs.config1(n=32, steps=110, pml=6, margin=4)
    ids = np.zeros_like(w.ids)
    ids[12:20] = 1
    o = O.OracleSim(w.p)
    o.set_materials(ids, w.eps_r)
    o.step(110)
    for c in w.p.components():
        f = o.field(c)
        assert np.linalg.norm(f) > 0, c
        py, pz = MIRROR_PARITY[c]
        assert mirror_asymmetry(f, 1, py) == 0.0, (c, "y")
        assert mirror_asymmetry(f, 2, pz) == 0.0, (c, "z")
        assert mirror_asymmetry(f, 1, -py) > 1.9, (c, "y parity flipped")
        assert mirror_asymmetry(f, 2, -pz) > 1.9, (c, "z parity flipped")
        g = f.copy()
        g[5, 7, 9] += 1e-3 * np.abs(f).max()      # one wrong cell off the centre planes
        assert mirror_asymmetry(g, 1, py) > 1e-5
    o2 = O.OracleSim(w.p)                         # the cube: not a mirror image
    o2.set_materials(w.ids, w.eps_r)
    o2.step(110)
    assert mirror_asymmetry(o2.field(2), 2, +1) > 1e-3
    o.close()
    o2.close()


def test_fp32_storage_build_is_the_same_algorithm():
    """The FP32-storage build of the oracle (every stored value rounded to
    float, FP64 arithmetic) follows the FP64 checker to FP32 rounding on the
    cube workload, and is not the FP64 build (its fields differ)."""
    from luxum_amd import configs
    w = configs.config1(n=32, steps=80, pml=6, margin=4)
    runs = []
    for storage in ("fp64", "fp32"):
        o = O.OracleSim(w.p, threads=4, storage=storage)
        o.set_materials(w.ids, w.eps_r)
        o.step(80)
        runs.append([o.field(c) for c in range(6)])
        o.close()
    num = sum(float(np.sum((a - b) ** 2)) for a, b in zip(*runs))
    den = sum(float(np.sum(a ** 2)) for a in runs[0])
    err = (num / den) ** 0.5
    assert 1e-9 < err < 1e-6
    for a, b in zip(*runs):
        assert np.all(b == b.astype(np.float32))     # every stored value is a float
