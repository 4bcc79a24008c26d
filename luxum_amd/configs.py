"""Workload builders for the benchmark configurations of the Yee timestep.

Every builder returns a ``Workload`` (parameters, uint8 material ids in the
reference's C order, the eps_r table and probe points).  ``n``/``steps``
arguments shrink a configuration for parity tests while keeping its
structure (CPML thickness, TF/SF margin, relative geometry).

Units: lambda0 = 1 um, dx = lambda0/20 = 50 nm, Courant S = 0.5.
CPML grading: m = 3, sigma_max = 0.8(m+1)/(eta0*dx) (ratio 1 to the
reference's sigma_optimal, maxwell.cc:83-84), kappa_max = 5, a_max = 0.05,
linear a.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .params import (LX_EY, LX_EZ, LX_FP16, LX_FP32, LX_INC_AUX1D, LX_SRC_SOFT, LX_WAVE_GAUSSIAN_SINE,
                     LxParams, default_params)

LAMBDA0 = 1e-6
DX = LAMBDA0 / 20


@dataclasses.dataclass
class Workload:
    name: str
    p: LxParams
    ids: np.ndarray          # uint8, shape p.dims, C order (k fastest)
    eps_r: np.ndarray        # float64 table
    probes: np.ndarray       # int32 (n, 4): component, i, j, k

    @property
    def cells(self) -> int:
        nx, ny, nz = self.p.dims
        return nx * ny * nz


def _cpml(p: LxParams, cells: int) -> None:
    p.use_pml = 1
    p.pml_thickness = cells * p.dx
    p.sigma_ratio = 1.0
    p.sigma_m = 3.0
    p.kappa_max = 5.0
    p.a_max = 0.05
    p.a_m = 1.0


def config0(n: int = 512, steps: int = 1000, precision: int = LX_FP32) -> Workload:
    """2D TM, n x n, CPML 16, TF/SF box 20 cells inside the CPML, normally
    incident Gaussian-modulated sinusoid (carrier lambda0, FWHM 3 periods),
    centred (n/4)^2 square of n = 3.3 (eps_r = 10.89)."""
    p = default_params(is_planar=1, nx=n, ny=1, nz=n, dx=DX, dy=0.0, dz=DX, courant=0.5, num_timesteps=steps,
                       precision=precision)
    _cpml(p, 16)
    p.use_pw = 1
    p.incident = LX_INC_AUX1D
    p.k_pw[:] = (1.0, 0.0, 0.0)
    p.e_pw[:] = (0.0, 1.0, 0.0)
    lo, hi = 16 + 20, n - 1 - 16 - 20
    p.tf_lo[:] = (lo, 0, lo)
    p.tf_hi[:] = (hi, 0, hi)
    p.pw_wave = LX_WAVE_GAUSSIAN_SINE
    p.pw_lambda0 = LAMBDA0
    p.pw_fwhm = 3.0
    p.pw_delay = 4.0
    p.src_comp = LX_EY
    ids = np.zeros((n, 1, n), dtype=np.uint8)
    a, b = n // 2 - n // 8, n // 2 + n // 8
    ids[a:b, 0, a:b] = 1
    eps = np.array([1.0, 3.3 ** 2])
    c = n // 2
    probes = np.array([[LX_EY, n * 3 // 16, 0, c], [LX_EY, c, 0, c], [LX_EY, n * 13 // 16, 0, c],
                       [LX_EY, c, 0, n * 13 // 16]], dtype=np.int32)
    return Workload("2D TM CPML %dx%d" % (n, n), p, ids, eps, probes)


def config1(n: int = 256, steps: int = 2000, precision: int = LX_FP32, pml: int = 12, margin: int = 16) -> Workload:
    """3D Yee, n^3, CPML 12 on all faces, TF/SF box 16 cells inside the CPML,
    plane wave along +x, Ez polarised Gaussian-modulated pulse, centred
    dielectric cube of side n/4 with eps_r = 6, 8 Ez probes."""
    p = default_params(is_planar=0, nx=n, ny=n, nz=n, dx=DX, dy=DX, dz=DX, courant=0.5, num_timesteps=steps,
                       precision=precision)
    _cpml(p, pml)
    p.use_pw = 1
    p.incident = LX_INC_AUX1D
    p.k_pw[:] = (1.0, 0.0, 0.0)
    p.e_pw[:] = (0.0, 0.0, 1.0)
    lo, hi = pml + margin, n - 1 - pml - margin
    p.tf_lo[:] = (lo, lo, lo)
    p.tf_hi[:] = (hi, hi, hi)
    p.pw_wave = LX_WAVE_GAUSSIAN_SINE
    p.pw_lambda0 = LAMBDA0
    p.pw_fwhm = 3.0
    p.pw_delay = 4.0
    ids = np.zeros((n, n, n), dtype=np.uint8)
    a, b = n // 2 - n // 8, n // 2 + n // 8
    ids[a:b, a:b, a:b] = 1
    eps = np.array([1.0, 6.0])
    q0, q1 = n * 11 // 32, n * 21 // 32
    probes = np.array([[LX_EZ, x, y, z] for x in (q0, q1) for y in (q0, q1) for z in (q0, q1)], dtype=np.int32)
    return Workload("3D CPML cube %d^3" % n, p, ids, eps, probes)


def box3d(nx: int, ny: int, nz: int, pml: int = 8, margin: int = 5, precision: int = LX_FP32,
          steps: int = 400) -> Workload:
    """3D CPML grid of uneven extents: config 1's excitation (aux-1D plane
    wave along +x, Ez polarised, TF/SF `margin` cells inside the CPML), a
    dielectric block (eps_r 4) off the centre and 4 Ez probes."""
    p = default_params(is_planar=0, nx=nx, ny=ny, nz=nz, dx=DX, dy=DX, dz=DX, courant=0.5, num_timesteps=steps,
                       precision=precision)
    _cpml(p, pml)
    p.use_pw = 1
    p.incident = LX_INC_AUX1D
    p.k_pw[:] = (1.0, 0.0, 0.0)
    p.e_pw[:] = (0.0, 0.0, 1.0)
    lo = pml + margin
    p.tf_lo[:] = (lo, lo, lo)
    p.tf_hi[:] = (nx - 1 - lo, ny - 1 - lo, nz - 1 - lo)
    p.pw_wave = LX_WAVE_GAUSSIAN_SINE
    p.pw_lambda0 = LAMBDA0
    p.pw_fwhm = 3.0
    p.pw_delay = 4.0
    ids = np.zeros((nx, ny, nz), dtype=np.uint8)
    ids[nx // 3:nx // 2 + 7, ny // 3:ny // 2 + 3, nz // 3:nz // 2 + 5] = 1
    eps = np.array([1.0, 4.0])
    probes = np.array([[LX_EZ, x, ny // 2, z] for x in (nx // 4, 3 * nx // 4) for z in (nz // 3, 2 * nz // 3)],
                      dtype=np.int32)
    return Workload("3D CPML box %dx%dx%d" % (nx, ny, nz), p, ids, eps, probes)


def config2(nxy: int = 1280, nz: int = 640, steps: int = 4000, precision: int = LX_FP32, pitch: int = 32,
            side: int = 13, slab: int = 64) -> Workload:
    """3D photonic-crystal slab waveguide: square lattice of square rods
    (eps_r 12, pitch 32, side 13) in a 64-cell slab in air, one rod row
    removed at mid-y, soft Ez dipole Gaussian-modulated at a/lambda = 0.3,
    CPML 16."""
    p = default_params(is_planar=0, nx=nxy, ny=nxy, nz=nz, dx=DX, dy=DX, dz=DX, courant=0.5, num_timesteps=steps,
                       precision=precision)
    _cpml(p, 16)
    lam = pitch * DX / 0.3
    p.use_ps = 1
    p.src_kind = LX_SRC_SOFT
    p.src_comp = LX_EZ
    p.src_amp = 1.0
    p.src_wave = LX_WAVE_GAUSSIAN_SINE
    p.src_lambda0 = lam
    p.src_fwhm = 3.0
    p.src_delay = 4.0
    c = nxy // 2
    p.src_pos[:] = (c, c, nz // 2)
    z0, z1 = nz // 2 - slab // 2, nz // 2 + slab // 2
    xs = np.arange(nxy)
    in_rod_1d = (xs % pitch) < side
    row = (xs // pitch)
    defect_row = (c // pitch)
    mask_x = in_rod_1d[:, None]
    mask_y = (in_rod_1d & (row != defect_row))[None, :]
    rods = mask_x & mask_y
    ids = np.zeros((nxy, nxy, nz), dtype=np.uint8)
    ids[:, :, z0:z1] = rods[:, :, None].astype(np.uint8)
    eps = np.array([1.0, 12.0])
    probes = np.array([[LX_EZ, c + d, c, nz // 2] for d in (0, 64, 128, 256)], dtype=np.int32)
    return Workload("3D PhC slab %dx%dx%d" % (nxy, nxy, nz), p, ids, eps, probes)


def phc2d(n: int, seed: int, steps: int = 3000, precision: int = LX_FP32) -> Workload:
    """One simulation of the batched 2D workload: a random photonic crystal of
    circular rods (eps_r ~ U[2,12], pitch 16-40 cells, radius ~ U[0.2,0.45]
    pitch, 0-2 rows removed), TF/SF plane wave as config 0, CPML 16, with a
    transmitted-flux probe line of Ey near the far side."""
    rng = np.random.default_rng(seed)
    w = config0(n=n, steps=steps, precision=precision)
    p = w.p
    pitch = int(rng.integers(16, 41))
    radius = float(rng.uniform(0.2, 0.45)) * pitch
    er = float(rng.uniform(2.0, 12.0))
    lo, hi = p.tf_lo[0], p.tf_hi[0]
    ii, kk = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x0, x1 = lo + 16, hi - 16
    cx = (ii - x0) % pitch - pitch / 2
    ck = (kk - lo) % pitch - pitch / 2
    rod = (cx ** 2 + ck ** 2 <= radius ** 2) & (ii >= x0) & (ii < x1) & (kk >= lo) & (kk <= hi)
    rows = (kk - lo) // pitch
    nrem = int(rng.integers(0, 3))
    nrows = max(int((hi - lo) // pitch), 1)
    for r in rng.choice(nrows, size=min(nrem, nrows), replace=False):
        rod &= rows != r
    ids = rod.astype(np.uint8)[:, None, :]
    eps = np.array([1.0, er])
    xp = hi - 8
    ks = np.linspace(lo + 4, hi - 4, 8).astype(np.int32)
    probes = np.array([[LX_EY, xp, 0, k] for k in ks], dtype=np.int32)
    return Workload("2D PhC %d^2 seed %d" % (n, seed), p, ids, eps, probes)


def batch_sizes(count: int = 256, seed: int = 2024):
    rng = np.random.default_rng(seed)
    sizes = np.arange(256, 2049, 128)
    return [int(x) for x in rng.choice(sizes, size=count)]


def scatterer2d(n: int, rng, precision: int = LX_FP32, steps: int = 3000) -> Workload:
    """2D request of the served stream: config 0's set-up (CPML 16, TF/SF
    plane wave +x, Ey) around one dielectric square or disk, eps_r ~
    U[1.5, 12], size 10-40% of the domain, at a random place inside the TF
    box."""
    w = config0(n=n, steps=steps, precision=precision)
    lo, hi = w.p.tf_lo[0], w.p.tf_hi[0]
    size = max(2, int(rng.uniform(0.1, 0.4) * n))
    size = min(size, hi - lo - 2)
    ci, ck = (int(rng.integers(lo + 1 + size // 2, hi - size // 2)) for _ in range(2))
    er = float(rng.uniform(1.5, 12.0))
    ii, kk = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    if rng.random() < 0.5:
        body = (np.abs(ii - ci) <= size // 2) & (np.abs(kk - ck) <= size // 2)
        shape = "square"
    else:
        body = (ii - ci) ** 2 + (kk - ck) ** 2 <= (size / 2) ** 2
        shape = "disk"
    ids = body.astype(np.uint8)[:, None, :]
    return Workload("2D %s %d^2 eps %.2f" % (shape, n, er), w.p, ids, np.array([1.0, er]), w.probes)


def scatterer3d(n: int, rng, precision: int = LX_FP32, steps: int = 3000) -> Workload:
    """3D request of the served stream: config 1's set-up (CPML 12, TF/SF 16
    inside, plane wave +x, Ez) around one dielectric cube or sphere, eps_r ~
    U[1.5, 12], size 10-40% of the domain."""
    w = config1(n=n, steps=steps, precision=precision)
    lo, hi = w.p.tf_lo[0], w.p.tf_hi[0]
    size = max(2, int(rng.uniform(0.1, 0.4) * n))
    size = min(size, hi - lo - 2)
    c = [int(rng.integers(lo + 1 + size // 2, hi - size // 2)) for _ in range(3)]
    er = float(rng.uniform(1.5, 12.0))
    ax = np.arange(n)
    if rng.random() < 0.5:
        m = [np.abs(ax - c[a]) <= size // 2 for a in range(3)]
        body = m[0][:, None, None] & m[1][None, :, None] & m[2][None, None, :]
        shape = "cube"
    else:
        d = [(ax - c[a]) ** 2 for a in range(3)]
        body = (d[0][:, None, None] + d[1][None, :, None] + d[2][None, None, :]) <= (size / 2) ** 2
        shape = "sphere"
    return Workload("3D %s %d^3 eps %.2f" % (shape, n, er), w.p, body.astype(np.uint8), np.array([1.0, er]),
                    w.probes)


def serve_stream(count: int, rate: float, seed: int = 7, steps_range=(1000, 5000), frac3d: float = 0.2,
                 frac_fp16: float = 0.5, sizes2d=None, sizes3d=None):
    """Seeded request stream of the served workload: Poisson arrivals at
    `rate` requests/s; 80% 2D (256^2-2048^2), 20% 3D (64^3-320^3) scatterers;
    step counts uniform over steps_range; FP32 or FP16 storage.  Returns a
    list of (workload, steps, arrival seconds)."""
    rng = np.random.default_rng(seed)
    sizes2d = list(range(256, 2049, 128)) if sizes2d is None else list(sizes2d)
    sizes3d = list(range(64, 321, 32)) if sizes3d is None else list(sizes3d)
    t = 0.0
    out = []
    for _ in range(count):
        t += float(rng.exponential(1.0 / rate))
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        prec = LX_FP16 if rng.random() < frac_fp16 else LX_FP32
        if rng.random() < frac3d:
            w = scatterer3d(int(rng.choice(sizes3d)), rng, prec, steps)
        else:
            w = scatterer2d(int(rng.choice(sizes2d)), rng, prec, steps)
        out.append((w, steps, t))
    return out


def workload_bytes_per_step(p: LxParams, slab_cells) -> int:
    """Algorithmic HBM bytes of one full timestep: every stored field value is
    read twice and written once, plus the uint8 id and the CPML psi traffic
    of the slab cells (4 psi values per slab cell in 3D, 2 in 2D, read and
    written, FP32)."""
    s = 2 if p.precision == LX_FP16 else 4
    nx, ny, nz = p.dims
    n = nx * ny * nz
    if p.is_planar:
        return n * (9 * s + 1) + sum(slab_cells) * 16
    return n * (18 * s + 1) + sum(slab_cells) * 32


__all__ = ["Workload", "config0", "config1", "config2", "phc2d", "batch_sizes", "workload_bytes_per_step",
           "scatterer2d", "scatterer3d", "serve_stream",
           "LX_FP16", "LX_FP32"]
