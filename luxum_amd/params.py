"""Simulation parameters: ctypes mirror of ``lx_params`` (include/luxum.h).

The first five groups of fields mirror class ``reader`` of the reference
(lib/io/reader.h:52-167, filled by reader::readYAML, lib/io/reader.cc:68-130);
the last group is the optional sixth document of parameters.yaml.
"""
from __future__ import annotations

import ctypes as C
import math

LX_EX, LX_EY, LX_EZ, LX_HX, LX_HY, LX_HZ = range(6)
COMP_NAMES = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
LX_FP32, LX_FP16 = 0, 1
LX_INC_ANALYTIC, LX_INC_AUX1D = 0, 1
LX_WAVE_GAUSSIAN, LX_WAVE_GAUSSIAN_SINE = 0, 1
LX_SRC_HARD_CW, LX_SRC_SOFT = 0, 1

# Yee stagger (x, y, z) per component (lib/field/vfield.cc:75-84).
STAGGER = ((1, 0, 0), (0, 1, 0), (0, 0, 1), (0, 1, 1), (1, 0, 1), (1, 1, 0))


class LxParams(C.Structure):
    _fields_ = [
        ("is_planar", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("courant", C.c_double),
        ("num_timesteps", C.c_int32),
        ("npx", C.c_int32), ("npy", C.c_int32), ("n_threads", C.c_int32),
        ("use_ps", C.c_int32),
        ("src_pos", C.c_int32 * 3),
        ("src_amp", C.c_double), ("src_omega", C.c_double),
        ("use_pw", C.c_int32),
        ("k_pw", C.c_double * 3),
        ("e_pw", C.c_double * 3),
        ("off_pw", C.c_double),
        ("tf_lo", C.c_int32 * 3),
        ("tf_hi", C.c_int32 * 3),
        ("use_pml", C.c_int32),
        ("pml_thickness", C.c_double),
        ("sigma_ratio", C.c_double),
        ("sigma_m", C.c_double),
        ("kappa_max", C.c_double),
        ("a_max", C.c_double),
        ("a_m", C.c_double),
        ("precision", C.c_int32),
        ("incident", C.c_int32),
        ("pw_sigma", C.c_double),
        ("pw_wave", C.c_int32),
        ("pw_lambda0", C.c_double), ("pw_fwhm", C.c_double), ("pw_delay", C.c_double),
        ("src_kind", C.c_int32), ("src_comp", C.c_int32), ("src_wave", C.c_int32),
        ("src_lambda0", C.c_double), ("src_fwhm", C.c_double), ("src_delay", C.c_double),
        ("snapshot_every", C.c_int32),
        ("monitor_every", C.c_int32),
        ("storage_scale", C.c_double),
    ]

    def copy(self) -> "LxParams":
        q = LxParams()
        C.pointer(q)[0] = self
        return q

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out

    # convenience ------------------------------------------------------
    @property
    def dims(self):
        return (self.nx, 1 if self.is_planar else self.ny, self.nz)

    def field_shape(self, comp: int):
        """Host (reference) extents of a component, C order, z fastest."""
        nx, ny, nz = self.dims
        if self.is_planar:
            return {LX_EY: (nx, 1, nz), LX_HX: (nx, 1, nz - 1), LX_HZ: (nx - 1, 1, nz)}[comp]
        s = STAGGER[comp]
        return (nx - s[0], ny - s[1], nz - s[2])

    def components(self):
        return (LX_EY, LX_HX, LX_HZ) if self.is_planar else tuple(range(6))


def default_params(**over) -> LxParams:
    """Same values as lx_params_default() in the C ABI."""
    p = LxParams()
    p.is_planar = 0
    p.nx = p.ny = p.nz = 32
    p.dx = p.dy = p.dz = 5e-8
    p.courant = 0.5
    p.num_timesteps = 100
    p.npx = p.npy = p.n_threads = 1
    p.use_ps = 0
    p.src_amp = 0.0
    p.src_omega = 0.0
    p.use_pw = 0
    p.k_pw[:] = (1.0, 0.0, 0.0)
    p.e_pw[:] = (0.0, 0.0, 1.0)
    p.off_pw = 0.0
    p.use_pml = 0
    p.pml_thickness = 0.0
    p.sigma_ratio = 1.0
    p.sigma_m = 3.0
    p.kappa_max = 5.0
    p.a_max = 0.05
    p.a_m = 1.0
    p.precision = LX_FP32
    p.incident = LX_INC_ANALYTIC
    p.pw_sigma = 1e-4 / 0.5
    p.pw_wave = LX_WAVE_GAUSSIAN_SINE
    p.pw_lambda0 = 1e-6
    p.pw_fwhm = 3.0
    p.pw_delay = 4.0
    p.src_kind = LX_SRC_HARD_CW
    p.src_comp = LX_EZ
    p.src_wave = LX_WAVE_GAUSSIAN_SINE
    p.src_lambda0 = 1e-6
    p.src_fwhm = 3.0
    p.src_delay = 4.0
    p.snapshot_every = 0
    p.monitor_every = 0
    p.storage_scale = 0.0
    for k, v in over.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        if isinstance(v, (tuple, list)):
            getattr(p, k)[:] = v
        else:
            setattr(p, k, v)
    return p


def constants(p: LxParams):
    """eps0, mu0, c, dt as src/solvers/maxwell.cc:63-67 computes them."""
    eps0 = 8.85 * math.pow(10, -12)
    mu0 = math.pi * 4 * math.pow(10, -7)
    c = 1 / math.pow(eps0 * mu0, 0.5)
    dt = p.courant * p.dx / c
    return eps0, mu0, c, dt
