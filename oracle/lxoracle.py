"""ctypes wrapper of the FP64 oracle (liblxoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from luxum_amd.params import LxParams

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblxoracle.so")
# the same restatement with every stored value rounded to float (FP64
# arithmetic): the precision floor of FP32 storage
LIB_PATH_F32S = os.path.join(HERE, "liblxoracle_f32s.so")

SIGMA, KAPPA, ALPHA, B, CC = range(5)

_libs = {}


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(storage: str = "fp64"):
    L = _libs.get(storage)
    if L is None:
        path = LIB_PATH if storage == "fp64" else LIB_PATH_F32S
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.POINTER(LxParams)
        vp = C.c_void_p
        L.lxo_create.argtypes = [P, C.c_int, C.POINTER(vp)]
        L.lxo_destroy.argtypes = [vp]
        L.lxo_set_materials.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_int]
        L.lxo_set_materials3.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_int]
        L.lxo_set_probes.argtypes = [vp, C.c_void_p, C.c_int]
        L.lxo_step.argtypes = [vp, C.c_int]
        L.lxo_read_field.argtypes = [vp, C.c_int, C.c_void_p, C.c_size_t]
        L.lxo_set_field.argtypes = [vp, C.c_int, C.c_void_p, C.c_size_t]
        L.lxo_read_probes.argtypes = [vp, C.c_void_p, C.c_size_t, C.POINTER(C.c_int)]
        L.lxo_field_size.argtypes = [P, C.c_int, C.c_int * 3]
        L.lxo_profile.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int]
        L.lxo_constants.argtypes = [P, C.c_double * 6]
        L.lxo_max_e2.argtypes = [vp]
        L.lxo_max_e2.restype = C.c_double
        L.lxo_energy.argtypes = [vp]
        L.lxo_energy.restype = C.c_double
        L.lxo_aux_line.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        _libs[storage] = L
    return L


def profile(p: LxParams, dim: int, is_e: bool, kind: int) -> np.ndarray:
    n = (p.nx, p.ny, p.nz)[dim] - (0 if is_e else 1)
    out = np.zeros(n, dtype=np.float64)
    rc = lib().lxo_profile(C.byref(p), dim, int(is_e), kind, out.ctypes.data, n)
    assert rc == 0
    return out


def ref_constants(p: LxParams):
    a = (C.c_double * 6)()
    lib().lxo_constants(C.byref(p), a)
    return dict(eps0=a[0], mu0=a[1], c=a[2], dt=a[3], sigma_opt=a[4], sigma_max=a[5])


class OracleSim:
    """FP64 reference-structured simulation on the CPU."""

    def __init__(self, p: LxParams, threads: int | None = None, storage: str = "fp64"):
        """storage: "fp64" (the checker) or "fp32" (stored values rounded to
        float, FP64 arithmetic: the floor of FP32 storage)."""
        self.p = p.copy()
        self._h = C.c_void_p()
        self._L = lib("fp64" if storage == "fp64" else "fp32")
        th = threads if threads is not None else min(os.cpu_count() or 1, 16)
        rc = self._L.lxo_create(C.byref(self.p), th, C.byref(self._h))
        if rc != 0:
            raise ValueError(f"lxo_create failed ({rc})")
        self.n_probes = 0

    def close(self):
        if self._h:
            self._L.lxo_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_materials(self, ids: np.ndarray, eps_r):
        """eps_r: one permittivity per id, or an (n_ids, 3) array of the
        permittivities that Ex, Ey and Ez take for each id."""
        ids = np.ascontiguousarray(ids, dtype=np.uint8)
        eps = np.ascontiguousarray(eps_r, dtype=np.float64)
        assert ids.size == self.p.dims[0] * self.p.dims[1] * self.p.dims[2]
        if eps.ndim == 2:
            assert eps.shape[1] == 3
            rc = self._L.lxo_set_materials3(self._h, ids.ctypes.data, eps.ctypes.data, eps.shape[0])
        else:
            rc = self._L.lxo_set_materials(self._h, ids.ctypes.data, eps.ctypes.data, eps.size)
        if rc != 0:
            raise ValueError("bad material ids")

    def set_probes(self, probes):
        arr = np.ascontiguousarray(np.asarray(probes, dtype=np.int32).reshape(-1, 4))
        rc = self._L.lxo_set_probes(self._h, arr.ctypes.data, arr.shape[0])
        if rc != 0:
            raise ValueError("bad probes")
        self.n_probes = arr.shape[0]

    def step(self, n: int):
        self._L.lxo_step(self._h, int(n))

    def field(self, comp: int) -> np.ndarray:
        shape = self.p.field_shape(comp)
        out = np.empty(shape, dtype=np.float64)
        rc = self._L.lxo_read_field(self._h, comp, out.ctypes.data, out.size)
        if rc != 0:
            raise ValueError("bad component")
        return out

    def set_field(self, comp: int, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == tuple(self.p.field_shape(comp))
        rc = self._L.lxo_set_field(self._h, comp, a.ctypes.data, a.size)
        if rc != 0:
            raise ValueError("bad component")

    def probes(self) -> np.ndarray:
        cap = 1 << 20
        buf = np.empty(cap * max(self.n_probes, 1), dtype=np.float64)
        ns = C.c_int()
        rc = self._L.lxo_read_probes(self._h, buf.ctypes.data, buf.size, C.byref(ns))
        assert rc == 0
        return buf[: ns.value * self.n_probes].reshape(ns.value, self.n_probes).copy()

    def max_e2(self) -> float:
        return self._L.lxo_max_e2(self._h)

    def energy(self) -> float:
        return self._L.lxo_energy(self._h)

    def aux_line(self):
        n = 1 << 16
        e = np.zeros(n)
        h = np.zeros(n)
        x0 = C.c_int()
        L = self._L.lxo_aux_line(self._h, e.ctypes.data, h.ctypes.data, n, C.byref(x0))
        if L < 0:
            return None
        return e[:L].copy(), h[: L - 1].copy(), x0.value
