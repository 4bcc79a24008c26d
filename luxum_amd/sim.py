"""Python host API over the C ABI: Context, Simulation, Batch.

Method names follow the reference's solver vocabulary where one exists
(step = one iteration of the solve* time loop, field() = the vfield
components the reference writes with fileio).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import SimInfo, check, lib
from .params import LxParams, default_params


class Context:
    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        check(lib().lx_ctx_create(device, C.byref(self._h)))
        self.device = device

    def mem_info(self):
        fr, tot = C.c_int64(), C.c_int64()
        check(lib().lx_ctx_mem_info(self._h, C.byref(fr), C.byref(tot)))
        return fr.value, tot.value

    def close(self):
        if self._h:
            lib().lx_ctx_destroy(self._h)
            self._h = C.c_void_p()


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class Simulation:
    """One Yee simulation resident on the GPU."""

    def __init__(self, p: LxParams, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.p = p.copy()
        self._h = C.c_void_p()
        check(lib().lx_sim_create(self.ctx._h, C.byref(self.p), C.byref(self._h)))
        self.n_probes = 0

    # lifecycle ---------------------------------------------------------
    def close(self):
        if self._h:
            lib().lx_sim_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # set-up --------------------------------------------------------------
    def set_materials(self, ids: np.ndarray, eps_r) -> None:
        """eps_r: one relative permittivity per id (every E component), or an
        (n_ids, 3) array of the permittivities that Ex, Ey and Ez take."""
        ids = np.ascontiguousarray(ids, dtype=np.uint8)
        eps = np.ascontiguousarray(eps_r, dtype=np.float64)
        nx, ny, nz = self.p.dims
        if ids.size != nx * ny * nz:
            raise ValueError(f"ids has {ids.size} entries, grid has {nx * ny * nz}")
        if eps.ndim == 2:
            if eps.shape[1] != 3:
                raise ValueError("per-component eps_r must have shape (n_ids, 3)")
            check(lib().lx_sim_set_materials3(self._h, ids.ctypes.data, eps.ctypes.data, eps.shape[0]))
        else:
            check(lib().lx_sim_set_materials(self._h, ids.ctypes.data, eps.ctypes.data, eps.size))

    def set_probes(self, probes) -> None:
        arr = np.ascontiguousarray(np.asarray(probes, dtype=np.int32).reshape(-1, 4))
        check(lib().lx_sim_set_probes(self._h, arr.ctypes.data if arr.size else None, arr.shape[0]))
        self.n_probes = arr.shape[0]

    # time loop -----------------------------------------------------------
    def step(self, n: int) -> None:
        check(lib().lx_sim_step(self._h, int(n)))

    def sync(self) -> None:
        check(lib().lx_sim_sync(self._h))

    # outputs -------------------------------------------------------------
    def field(self, comp: int) -> np.ndarray:
        out = np.empty(self.p.field_shape(comp), dtype=np.float64)
        check(lib().lx_sim_read_field(self._h, comp, out.ctypes.data, out.size))
        return out

    def snapshot_async(self, comp: int, out: np.ndarray) -> None:
        assert out.dtype == np.float64 and out.flags.c_contiguous
        check(lib().lx_sim_snapshot_async(self._h, comp, out.ctypes.data, out.size))

    def probes(self) -> np.ndarray:
        info = self.info()
        n = max(info.step, 1) * max(self.n_probes, 1)
        buf = np.zeros(n, dtype=np.float64)
        ns = C.c_int32()
        check(lib().lx_sim_read_probes(self._h, buf.ctypes.data, buf.size, C.byref(ns)))
        return buf[: ns.value * self.n_probes].reshape(ns.value, self.n_probes).copy()

    def monitor(self):
        mx, en = C.c_double(), C.c_double()
        check(lib().lx_sim_monitor(self._h, C.byref(mx), C.byref(en)))
        return mx.value, en.value

    def max_e2(self) -> float:
        """The reference monitor alone (max|E|^2), from the E components."""
        mx = C.c_double()
        check(lib().lx_sim_max_e2(self._h, C.byref(mx)))
        return mx.value

    def max_e2_start(self) -> None:
        """Start the reference monitor of the current state beside the next step."""
        check(lib().lx_sim_max_e2_start(self._h))

    def max_e2_finish(self) -> float:
        """The started monitor's max|E|^2 (waits for it)."""
        mx = C.c_double()
        check(lib().lx_sim_max_e2_finish(self._h, C.byref(mx)))
        return mx.value

    def info(self) -> SimInfo:
        inf = SimInfo()
        check(lib().lx_sim_get_info(self._h, C.byref(inf)))
        return inf

    def stream(self) -> int:
        return lib().lx_sim_stream(self._h) or 0

    def autotune(self) -> int:
        kz = C.c_int32()
        check(lib().lx_sim_autotune(self._h, C.byref(kz)))
        return kz.value

    def time_kernels(self, n: int):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        check(lib().lx_sim_time_kernels(self._h, int(n), C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value


class Batch:
    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self._h = C.c_void_p()
        check(lib().lx_batch_create(self.ctx._h, C.byref(self._h)))
        self.sims: list[Simulation] = []

    def add(self, sim: Simulation) -> None:
        check(lib().lx_batch_add(self._h, sim._h))
        self.sims.append(sim)

    def remove(self, sim: Simulation) -> None:
        check(lib().lx_batch_remove(self._h, sim._h))
        self.sims.remove(sim)

    def step(self, n: int) -> None:
        check(lib().lx_batch_step(self._h, int(n)))

    def sync(self) -> None:
        check(lib().lx_batch_sync(self._h))

    def close(self):
        if self._h:
            lib().lx_batch_destroy(self._h)
            self._h = C.c_void_p()


LX_OUT_FIELD, LX_OUT_E2 = 0, 1


class Writer:
    """Asynchronous HDF5 snapshot writer in the reference's layout (one file,
    datasets "t=<t>"): LX_OUT_FIELD writes one component, LX_OUT_E2 the 3D
    solvers' |E|^2 at equal indices."""

    def __init__(self, sim: "Simulation", path: str, what: int = LX_OUT_FIELD, comp: int = 2):
        self.sim = sim
        self._h = C.c_void_p()
        check(lib().lx_writer_open(sim._h, path.encode(), int(what), int(comp), C.byref(self._h)))

    def write(self, t: float) -> None:
        check(lib().lx_writer_write(self._h, float(t)))

    def count(self) -> int:
        return int(lib().lx_writer_count(self._h))

    def close(self) -> None:
        if self._h:
            h, self._h = self._h, C.c_void_p()
            check(lib().lx_writer_close(h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def h5_dataset_name(t: float) -> str:
    buf = C.create_string_buffer(64)
    check(lib().lx_h5_dataset_name(float(t), buf, 64))
    return buf.value.decode()


def h5_write(path: str, t, data: np.ndarray, append: bool = False) -> None:
    """Write dataset "t=<t>" (t a number) or the dataset named t (a str)."""
    a = np.ascontiguousarray(data, dtype=np.float64)
    if a.ndim != 3:
        raise ValueError("datasets are 3D")
    dims = (C.c_int64 * 3)(*a.shape)
    if isinstance(t, str):
        check(lib().lx_h5_write_dataset(path.encode(), t.encode(), a.ctypes.data, dims, int(append)))
    else:
        check(lib().lx_h5_write(path.encode(), float(t), a.ctypes.data, dims, int(append)))


def h5_read(path: str, dataset: str) -> np.ndarray:
    dims = (C.c_int64 * 3)()
    check(lib().lx_h5_read(path.encode(), dataset.encode(), None, 0, dims))
    out = np.empty(tuple(dims), dtype=np.float64)
    check(lib().lx_h5_read(path.encode(), dataset.encode(), out.ctypes.data, out.size, dims))
    return out


def load_yaml(path: str) -> LxParams:
    """reader::readYAML + checkData (lib/io/reader.cc:68-162) through the C ABI."""
    p = default_params()
    check(lib().lx_params_load_yaml(path.encode(), C.byref(p)))
    return p


def params_check(p: LxParams) -> None:
    check(lib().lx_params_check(C.byref(p)))


__all__ = ["Context", "Simulation", "Batch", "Writer", "load_yaml", "params_check", "default_context", "h5_read",
           "h5_write", "h5_dataset_name", "LX_OUT_FIELD", "LX_OUT_E2", "_lib"]
