"""Serving: continuous batching of simulation requests (lx_server_* in the C ABI).

A ``Server`` takes requests (a workload: parameters, material ids, eps_r,
probes; a step count; an arrival time in seconds after the start of ``run``),
admits arrived requests while their device footprints fit under the HBM cap,
and advances every active request together tick by tick.  ``plain=True``
runs the same requests one at a time with the single-simulation kernels.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import Request, RequestInfo, ServerStats, check, lib
from .params import LxParams
from .sim import Context, default_context

LX_SERVE_BATCHED, LX_SERVE_PLAIN = 0, 1
LX_REQ_QUEUED, LX_REQ_ACTIVE, LX_REQ_DONE, LX_REQ_FAILED = 0, 1, 2, 3


def footprint(p: LxParams, batched: bool = True, steps: int = 0, n_probes: int = 0) -> int:
    """Upper bound of the device bytes of a simulation with parameters p."""
    out = C.c_int64()
    check(lib().lx_params_footprint(C.byref(p), int(batched), int(steps), int(n_probes), C.byref(out)))
    return out.value


@dataclass
class Result:
    state: int
    steps_done: int
    arrival: float
    admitted: float
    finished: float
    footprint: int
    cells: int

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class Server:
    def __init__(self, ctx: Context | None = None, hbm_cap: int = 200 * 10 ** 9, max_active: int = 1024,
                 plain: bool = False):
        self.ctx = ctx or default_context()
        self._h = C.c_void_p()
        check(lib().lx_server_create(self.ctx._h, LX_SERVE_PLAIN if plain else LX_SERVE_BATCHED, int(hbm_cap),
                                     int(max_active), C.byref(self._h)))
        self._meta: list[tuple] = []   # (params, n_probes, steps) per request id

    def submit(self, p: LxParams, ids: np.ndarray, eps_r, steps: int, arrival: float = 0.0, probes=None,
               keep_fields: bool = False) -> int:
        ids = np.ascontiguousarray(ids, dtype=np.uint8)
        eps = np.ascontiguousarray(eps_r, dtype=np.float64)
        pr = None if probes is None else np.ascontiguousarray(np.asarray(probes, dtype=np.int32).reshape(-1, 4))
        req = Request()
        req.params = C.cast(C.pointer(p), C.c_void_p)
        req.ids = ids.ctypes.data
        req.n_cells = ids.size
        req.eps_r = eps.ctypes.data
        req.n_ids = eps.size
        req.probes = pr.ctypes.data if pr is not None and pr.size else None
        req.n_probes = 0 if pr is None else pr.shape[0]
        req.steps = int(steps)
        req.arrival = float(arrival)
        req.keep_fields = int(keep_fields)
        rid = C.c_int64()
        check(lib().lx_server_submit(self._h, C.byref(req), C.byref(rid)))
        self._meta.append((p, req.n_probes, int(steps)))
        return rid.value

    def run(self) -> None:
        check(lib().lx_server_run(self._h))

    def result(self, rid: int) -> Result:
        inf = RequestInfo()
        check(lib().lx_server_info(self._h, int(rid), C.byref(inf)))
        return Result(inf.state, inf.steps_done, inf.arrival, inf.admitted, inf.finished, inf.footprint, inf.cells)

    def probes(self, rid: int) -> np.ndarray:
        p, n_probes, steps = self._meta[rid]
        out = np.empty((steps, n_probes), dtype=np.float64)
        check(lib().lx_server_probes(self._h, int(rid), out.ctypes.data, out.size))
        return out

    def field(self, rid: int, comp: int) -> np.ndarray:
        p = self._meta[rid][0]
        out = np.empty(p.field_shape(comp), dtype=np.float64)
        check(lib().lx_server_field(self._h, int(rid), int(comp), out.ctypes.data, out.size))
        return out

    def stats(self) -> ServerStats:
        st = ServerStats()
        check(lib().lx_server_get_stats(self._h, C.byref(st)))
        return st

    def close(self) -> None:
        if self._h:
            lib().lx_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Server", "Result", "footprint", "LX_REQ_DONE", "LX_REQ_FAILED"]
