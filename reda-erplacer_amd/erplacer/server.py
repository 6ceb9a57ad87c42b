"""Serving form: many independent placements on one GPU with continuous batching.

``Server`` wraps the erp_server_* entry points of liberp_gpu.so.  A worker thread
inside the library admits submitted designs while job slots and the device
memory budget allow, keeps their state resident, advances the jobs that share a
bin grid together in one captured hipGraph of ``check_every`` batched
iterations (``eager=True``: the same kernels launched directly, for tracers),
and retires a job once its overflow drops below its threshold or it
reaches its iteration limit.  A job's result equals
``Placement.run_until(max_iterations, stop_overflow, check_every)`` on the same
design alone, bit for bit.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as L


def estimate_device_bytes(design, num_bins, target_density=float("nan"), **cfg) -> int:
    """Device memory a context for this design would hold (upper estimate)."""
    if isinstance(num_bins, int):
        num_bins = (num_bins, num_bins)
    ds = L.design_struct(design)
    c = L.default_config(num_bins_x=num_bins[0], num_bins_y=num_bins[1],
                         target_density=target_density, **cfg)
    v = C.c_int64()
    L.check(L.load().erp_estimate_device_bytes(C.byref(ds), C.byref(c), C.byref(v)))
    return v.value


class Server:
    def __init__(self, hbm_budget_bytes: int | None = None, max_active: int | None = None,
                 check_every: int | None = None, device: int = 0, eager: bool = False):
        self._lib = L.load()
        cfg = L.ServerConfig()
        self._lib.erp_server_config_default(C.byref(cfg))
        if hbm_budget_bytes is not None:
            cfg.hbm_budget_bytes = int(hbm_budget_bytes)
        if max_active is not None:
            cfg.max_active = int(max_active)
        if check_every is not None:
            cfg.check_every = int(check_every)
        cfg.device = device
        cfg.eager = 1 if eager else 0
        self.config = cfg
        h = C.c_void_p()
        L.check(self._lib.erp_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._nodes = {}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.erp_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, design, num_bins, max_iterations: int, stop_overflow: float,
               target_density=float("nan"), gamma_scale=4.0, density_weight0=8e-5) -> int:
        if isinstance(num_bins, int):
            num_bins = (num_bins, num_bins)
        ds = L.design_struct(design)
        c = L.default_config(num_bins_x=num_bins[0], num_bins_y=num_bins[1],
                             target_density=target_density, gamma_scale=gamma_scale,
                             density_weight0=density_weight0, spectral_mode=L.SPECTRAL_FFT,
                             device=self.config.device)
        jid = C.c_int64()
        L.check(self._lib.erp_server_submit(self._h, C.byref(ds), C.byref(c),
                                            C.c_int64(max_iterations), C.c_double(stop_overflow),
                                            C.byref(jid)))
        self._nodes[jid.value] = int(design.num_nodes)
        return jid.value

    def poll(self, capacity: int = 256) -> list[dict]:
        buf = (L.JobResult * capacity)()
        n = C.c_int32()
        L.check(self._lib.erp_server_poll(self._h, buf, C.c_int32(capacity), C.byref(n)))
        return [buf[k].as_dict() for k in range(n.value)]

    def wait(self, job_id: int) -> dict:
        r = L.JobResult()
        L.check(self._lib.erp_server_wait(self._h, C.c_int64(job_id), C.byref(r)))
        return r.as_dict()

    def drain(self):
        L.check(self._lib.erp_server_drain(self._h))

    def positions(self, job_id: int):
        n = self._nodes[job_id]
        x = np.empty(n, np.float32)
        y = np.empty(n, np.float32)
        L.check(self._lib.erp_server_get_positions(self._h, C.c_int64(job_id),
                                                   x.ctypes.data_as(C.c_void_p),
                                                   y.ctypes.data_as(C.c_void_p)))
        return x, y

    def release(self, job_id: int):
        L.check(self._lib.erp_server_release(self._h, C.c_int64(job_id)))
        self._nodes.pop(job_id, None)

    def stats(self) -> dict:
        st = L.ServerStat()
        L.check(self._lib.erp_server_stats(self._h, C.byref(st)))
        return st.as_dict()
