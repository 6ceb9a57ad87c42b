"""Host-side mirror of the reference's placement loop and leaf functions.

``Placement`` stands for ``Placement`` of placement.rs:142-228: ``new`` builds
the device context (Bins, PlacementSolution, PotentialComputation,
HpwlComputation, Scheduler, Nesterov), ``step`` is ``Placement::step``.  The
module-level functions are the reference's leaf functions with the same names
and argument meaning (density.rs, dct2d.rs, idct2.rs, idct_idxst.rs,
idxst_idct.rs, eforce.rs).  Everything runs in liberp_gpu.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as L


class Placement:
    """placement.rs:142-228 on the GPU."""

    def __init__(self, design, num_bins=512, target_density=float("nan"),
                 spectral_mode=L.SPECTRAL_FFT, device=0, gamma_scale=4.0, density_weight0=8e-5):
        if isinstance(num_bins, int):
            num_bins = (num_bins, num_bins)
        self.design = design
        self.m, self.n = num_bins
        self._ds = L.design_struct(design)
        self.config = L.default_config(num_bins_x=self.m, num_bins_y=self.n,
                                       target_density=target_density, spectral_mode=spectral_mode,
                                       device=device, gamma_scale=gamma_scale,
                                       density_weight0=density_weight0)
        self._lib = L.load()
        h = C.c_void_p()
        L.check(self._lib.erp_ctx_create(C.byref(self._ds), C.byref(self.config), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.erp_ctx_destroy(self._h)
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

    # ------------------------------------------------------------- the loop
    def step(self) -> dict:
        """Placement::step (placement.rs:180-228)."""
        info = L.StepInfo()
        L.check(self._lib.erp_step(self._h, C.byref(info)))
        return info.as_dict()

    def run(self, iterations: int) -> dict:
        """`iterations` steps, replayed from captured hipGraphs."""
        info = L.StepInfo()
        L.check(self._lib.erp_run(self._h, C.c_int64(iterations), C.byref(info)))
        return info.as_dict()

    def run_until(self, max_iterations: int, stop_overflow: float, check_every: int = 10) -> dict:
        info = L.StepInfo()
        L.check(self._lib.erp_run_until(self._h, C.c_int64(max_iterations), C.c_double(stop_overflow),
                                        C.c_int32(check_every), C.byref(info)))
        return info.as_dict()

    def reset(self):
        L.check(self._lib.erp_ctx_reset(self._h))

    # ------------------------------------------------------------- phases
    def compute_hpwl(self, gamma: float | None = None) -> float:
        """compute_hpwl (hpwl.rs:139) with the scheduler's gamma or an explicit one."""
        v = C.c_double()
        if gamma is None:
            L.check(self._lib.erp_compute_hpwl(self._h, C.byref(v)))
        else:
            L.check(self._lib.erp_compute_hpwl_gamma(self._h, C.c_double(gamma), C.byref(v)))
        return v.value

    def compute_potential(self):
        """compute_potential (epotential.rs:225) -> (energy, overflow)."""
        e, o = C.c_double(), C.c_double()
        L.check(self._lib.erp_compute_potential(self._h, C.byref(e), C.byref(o)))
        return e.value, o.value

    def nesterov_step(self):
        L.check(self._lib.erp_nesterov_step(self._h))

    def scheduler_update(self):
        L.check(self._lib.erp_scheduler_update(self._h))

    # ------------------------------------------------------------- state
    def info(self) -> dict:
        info = L.StepInfo()
        L.check(self._lib.erp_get_step_info(self._h, C.byref(info)))
        return info.as_dict()

    def positions(self):
        N = self.design.num_nodes
        x = np.zeros(N, np.float32)
        y = np.zeros(N, np.float32)
        L.check(self._lib.erp_get_positions(self._h, L._ptr(x), L._ptr(y)))
        return x, y

    def set_positions(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        L.check(self._lib.erp_set_positions(self._h, L._ptr(x), L._ptr(y)))

    def map(self, which: int) -> np.ndarray:
        out = np.zeros(self.m * self.n, np.float32)
        L.check(self._lib.erp_get_map(self._h, C.c_int(which), L._ptr(out)))
        return out.reshape(self.m, self.n)

    def cell_grads(self, which: int):
        C_ = self.design.num_movable
        gx = np.zeros(C_, np.float32)
        gy = np.zeros(C_, np.float32)
        L.check(self._lib.erp_get_cell_grads(self._h, C.c_int(which), L._ptr(gx), L._ptr(gy)))
        return gx, gy

    def set_keep_density(self, keep: bool):
        L.check(self._lib.erp_set_keep_density(self._h, C.c_int(int(keep))))

    def set_compute_potential(self, on: bool):
        """Potential map and energy every step (default on); off skips one of the
        three inverse transforms -- Placement::step discards both (placement.rs:185)."""
        L.check(self._lib.erp_set_compute_potential(self._h, C.c_int(1 if on else 0)))

    def set_scatter_kind(self, kind: int):
        """Scatter kernel of the loop (lib.SCATTER_AUTO / _CELL_MAJOR / _OWNER)."""
        L.check(self._lib.erp_set_scatter_kind(self._h, C.c_int(kind)))

    def scatter_kind(self) -> int:
        v = C.c_int32()
        L.check(self._lib.erp_get_scatter_kind(self._h, C.byref(v)))
        return v.value

    def set_store_gradients(self, store: bool):
        L.check(self._lib.erp_set_store_gradients(self._h, C.c_int(int(store))))

    def density_fixed(self, movable=True):
        B = self.m * self.n
        qm = np.zeros(B, np.int64) if movable else None
        qf = np.zeros(B, np.int64)
        sc = C.c_double()
        L.check(self._lib.erp_get_density_fixed(self._h, L._ptr(qm) if movable else None,
                                                L._ptr(qf), C.byref(sc)))
        return (qm.reshape(self.m, self.n) if movable else None), qf.reshape(self.m, self.n), sc.value

    def setup(self) -> dict:
        N = self.design.num_nodes
        arrs = [np.zeros(N, np.float32) for _ in range(5)]
        sc = np.zeros(4, np.float64)
        L.check(self._lib.erp_get_setup(self._h, *[L._ptr(a) for a in arrs], L._ptr(sc)))
        cw, ch, ox, oy, ratio = arrs
        return dict(cw=cw, ch=ch, ox=ox, oy=oy, ratio=ratio, bw=sc[0], bh=sc[1],
                    target_density=sc[2], fx_scale=sc[3])

    def final_hpwl_dbu(self) -> int:
        v = C.c_int64()
        L.check(self._lib.erp_final_hpwl_dbu(self._h, C.byref(v)))
        return v.value

    def device_bytes(self) -> int:
        """Device memory this context holds."""
        v = C.c_int64()
        L.check(self._lib.erp_ctx_device_bytes(self._h, C.byref(v)))
        return v.value

    def set_timing(self, enabled: bool):
        L.check(self._lib.erp_set_timing(self._h, C.c_int(int(enabled))))

    def timing(self) -> dict:
        cap = 64
        ms = (C.c_double * cap)()
        names = (C.c_char_p * cap)()
        cnt = C.c_int32()
        L.check(self._lib.erp_get_timing(self._h, ms, C.c_int32(cap), C.byref(cnt), names))
        return {names[i].decode(): ms[i] for i in range(cnt.value)}


# ------------------------------------------------------------------ leaf functions
def _spectral(fn, x, mode):
    x = np.ascontiguousarray(x, np.float32)
    m, n = x.shape
    out = np.zeros((m, n), np.float32)
    L.check(getattr(L.load(), fn)(L._ptr(x), L._ptr(out), C.c_int32(m), C.c_int32(n), C.c_int32(mode)))
    return out


def dct2_fft2_forward(x, mode=L.SPECTRAL_FFT):
    """dct2d.rs:256-279: dctn(x, type 2) / (m n)."""
    return _spectral("erp_dct2_fft2_forward", x, mode)


def idct2_fft2_forward(x, mode=L.SPECTRAL_FFT):
    """idct2.rs:144-166: dctn(x, type 3)."""
    return _spectral("erp_idct2_fft2_forward", x, mode)


def idct_idxst_forward(x, mode=L.SPECTRAL_FFT):
    """idct_idxst.rs:137-161."""
    return _spectral("erp_idct_idxst_forward", x, mode)


def idxst_idct_forward(x, mode=L.SPECTRAL_FFT):
    """idxst_idct.rs:156-180."""
    return _spectral("erp_idxst_idct_forward", x, mode)


def _f32(*arrs):
    return [np.ascontiguousarray(a, np.float32) for a in arrs]


def compute_density_map(x, y, cw, ch, ox, oy, ratio, m, n, bin_w, bin_h, scale, initial=None):
    """density.rs:240-299 movable scatter (fixed-point accumulation) + initial.
    Returns (map, int64 accumulators)."""
    x, y, cw, ch, ox, oy, ratio = _f32(x, y, cw, ch, ox, oy, ratio)
    nc = x.shape[0]
    out = np.zeros(m * n, np.float32)
    q = np.zeros(m * n, np.int64)
    init = None if initial is None else np.ascontiguousarray(initial, np.float32).reshape(-1)
    L.check(L.load().erp_compute_density_map(
        C.c_int64(nc), L._ptr(x), L._ptr(y), L._ptr(cw), L._ptr(ch), L._ptr(ox), L._ptr(oy),
        L._ptr(ratio), C.c_int32(m), C.c_int32(n), C.c_float(bin_w), C.c_float(bin_h),
        C.c_double(scale), None if init is None else L._ptr(init), L._ptr(out), L._ptr(q)))
    return out.reshape(m, n), q.reshape(m, n)


def compute_initial_density_map(x, y, cw, ch, m, n, bin_w, bin_h, die_w, die_h, scale):
    """density.rs:301-325 (fixed cells against bin centres).  Returns (map, q)."""
    x, y, cw, ch = _f32(x, y, cw, ch)
    out = np.zeros(m * n, np.float32)
    q = np.zeros(m * n, np.int64)
    L.check(L.load().erp_compute_initial_density_map(
        C.c_int64(x.shape[0]), L._ptr(x), L._ptr(y), L._ptr(cw), L._ptr(ch), C.c_int32(m),
        C.c_int32(n), C.c_float(bin_w), C.c_float(bin_h), C.c_float(die_w), C.c_float(die_h),
        C.c_double(scale), L._ptr(out), L._ptr(q)))
    return out.reshape(m, n), q.reshape(m, n)


def compute_electric_force(field_x, field_y, x, y, cw, ch, ox, oy, ratio, bin_w, bin_h):
    """eforce.rs:15-91."""
    fx, fy = _f32(field_x, field_y)
    m, n = fx.shape
    x, y, cw, ch, ox, oy, ratio = _f32(x, y, cw, ch, ox, oy, ratio)
    nc = x.shape[0]
    gx = np.zeros(nc, np.float32)
    gy = np.zeros(nc, np.float32)
    L.check(L.load().erp_compute_electric_force(
        C.c_int64(nc), C.c_int32(m), C.c_int32(n), L._ptr(fx), L._ptr(fy), L._ptr(x), L._ptr(y),
        L._ptr(cw), L._ptr(ch), L._ptr(ox), L._ptr(oy), L._ptr(ratio), C.c_float(bin_w),
        C.c_float(bin_h), L._ptr(gx), L._ptr(gy)))
    return gx, gy
