"""ctypes access to the CPU oracle (oracle/liberp_oracle.so) for the tests.

The oracle is the checker.  Only tests/, the smoke check and bench.py's CPU
baseline load it; the product library never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liberp_oracle.so")

PI_F32, PI_EXACT = 0, 1
MAP_DENSITY, MAP_AUV, MAP_FIELD_X, MAP_FIELD_Y, MAP_POTENTIAL, MAP_INITIAL = range(6)
GRAD_WL, GRAD_DENSITY = 0, 1

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)


class EroDesign(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_int64), ("num_movable", C.c_int64),
        ("x", _fp), ("y", _fp), ("w", _fp), ("h", _fp), ("num_pins", _fp),
        ("num_pins_total", C.c_int64),
        ("pin_offset_x", _fp), ("pin_offset_y", _fp), ("pin_node", _ip),
        ("num_nets", C.c_int64), ("n_2nets", C.c_int64),
        ("net_start", _ip), ("net_pins", _ip),
        ("node_pin_start", _ip), ("node_pins", _ip),
        ("die_w", C.c_double), ("die_h", C.c_double),
        ("movable_area", C.c_double), ("cell_utilization", C.c_double),
    ]


class EroConfig(C.Structure):
    _fields_ = [("num_bins_x", C.c_int64), ("num_bins_y", C.c_int64),
                ("target_density", C.c_double), ("pi_mode", C.c_int),
                ("nthreads", C.c_int), ("fixed_point", C.c_int)]


class EroStepOut(C.Structure):
    _fields_ = [("hpwl", C.c_double), ("overflow", C.c_double), ("energy", C.c_double),
                ("density_weight", C.c_double), ("gamma", C.c_double),
                ("alpha", C.c_double), ("beta", C.c_double),
                ("iteration", C.c_int64), ("should_stop", C.c_int)]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.ero_fx_scale.restype = C.c_double
        for sfx in ("_f64", "_f32", "_f32d"):
            getattr(_lib, "ero_create" + sfx).restype = C.c_void_p
            getattr(_lib, "ero_compute_hpwl" + sfx).restype = C.c_double if sfx == "_f64" else C.c_float
            getattr(_lib, "ero_wa_value" + sfx).restype = C.c_double if sfx == "_f64" else C.c_float
    return _lib


def _dt(sfx):
    return np.float64 if sfx == "_f64" else np.float32


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _real(sfx):
    return C.c_double if sfx == "_f64" else C.c_float


def make_design_struct(d):
    """EroDesign over the arrays of a generator Design (keeps references alive)."""
    keep = []

    def f(a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        keep.append(a)
        return a.ctypes.data_as(_fp)

    def i(a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        keep.append(a)
        return a.ctypes.data_as(_ip)

    s = EroDesign(
        num_nodes=d.num_nodes, num_movable=d.num_movable,
        x=f(d.x), y=f(d.y), w=f(d.w), h=f(d.h), num_pins=f(d.num_pins),
        num_pins_total=d.num_pin_total,
        pin_offset_x=f(d.pin_offset_x), pin_offset_y=f(d.pin_offset_y), pin_node=i(d.pin_node),
        num_nets=d.num_nets, n_2nets=d.n_2nets,
        net_start=i(d.net_start), net_pins=i(d.net_pins),
        node_pin_start=i(d.node_pin_start), node_pins=i(d.node_pins),
        die_w=d.die_w, die_h=d.die_h, movable_area=d.movable_area,
        cell_utilization=d.cell_utilization)
    s._keep = keep
    return s


# ---------------------------------------------------------------- leaf wrappers
def make_expk(n, pi_mode=PI_F32, sfx="_f64"):
    dt = _dt(sfx)
    re = np.zeros(n, dt)
    im = np.zeros(n, dt)
    getattr(lib(), "ero_make_expk" + sfx)(C.c_int64(n), C.c_int(pi_mode), _p(re), _p(im))
    return re + 1j * im


def frequency_matrices(m, n, bw, bh, sfx="_f64"):
    dt = _dt(sfx)
    a, b, c = (np.zeros(m * n, dt) for _ in range(3))
    R = _real(sfx)
    getattr(lib(), "ero_frequency_matrices" + sfx)(C.c_int64(m), C.c_int64(n), R(bw), R(bh),
                                                   _p(a), _p(b), _p(c))
    return a.reshape(m, n), b.reshape(m, n), c.reshape(m, n)


def fft(z, inverse=False, sfx="_f64"):
    dt = _dt(sfx)
    re = np.ascontiguousarray(z.real, dt).copy()
    im = np.ascontiguousarray(z.imag, dt).copy()
    getattr(lib(), "ero_fft" + sfx)(_p(re), _p(im), C.c_int64(z.shape[0]), C.c_int(int(inverse)))
    return re + 1j * im


def _spectral(name, x, pi_mode=PI_F32, nthreads=1, sfx="_f64"):
    dt = _dt(sfx)
    x = np.ascontiguousarray(x, dt)
    m, n = x.shape
    out = np.zeros((m, n), dt)
    getattr(lib(), name + sfx)(_p(x), _p(out), C.c_int64(m), C.c_int64(n), C.c_int(pi_mode),
                               C.c_int(nthreads))
    return out


def dct2_fft2_forward(x, **kw):
    return _spectral("ero_dct2_fft2_forward", x, **kw)


def idct2_fft2_forward(x, **kw):
    return _spectral("ero_idct2_fft2_forward", x, **kw)


def idct_idxst_forward(x, **kw):
    return _spectral("ero_idct_idxst_forward", x, **kw)


def idxst_idct_forward(x, **kw):
    return _spectral("ero_idxst_idct_forward", x, **kw)


def dct2d_preprocess(x, sfx="_f64"):
    dt = _dt(sfx)
    x = np.ascontiguousarray(x, dt)
    m, n = x.shape
    out = np.zeros(m * n, dt)
    getattr(lib(), "ero_dct2d_preprocess" + sfx)(_p(x), _p(out), C.c_int64(m), C.c_int64(n))
    return out.reshape(m, n)


def irfft_2d(spec, n, nthreads=1, sfx="_f64"):
    dt = _dt(sfx)
    m = spec.shape[0]
    re = np.ascontiguousarray(spec.real, dt).copy()
    im = np.ascontiguousarray(spec.imag, dt).copy()
    out = np.zeros((m, n), dt)
    getattr(lib(), "ero_irfft_2d" + sfx)(_p(re), _p(im), _p(out), C.c_int64(m), C.c_int64(n),
                                         C.c_int(nthreads))
    return out


def _cells(args, dt):
    return [np.ascontiguousarray(a, dt) for a in args]


def density_map(kind, x, y, sx, sy, ox, oy, ratio, m, n, bw, bh, initial, nthreads=1, sfx="_f64"):
    dt = _dt(sfx)
    R = _real(sfx)
    x, y, sx, sy, ox, oy, ratio = _cells((x, y, sx, sy, ox, oy, ratio), dt)
    initial = np.ascontiguousarray(initial, dt).reshape(-1)
    out = np.zeros(m * n, dt)
    args = [C.c_int64(x.shape[0]), _p(x), _p(y), _p(sx), _p(sy), _p(ox), _p(oy), _p(ratio),
            C.c_int64(m), C.c_int64(n), R(bw), R(bh), _p(initial), _p(out)]
    if kind == "sequential":
        getattr(lib(), "ero_density_map_sequential" + sfx)(*args)
    else:
        getattr(lib(), "ero_density_map_parallel" + sfx)(*args, C.c_int(nthreads))
    return out.reshape(m, n)


def density_map_fx64(x, y, sx, sy, ox, oy, ratio, m, n, bw, bh, scale, sfx="_f32"):
    dt = _dt(sfx)
    R = _real(sfx)
    x, y, sx, sy, ox, oy, ratio = _cells((x, y, sx, sy, ox, oy, ratio), dt)
    q = np.zeros(m * n, np.int64)
    getattr(lib(), "ero_density_map_fx64" + sfx)(
        C.c_int64(x.shape[0]), _p(x), _p(y), _p(sx), _p(sy), _p(ox), _p(oy), _p(ratio),
        C.c_int64(m), C.c_int64(n), R(bw), R(bh), C.c_double(scale), _p(q))
    return q.reshape(m, n)


def initial_density_map(x, y, sx, sy, m, n, bw, bh, die_w, die_h, nthreads=1, sfx="_f64"):
    dt = _dt(sfx)
    R = _real(sfx)
    x, y, sx, sy = _cells((x, y, sx, sy), dt)
    out = np.zeros(m * n, dt)
    getattr(lib(), "ero_initial_density_map" + sfx)(
        C.c_int64(x.shape[0]), _p(x), _p(y), _p(sx), _p(sy), C.c_int64(m), C.c_int64(n),
        R(bw), R(bh), R(die_w), R(die_h), C.c_int(nthreads), _p(out))
    return out.reshape(m, n)


def initial_density_fx64(x, y, sx, sy, m, n, bw, bh, die_w, die_h, scale, sfx="_f32"):
    dt = _dt(sfx)
    R = _real(sfx)
    x, y, sx, sy = _cells((x, y, sx, sy), dt)
    q = np.zeros(m * n, np.int64)
    getattr(lib(), "ero_initial_density_fx64" + sfx)(
        C.c_int64(x.shape[0]), _p(x), _p(y), _p(sx), _p(sy), C.c_int64(m), C.c_int64(n),
        R(bw), R(bh), R(die_w), R(die_h), C.c_double(scale), _p(q))
    return q.reshape(m, n)


def fx_scale(mass_bound):
    return lib().ero_fx_scale(C.c_double(mass_bound))


def fx_finalize(q_mov, q_fix, scale, coef):
    q_mov = np.ascontiguousarray(q_mov, np.int64).reshape(-1)
    q_fix = np.ascontiguousarray(q_fix, np.int64).reshape(-1)
    rho = np.zeros(q_mov.shape[0], np.float32)
    ovf = C.c_int64(0)
    lib().ero_fx_finalize_f32(C.c_int64(q_mov.shape[0]), _p(q_mov), _p(q_fix),
                              C.c_double(scale), C.c_double(coef), _p(rho), C.byref(ovf))
    return rho, ovf.value


def electric_force(fx, fy, x, y, sx, sy, ox, oy, ratio, bw, bh, nthreads=1, sfx="_f64"):
    dt = _dt(sfx)
    R = _real(sfx)
    m, n = fx.shape
    fx = np.ascontiguousarray(fx, dt)
    fy = np.ascontiguousarray(fy, dt)
    x, y, sx, sy, ox, oy, ratio = _cells((x, y, sx, sy, ox, oy, ratio), dt)
    nc = x.shape[0]
    gx = np.zeros(nc, dt)
    gy = np.zeros(nc, dt)
    getattr(lib(), "ero_compute_electric_force" + sfx)(
        C.c_int64(nc), C.c_int64(m), C.c_int64(n), _p(fx), _p(fy), _p(x), _p(y), _p(sx),
        _p(sy), _p(ox), _p(oy), _p(ratio), R(bw), R(bh), _p(gx), _p(gy), C.c_int(nthreads))
    return gx, gy


def compute_hpwl(ds, d, x, y, gamma, nthreads=1, sfx="_f64"):
    """ds: EroDesign; d: generator Design.  x, y: positions of all nodes."""
    dt = _dt(sfx)
    R = _real(sfx)
    x = np.ascontiguousarray(x, dt)
    y = np.ascontiguousarray(y, dt)
    gx = np.zeros(d.num_movable, dt)
    gy = np.zeros(d.num_movable, dt)
    pgx = np.zeros(d.num_pin_total, dt)
    pgy = np.zeros(d.num_pin_total, dt)
    h = getattr(lib(), "ero_compute_hpwl" + sfx)(C.byref(ds), _p(x), _p(y), R(gamma), _p(gx),
                                                _p(gy), _p(pgx), _p(pgy), C.c_int(nthreads))
    return float(h), gx, gy, pgx, pgy


def wa_value(ds, x, y, gamma, sfx="_f64"):
    dt = _dt(sfx)
    R = _real(sfx)
    x = np.ascontiguousarray(x, dt)
    y = np.ascontiguousarray(y, dt)
    return float(getattr(lib(), "ero_wa_value" + sfx)(C.byref(ds), _p(x), _p(y), R(gamma)))


class Placer:
    """The oracle's Placement::step loop (placement.rs:180-228)."""

    def __init__(self, d, grid, target_density=float("nan"), sfx="_f64", nthreads=1,
                 fixed_point=False, pi_mode=PI_F32):
        self.d = d
        self.sfx = sfx
        self.dt = _dt(sfx)
        self.ds = make_design_struct(d)
        if isinstance(grid, int):
            grid = (grid, grid)
        self.m, self.n = grid
        self.cfg = EroConfig(num_bins_x=self.m, num_bins_y=self.n,
                             target_density=target_density, pi_mode=pi_mode,
                             nthreads=nthreads, fixed_point=int(fixed_point))
        self.h = getattr(lib(), "ero_create" + sfx)(C.byref(self.ds), C.byref(self.cfg))

    def __del__(self):
        if getattr(self, "h", None):
            getattr(lib(), "ero_destroy" + self.sfx)(C.c_void_p(self.h))
            self.h = None

    def step(self):
        out = EroStepOut()
        getattr(lib(), "ero_step" + self.sfx)(C.c_void_p(self.h), C.byref(out))
        return {k: getattr(out, k) for k, _ in EroStepOut._fields_}

    def positions(self):
        x = np.zeros(self.d.num_nodes, self.dt)
        y = np.zeros(self.d.num_nodes, self.dt)
        getattr(lib(), "ero_get_positions" + self.sfx)(C.c_void_p(self.h), _p(x), _p(y))
        return x, y

    def set_positions(self, x, y):
        x = np.ascontiguousarray(x, self.dt)
        y = np.ascontiguousarray(y, self.dt)
        getattr(lib(), "ero_set_positions" + self.sfx)(C.c_void_p(self.h), _p(x), _p(y))

    def map(self, which):
        out = np.zeros(self.m * self.n, self.dt)
        getattr(lib(), "ero_get_map" + self.sfx)(C.c_void_p(self.h), C.c_int(which), _p(out))
        return out.reshape(self.m, self.n)

    def cell_grads(self, which):
        gx = np.zeros(self.d.num_movable, self.dt)
        gy = np.zeros(self.d.num_movable, self.dt)
        getattr(lib(), "ero_get_cell_grads" + self.sfx)(C.c_void_p(self.h), C.c_int(which),
                                                       _p(gx), _p(gy))
        return gx, gy

    def setup(self):
        N = self.d.num_nodes
        arrs = [np.zeros(N, self.dt) for _ in range(5)]
        sc = np.zeros(6, np.float64)
        getattr(lib(), "ero_get_setup" + self.sfx)(C.c_void_p(self.h), *[_p(a) for a in arrs],
                                                  _p(sc))
        cw, ch, ox, oy, ratio = arrs
        return dict(cw=cw, ch=ch, ox=ox, oy=oy, ratio=ratio, bw=sc[0], bh=sc[1],
                    target_density=sc[2], fx_scale=sc[3], gamma=sc[4], density_weight=sc[5])


def scheduler_sequence(hpwl, ovf, base_gamma, dw0, sfx="_f64"):
    dt = _dt(sfx)
    R = _real(sfx)
    hpwl = np.ascontiguousarray(hpwl, dt)
    ovf = np.ascontiguousarray(ovf, dt)
    n = hpwl.shape[0]
    dw = np.zeros(n, dt)
    gm = np.zeros(n, dt)
    stop = np.zeros(n, np.int32)
    getattr(lib(), "ero_scheduler_sequence" + sfx)(C.c_int64(n), _p(hpwl), _p(ovf), R(base_gamma),
                                                  R(dw0), _p(dw), _p(gm), _p(stop))
    return dw, gm, stop
