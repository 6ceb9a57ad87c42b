"""Host-side checks that need no GPU: the synthetic netlist honours the reda-db
contract the loop reads, the workloads have the stated shapes, and the HIP
library loads and exports every function include/erp_gpu.h declares."""
import ctypes
import math

import numpy as np
import pytest

from erplacer import lib as L
from erplacer import workloads as W
from erplacer.netlist import generate


@pytest.fixture(scope="module")
def d():
    return generate(5000, 5250, 64, 20, 0.10, 0.7, 20261203)


def test_contract(d):
    C, N, P, E = d.num_movable, d.num_nodes, d.num_pin_total, d.num_nets
    deg = np.diff(d.net_start)
    assert (deg[:d.n_2nets] == 2).all() and (deg[d.n_2nets:] != 2).all()   # hpwl.rs:173-181
    assert deg.min() >= 2 and deg.max() <= 64
    assert d.net_start[0] == 0 and d.net_start[-1] == P
    assert np.array_equal(np.sort(d.net_pins), np.arange(P))                 # each pin in one net
    for e in range(0, E, 97):
        nodes = d.pin_node[d.net_pins[d.net_start[e]:d.net_start[e + 1]]]
        assert len(set(nodes.tolist())) == nodes.shape[0]                     # distinct cells
    # macro_2_pins CSR consistent with pin_node, and num_pins as float
    counts = np.bincount(d.pin_node, minlength=N)
    assert np.array_equal(np.diff(d.node_pin_start), counts)
    assert np.array_equal(d.num_pins, counts.astype(np.float32))
    for c in range(0, N, 101):
        pins = d.node_pins[d.node_pin_start[c]:d.node_pin_start[c + 1]]
        assert (d.pin_node[pins] == c).all()
    # pin offsets inside their instance
    assert (d.pin_offset_x >= 0).all() and (d.pin_offset_x <= d.w[d.pin_node]).all()
    assert (d.pin_offset_y >= 0).all() and (d.pin_offset_y <= d.h[d.pin_node]).all()
    # cells: 1..8 sites wide, one row tall; all fit the die with the -1 clamp
    assert set(np.unique(d.w[:C] / 200).astype(int).tolist()) <= set(range(1, 9))
    assert (d.h[:C] == 2000).all()
    assert (d.die_w - d.w[:C] - 1 >= 0).all()


def test_macros(d):
    C, N = d.num_movable, d.num_nodes
    rect = np.stack([d.x[C:], d.y[C:], d.w[C:], d.h[C:]], 1).astype(np.float64)
    assert (rect[:, 0] >= 0).all() and (rect[:, 1] >= 0).all()
    assert (rect[:, 0] + rect[:, 2] <= d.die_w).all() and (rect[:, 1] + rect[:, 3] <= d.die_h).all()
    for i in range(rect.shape[0]):
        for j in range(i + 1, rect.shape[0]):
            a, b = rect[i], rect[j]
            assert not (a[0] < b[0] + b[2] and b[0] < a[0] + a[2] and a[1] < b[1] + b[3] and b[1] < a[1] + a[3])
    frac = (rect[:, 2] * rect[:, 3]).sum() / (d.die_w * d.die_h)
    assert 0.07 < frac < 0.11
    assert math.isclose(d.cell_utilization, d.movable_area / (d.die_w * d.die_h - (rect[:, 2] * rect[:, 3]).sum()))
    assert 0.66 < d.cell_utilization < 0.72


def test_degree_mix(d):
    deg = np.diff(d.net_start)
    assert abs((deg == 2).mean() - 0.60) < 0.01 and abs((deg == 3).mean() - 0.25) < 0.02


def test_seeded_and_deterministic():
    a = generate(800, 840, 32, 3, 0.1, 0.7, 5)
    b = generate(800, 840, 32, 3, 0.1, 0.7, 5)
    c = generate(800, 840, 32, 3, 0.1, 0.7, 6)
    assert np.array_equal(a.pin_node, b.pin_node) and np.array_equal(a.x, b.x)
    assert not np.array_equal(a.pin_node, c.pin_node)


def test_stress_tail_exponent():
    """2% of all nets above 32 pins for the 2M-cell workload."""
    p = W.WORKLOADS["stress"]
    from erplacer.netlist import tail_alpha_for
    alpha = tail_alpha_for(p["frac_above_32"], 32, 4, p["cap"], 0.15)
    s = np.arange(4, p["cap"] + 1, dtype=float)
    share = (s[s > 32] ** -alpha).sum() / (s ** -alpha).sum()
    assert abs(0.15 * share - 0.02) < 1e-6


def test_serving_grid_rule():
    assert W.serving_grid(2_000) == 64
    assert W.serving_grid(150_000) == 256
    assert W.serving_grid(10 ** 7) == 512


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = L.declared_functions()
    assert len(names) >= 30
    for nm in names:
        assert hasattr(lib, nm), nm
    cfg = L.default_config()
    assert cfg.num_bins_x == 512 and cfg.gamma_scale == 4.0 and math.isnan(cfg.target_density)
    assert abs(cfg.density_weight0 - 8e-5) < 1e-20


def test_rust_binding_declares_every_entry_point():
    """The Rust extern blocks a maintainer copies declare every C-ABI function."""
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    with open(os.path.join(root, "INTEGRATION.md")) as f:
        text = f.read()
    bound = set(re.findall(r"pub fn (erp_[a-z0-9_]+)\s*\(", text))
    missing = [nm for nm in L.declared_functions() if nm not in bound]
    assert not missing, missing


def test_invalid_design_rejected_before_device_use():
    """Argument checks run before any HIP call (the reference panics instead)."""
    lib = L.load()
    h = ctypes.c_void_p()
    assert lib.erp_ctx_create(None, None, ctypes.byref(h)) == L.ERP_ERR_INVALID
    dd = generate(200, 210, 8, 1, 0.1, 0.7, 1)
    ds = L.design_struct(dd)
    bad = L.default_config(num_bins_x=100)
    assert lib.erp_ctx_create(ctypes.byref(ds), ctypes.byref(bad), ctypes.byref(h)) == L.ERP_ERR_UNSUPPORTED
    assert b"power of two" in lib.erp_last_error()


def test_server_arguments_checked_before_device_use():
    """The serving entry points reject bad configurations and estimate memory
    without touching a device."""
    from erplacer import server as S
    lib = L.load()
    cfg = L.ServerConfig()
    lib.erp_server_config_default(ctypes.byref(cfg))
    assert cfg.max_active > 0 and cfg.check_every > 0 and cfg.hbm_budget_bytes > 0
    h = ctypes.c_void_p()
    bad = L.ServerConfig(hbm_budget_bytes=1 << 30, max_active=0, check_every=10, device=0)
    assert lib.erp_server_create(ctypes.byref(bad), ctypes.byref(h)) == L.ERP_ERR_INVALID
    dd = generate(2000, 2100, 16, 4, 0.1, 0.7, 3)
    small = S.estimate_device_bytes(dd, 64)
    big = S.estimate_device_bytes(dd, 512)
    assert 0 < small < big
    # one int64 + eight float maps per bin at least
    assert big - small >= (512 * 512 - 64 * 64) * (8 + 8 * 4)
