"""Pins the oracle's WA wirelength gradient and HPWL (hpwl.rs:139-343):
finite differences of the WA value, translation invariance per net, the
two-pin closed form, and the exact HPWL."""
import math

import numpy as np
import pytest

import oracle_lib as O


@pytest.fixture(scope="module")
def design():
    from erplacer.netlist import generate
    d = generate(600, 640, 24, 3, 0.1, 0.7, 77)
    rng = np.random.default_rng(1)
    x = d.x.astype(np.float64).copy()
    y = d.y.astype(np.float64).copy()
    C = d.num_movable
    x[:C] = rng.uniform(0, d.die_w * 0.9, C)
    y[:C] = rng.uniform(0, d.die_h * 0.9, C)
    return d, O.make_design_struct(d), x, y


def test_hpwl_exact(design):
    d, ds, x, y = design
    h, *_ = O.compute_hpwl(ds, d, x, y, gamma=5000.0)
    px = x[d.pin_node] + d.pin_offset_x
    py = y[d.pin_node] + d.pin_offset_y
    ref = 0.0
    for e in range(d.num_nets):
        p = d.net_pins[d.net_start[e]:d.net_start[e + 1]]
        ref += px[p].max() - px[p].min() + py[p].max() - py[p].min()
    assert math.isclose(h, ref, rel_tol=1e-12)


def test_gradient_is_derivative_of_wa(design):
    """Central differences of sum_nets (WA+ - WA-) per movable cell."""
    d, ds, x, y = design
    gamma = 3000.0
    _, gx, gy, _, _ = O.compute_hpwl(ds, d, x, y, gamma)
    rng = np.random.default_rng(2)
    for c in rng.choice(d.num_movable, 25, replace=False):
        for dim, g in ((0, gx), (1, gy)):
            step = 0.5
            a = [x.copy(), y.copy()]
            b = [x.copy(), y.copy()]
            a[dim][c] += step
            b[dim][c] -= step
            fd = (O.wa_value(ds, *a, gamma) - O.wa_value(ds, *b, gamma)) / (2 * step)
            assert abs(fd - g[c]) <= 1e-5 + 1e-4 * abs(g[c]), (c, dim, fd, g[c])


def test_translation_invariance_per_net(design):
    d, ds, x, y = design
    _, _, _, pgx, pgy = O.compute_hpwl(ds, d, x, y, gamma=4000.0)
    for e in range(d.num_nets):
        p = d.net_pins[d.net_start[e]:d.net_start[e + 1]]
        assert abs(pgx[p].sum()) < 1e-9 and abs(pgy[p].sum()) < 1e-9


def test_two_pin_closed_form():
    """Pins d apart: WA = d tanh(d/2g), so the far pin's gradient is
    tanh(d/2g) + d/(2g) sech^2(d/2g) and the near pin's is its negative."""
    from erplacer.netlist import Design
    dd = 1000.0
    gamma = 800.0
    d = Design(name="two", seed=0, die_w=1e5, die_h=1e5, num_movable=2,
               x=np.array([0, 0], np.float32), y=np.array([0, 0], np.float32),
               w=np.array([10, 10], np.float32), h=np.array([10, 10], np.float32),
               num_pins=np.array([1, 1], np.float32),
               pin_offset_x=np.zeros(2, np.float32), pin_offset_y=np.zeros(2, np.float32),
               pin_node=np.array([0, 1], np.int32), net_start=np.array([0, 2], np.int32),
               net_pins=np.array([0, 1], np.int32), n_2nets=1,
               node_pin_start=np.array([0, 1, 2], np.int32), node_pins=np.array([0, 1], np.int32),
               movable_area=200.0, cell_utilization=0.01, params={})
    ds = O.make_design_struct(d)
    x = np.array([0.0, dd])
    y = np.array([5.0, 5.0])
    h, gx, gy, _, _ = O.compute_hpwl(ds, d, x, y, gamma)
    t = dd / (2 * gamma)
    ref = math.tanh(t) + t / math.cosh(t) ** 2
    assert math.isclose(h, dd, rel_tol=1e-15)
    assert math.isclose(gx[1], ref, rel_tol=1e-12) and math.isclose(gx[0], -ref, rel_tol=1e-12)
    assert gy[0] == 0 and gy[1] == 0
    assert math.isclose(O.wa_value(ds, x, y, gamma), dd * math.tanh(t), rel_tol=1e-12)
