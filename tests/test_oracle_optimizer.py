"""Pins the oracle's Nesterov step (nesterov.rs:27-158), inline preconditioner
(nesterov.rs:75-82) and Scheduler (scheduler.rs:6-130) with hand-computed
values, and the placement loop's invariants (placement.rs:109-139, 180-228)."""
import math

import numpy as np
import pytest

import oracle_lib as O

PHI = (1 + math.sqrt(5)) / 2


def test_scheduler_table():
    """Hand-derived values: the first update leaves the weight unchanged (mu = 1.05
    * 1.05^-1), overflow 1.0 / 0.55 / 0.1 give gamma coefficients 10 / 1 / 0.1,
    overflow 0 gives 10^(-11/9); a 2.375x HPWL drop trips should_stop."""
    hpwl = [100.0, 90.0, 95.0, 40.0]
    ovf = [1.0, 0.55, 0.1, 0.0]
    dw, gm, stop = O.scheduler_sequence(hpwl, ovf, base_gamma=10.0, dw0=1.0)
    assert math.isclose(dw[0], 1.0, rel_tol=1e-15)
    assert math.isclose(gm[0], 100.0, rel_tol=1e-12)
    assert math.isclose(dw[1], 1.05 * 0.9999 ** 2, rel_tol=1e-12)       # HPWL fell
    assert math.isclose(gm[1], 10.0, rel_tol=1e-12)
    assert math.isclose(dw[2], dw[1] * 1.05 * 1.05 ** (-5 / 100), rel_tol=1e-12)
    assert math.isclose(gm[2], 1.0, rel_tol=1e-12)
    assert math.isclose(gm[3], 10.0 * 10 ** (-11 / 9), rel_tol=1e-12)
    assert list(stop) == [0, 0, 0, 1]


def test_scheduler_clamps():
    dw, gm, _ = O.scheduler_sequence([100.0, 1e9, 1.0], [5.0, -5.0, 0.1], base_gamma=2.0, dw0=1.0)
    assert math.isclose(gm[0], 20.0, rel_tol=1e-15)      # coefficient capped at 10
    assert math.isclose(gm[1], 0.1, rel_tol=1e-15)       # floored at 0.05
    assert math.isclose(dw[1] / dw[0], 1.05 * 0.95, rel_tol=1e-12)   # power floored at 0.95
    # a long decreasing run floors the power at 0.98
    n = 400
    h = np.linspace(1e6, 1e5, n)
    dw, _, _ = O.scheduler_sequence(h, np.full(n, 0.5), base_gamma=1.0, dw0=1.0)
    assert math.isclose(dw[-1] / dw[-2], 1.05 * 0.98, rel_tol=1e-12)


@pytest.fixture(scope="module")
def small():
    from erplacer.netlist import generate
    return generate(1500, 1600, 16, 4, 0.1, 0.7, 9)


def test_initial_solution(small):
    """placement.rs:117-131: movable cells at die centre - size/2, fixed at DB."""
    d = small
    P = O.Placer(d, 64, target_density=1.0)
    x, y = P.positions()
    C = d.num_movable
    assert np.array_equal(x[:C], d.die_w / 2 - d.w[:C] / 2)
    assert np.array_equal(y[:C], d.die_h / 2 - d.h[:C] / 2)
    assert np.array_equal(x[C:], d.x[C:]) and np.array_equal(y[C:], d.y[C:])


def test_first_two_nesterov_steps(small):
    """Iteration 0: alpha = golden ratio, beta = 0, weight scaled by the L1 ratio,
    step x1 = clamp(x0 - phi * (g_wl + dw g_d) / min(pins + dw area, 1)).
    Iteration 1: beta = (phi - 1) / a1, a1 = (1 + sqrt(4 phi^2 + 1)) / 2."""
    d = small
    C = d.num_movable
    P = O.Placer(d, 64, target_density=1.0)
    x0, y0 = P.positions()
    r0 = P.step()
    gwx, gwy = P.cell_grads(O.GRAD_WL)
    gdx, gdy = P.cell_grads(O.GRAD_DENSITY)
    assert r0["beta"] == 0.0 and math.isclose(r0["alpha"], PHI, rel_tol=1e-15)
    ratio = (np.abs(gwx).sum() + np.abs(gwy).sum()) / (np.abs(gdx).sum() + np.abs(gdy).sum())
    dw = 8e-5 * ratio
    assert math.isclose(r0["density_weight"], dw, rel_tol=1e-12)   # scheduler mu = 1 at step 1
    area = d.w[:C].astype(np.float64) * d.h[:C]
    inv_p = 1.0 / np.minimum(d.num_pins[:C] + dw * area, 1.0)
    x1, y1 = P.positions()
    ex = np.clip(x0[:C] - PHI * (gwx + dw * gdx) * inv_p, 0, d.die_w - d.w[:C] - 1)
    ey = np.clip(y0[:C] - PHI * (gwy + dw * gdy) * inv_p, 0, d.die_h - d.h[:C] - 1)
    assert np.allclose(x1[:C], ex, rtol=0, atol=1e-9) and np.allclose(y1[:C], ey, rtol=0, atol=1e-9)
    r1 = P.step()
    a1 = (1 + math.sqrt(4 * PHI * PHI + 1)) / 2
    assert math.isclose(r1["beta"], (PHI - 1) / a1, rel_tol=1e-14)
    assert r1["alpha"] > 0 and r1["iteration"] == 2


def test_preconditioner_is_identity_for_connected_cells(small):
    """min(num_pins + dw*area, 1) = 1 whenever a cell has a pin (nesterov.rs:79)."""
    d = small
    assert (d.num_pins[:d.num_movable] >= 1).all()


def test_fixed_point_loop_tracks_float_loop(small):
    """The fixed-point density accumulation changes the FP64 loop by rounding only."""
    d = small
    a = O.Placer(d, 64, target_density=1.0)
    b = O.Placer(d, 64, target_density=1.0, fixed_point=True)
    for _ in range(15):
        ra, rb = a.step(), b.step()
        assert math.isclose(ra["hpwl"], rb["hpwl"], rel_tol=1e-9)
        assert abs(ra["overflow"] - rb["overflow"]) < 1e-9


def test_threads_do_not_change_first_iterations(small):
    d = small
    a = O.Placer(d, 64, target_density=1.0, nthreads=1)
    b = O.Placer(d, 64, target_density=1.0, nthreads=4)
    for _ in range(5):
        ra, rb = a.step(), b.step()
        assert math.isclose(ra["hpwl"], rb["hpwl"], rel_tol=1e-12)
