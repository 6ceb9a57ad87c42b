"""Edge cases of the placement loop on the GPU against the FP64 oracle: no
movable cells, no nets, a net far above the degree buckets (one wavefront
loops over its pins), a non-square bin grid, a single movable cell, and cells
pressed against the die boundary."""
import math

import numpy as np
import pytest

import oracle_lib as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _rand_cells(rng, nc, die):
    w = rng.integers(1, 9, nc).astype(np.float32) * 200
    h = np.full(nc, 2000, np.float32)
    x = rng.uniform(0, die - w - 1).astype(np.float32)
    y = rng.uniform(0, die - h - 1).astype(np.float32)
    return x, y, w, h


def _compare_steps(d, grid, steps, tol_h=1e-3, tol_o=1e-3, start_at_db=False):
    from erplacer.placement import Placement
    P = Placement(d, grid, target_density=1.0)
    ref = O.Placer(d, grid, target_density=1.0)
    if start_at_db:      # spread start instead of every cell at the die centre
        P.set_positions(d.x, d.y)
        ref.set_positions(d.x, d.y)
    for k in range(steps):
        a = P.step()
        b = ref.step()
        assert not a["nonfinite"]
        assert math.isclose(a["hpwl"], b["hpwl"], rel_tol=tol_h, abs_tol=1e-6), (k, a, b)
        assert abs(a["overflow"] - b["overflow"]) < tol_o, (k, a, b)
    return P, ref


def test_no_movable_cells(gpu_lib):
    """movable_area = 0: the reference's overflow is 0/0 (epotential.rs:271-277);
    the step runs, raises the non-finite flag like the oracle's NaN, and moves
    nothing."""
    from erplacer.netlist import from_nets
    from erplacer.placement import Placement
    die = 64 * 2000.0
    x = np.array([1000, 40000, 90000], np.float32)
    y = np.array([5000, 60000, 20000], np.float32)
    w = np.array([8000, 12000, 6000], np.float32)
    h = np.array([8000, 6000, 9000], np.float32)
    d = from_nets(x, y, w, h, 0, [[(0, 10, 10), (1, 20, 30)], [(0, 0, 0), (1, 5, 5), (2, 7, 7)]], die, die)
    P = Placement(d, 64, target_density=1.0)
    ref = O.Placer(d, 64, target_density=1.0)
    for _ in range(3):
        a = P.step()
        b = ref.step()
        assert a["hpwl"] == b["hpwl"]
        assert bool(a["nonfinite"]) == (not math.isfinite(b["overflow"]))
    px, py = P.positions()
    assert np.array_equal(px, x) and np.array_equal(py, y)


def test_no_nets(gpu_lib):
    """No nets: HPWL 0 and the L1-ratio density weight is NaN in the reference
    (nesterov.rs:47-55 with a zero wirelength gradient feeding
    scheduler.rs:102-113); the GPU flags it the same way and the first step's
    overflow matches."""
    from erplacer.netlist import from_nets
    from erplacer.placement import Placement
    rng = np.random.default_rng(2)
    die = 128 * 1500.0
    x, y, w, h = _rand_cells(rng, 3000, die)
    d = from_nets(x, y, w, h, 3000, [], die, die)
    P = Placement(d, 64, target_density=1.0)
    ref = O.Placer(d, 64, target_density=1.0)
    a = P.step()
    b = ref.step()
    assert a["hpwl"] == 0.0 and b["hpwl"] == 0.0
    assert abs(a["overflow"] - b["overflow"]) < 2e-5
    assert bool(a["nonfinite"]) == (not math.isfinite(b["density_weight"]))


def test_net_beyond_degree_buckets(gpu_lib):
    """One 1500-pin net (above the 512 bucket) plus ordinary nets: the step-0
    wirelength gradient matches the oracle."""
    from erplacer import lib as L
    from erplacer.netlist import from_nets
    from erplacer.placement import Placement
    rng = np.random.default_rng(3)
    die = 128 * 1500.0
    nc = 4000
    x, y, w, h = _rand_cells(rng, nc, die)
    # distinct pin offsets: at the all-at-the-centre start a net whose pins share
    # a coordinate has an exactly zero WA gradient along it, which the oracle's
    # (a + b x) e form only reproduces up to rounding noise
    nets = [[(int(c), float(rng.integers(0, 200)), float(rng.integers(0, 2000)))
             for c in rng.choice(nc, 1500, replace=False)]]
    for _ in range(3000):
        k = int(rng.integers(2, 6))
        nets.append([(int(c), float(rng.integers(0, 200)), float(rng.integers(0, 2000)))
                     for c in rng.choice(nc, k, replace=False)])
    d = from_nets(x, y, w, h, nc, nets, die, die)
    P = Placement(d, 64, target_density=1.0)
    ref = O.Placer(d, 64, target_density=1.0)
    h0 = P.compute_hpwl()
    r = ref.step()
    assert math.isclose(h0, r["hpwl"], rel_tol=1e-5)
    gx, gy = P.cell_grads(L.GRAD_WIRELENGTH)
    rx, ry = ref.cell_grads(O.GRAD_WL)
    assert rel_l2(gx, rx) < 1e-5 and rel_l2(gy, ry) < 1e-5
    # without macros the all-at-the-centre start has an exactly symmetric field
    # (zero density gradient, infinite L1 ratio in the reference too): step from
    # the spread DB positions instead
    _compare_steps(d, 64, 3, start_at_db=True)


def test_non_square_grid(gpu_lib):
    from erplacer.netlist import generate
    d = generate(3000, 3150, 16, 4, 0.1, 0.7, 9)
    from erplacer.placement import Placement
    P = Placement(d, (128, 64), target_density=1.0)
    ref = O.Placer(d, (128, 64), target_density=1.0)
    for k in range(5):
        a = P.step()
        b = ref.step()
        assert math.isclose(a["hpwl"], b["hpwl"], rel_tol=1e-3), k
        assert abs(a["overflow"] - b["overflow"]) < 1e-3, k


def test_single_movable_cell(gpu_lib):
    from erplacer.netlist import from_nets
    die = 64 * 1000.0
    x = np.array([0, 30000], np.float32)
    y = np.array([0, 30000], np.float32)
    w = np.array([800, 5000], np.float32)
    h = np.array([1000, 5000], np.float32)
    d = from_nets(x, y, w, h, 1, [[(0, 100, 100), (1, 2500, 2500)]], die, die)
    # one cell: the BB step grows to ~1e5 within a few steps and amplifies the
    # f32/f64 difference, so only the first steps are compared
    _compare_steps(d, 64, 3)


def test_cells_at_die_boundary(gpu_lib):
    """Cells placed at the clamp bounds (x = 0 and x = W - w - 1) after the
    first step are handled like the oracle (nesterov.rs:131-132)."""
    from erplacer.netlist import from_nets
    from erplacer.placement import Placement
    rng = np.random.default_rng(4)
    die = 64 * 1500.0
    x, y, w, h = _rand_cells(rng, 800, die)
    nets = [[(int(a), 0.0, 0.0), (int(b), 0.0, 0.0)] for a, b in rng.integers(0, 800, (900, 2)) if a != b]
    d = from_nets(x, y, w, h, 800, nets, die, die)
    P = Placement(d, 64, target_density=1.0)
    px, py = P.positions()
    px[:400] = 0.0
    px[400:] = (die - w[400:] - 1).astype(np.float32)
    P.set_positions(px, py)
    ref = O.Placer(d, 64, target_density=1.0)
    ref.set_positions(px, py)
    for k in range(3):
        a = P.step()
        b = ref.step()
        assert math.isclose(a["hpwl"], b["hpwl"], rel_tol=1e-3), k
    gx, _ = P.positions()
    assert gx.min() >= 0.0 and np.all(gx <= die - w - 1)
