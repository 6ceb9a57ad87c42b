"""Context API features the other GPU tests do not reach: the WA gradient at an
explicit gamma (compute_hpwl with a given gamma, hpwl.rs:139-343) against the
FP64 oracle, the per-kernel timing mode, and the fused loop without stored
per-cell gradients (both must leave the trajectory unchanged, bit for bit).
"""
import numpy as np
import pytest

import oracle_lib as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu

# the per-kernel timers of the FFT form's iteration (WA runs inside pass 1's launch,
# the field gather in the sorted order before the cell kernel; the tile scan runs
# inside the sort's placement pass)
FUSED_KERNELS = {"sort_keys", "sort_place", "scatter", "pass1_wa",
                 "pass2_col_spectral", "pass3_row_final", "field_sorted", "cell_fused",
                 "scalar", "update"}


@pytest.fixture(scope="module")
def design():
    from erplacer.netlist import generate
    # about one cell per bin at 64 x 64: the cell-major scatter
    return generate(4000, 4200, 32, 8, 0.10, 0.7, 21), 64


def _spread_positions(d, seed):
    """Movable cells at random legal corners, fixed ones at their coordinates."""
    rng = np.random.default_rng(seed)
    C = d.num_movable
    x = d.x.astype(np.float32).copy()
    y = d.y.astype(np.float32).copy()
    x[:C] = rng.uniform(0, d.die_w - d.w[:C] - 1).astype(np.float32)
    y[:C] = rng.uniform(0, d.die_h - d.h[:C] - 1).astype(np.float32)
    return x, y


def test_wa_gradient_at_explicit_gamma(gpu_lib, design):
    from erplacer import lib as L
    from erplacer.placement import Placement
    d, g = design
    x, y = _spread_positions(d, 3)
    ds = O.make_design_struct(d)
    with Placement(d, g, target_density=1.0) as P:
        P.set_positions(x, y)
        gamma0 = P.info()["gamma"]
        for gamma in (0.5 * gamma0, 3.0 * gamma0):
            h = P.compute_hpwl(gamma)
            gx, gy = P.cell_grads(L.GRAD_WIRELENGTH)
            h_ref, rgx, rgy, _, _ = O.compute_hpwl(ds, d, x.astype(np.float64), y.astype(np.float64), gamma)
            assert h == pytest.approx(h_ref, rel=1e-5)
            assert rel_l2(gx, rgx) < 1e-5 and rel_l2(gy, rgy) < 1e-5, gamma
            assert P.info()["gamma"] == gamma0          # the scheduler's gamma is untouched


def test_timing_mode_reports_kernels_and_keeps_trajectory(gpu_lib, design):
    from erplacer.placement import Placement
    d, g = design
    with Placement(d, g, target_density=1.0) as A, Placement(d, g, target_density=1.0) as B:
        B.set_timing(True)
        ia, ib = A.run(12), B.run(12)
        t = B.timing()
        assert FUSED_KERNELS <= set(t), set(t)
        assert all(v > 0.0 for v in t.values())
        B.set_timing(False)
        assert ia["hpwl"] == ib["hpwl"] and ia["overflow"] == ib["overflow"]
        xa, ya = A.positions()
        xb, yb = B.positions()
        assert np.array_equal(xa, xb) and np.array_equal(ya, yb)


def test_gradients_not_stored_keeps_trajectory(gpu_lib, design):
    from erplacer.placement import Placement
    d, g = design
    with Placement(d, g, target_density=1.0) as A, Placement(d, g, target_density=1.0) as B:
        B.set_store_gradients(False)
        ia, ib = A.run(30), B.run(30)
        assert ia["hpwl"] == ib["hpwl"] and ia["overflow"] == ib["overflow"]
        assert ia["density_weight"] == ib["density_weight"] and ia["gamma"] == ib["gamma"]
        xa, ya = A.positions()
        xb, yb = B.positions()
        assert np.array_equal(xa, xb) and np.array_equal(ya, yb)
