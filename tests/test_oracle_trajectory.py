"""The oracle's own whole-run spread on the 'small' workload (10k cells, 128x128,
300 iterations), which is what an end-state parity bound can be measured against.

The reference's Nesterov step keeps the Barzilai-Borwein step size in the field
it also uses as the momentum sequence (nesterov.rs:98-100,148-158): beta =
(alpha_prev - 1) / alpha_new.  The loop is therefore sensitive to its inputs.
tests/golden/small_band.json (written by tests/golden/make_small_band.py) holds
the final HPWL and overflow of 32 FP64 runs that differ by one FP64 ulp in one
movable coordinate (or only in how the density map is summed), and of 17 f32
runs (the reference binary's precision, main.rs:69).  These tests re-run part of
that ensemble, check it reproduces, and pin the two facts the GPU end-state
bound rests on:

* a one-ulp FP64 perturbation stays below 1e-9 relative for 100 iterations and
  then grows to a fraction of a percent, so the spread is the algorithm's, not
  an artefact of a large perturbation;
* in f32 coordinates the first two steps move cells by a few ulps only, so the
  step size of the third step (alpha = sqrt(sum dx^2 / sum dg^2), nesterov.rs:
  87-100) is known to ~1e-3: the f32 run leaves the FP64 run at step 3.
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

BAND = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "small_band.json")


@pytest.fixture(scope="module")
def band():
    with open(BAND) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def small():
    from erplacer import workloads as W
    return W.make("small")


def _trace(small, sfx="_f64", fixed_point=False, perturb=None, steps=None):
    d, g, iters, td = small
    P = O.Placer(d, g, target_density=td, sfx=sfx, fixed_point=fixed_point)
    out = []
    for k in range(iters if steps is None else steps):
        if perturb is not None and k == perturb[1]:
            x, y = P.positions()
            x[perturb[0]] = np.nextafter(x[perturb[0]], np.inf)
            P.set_positions(x, y)
        r = P.step()
        out.append((r["hpwl"], r["overflow"], r["alpha"]))
    return np.array(out), P


def test_band_fixture_reproduces(band, small):
    """Three runs of the committed ensemble, re-run: identical final scalars."""
    picks = [band["runs"][0], band["runs"][3],
             next(r for r in band["runs"] if r["sfx"] == "_f32" and r["cell"] is None)]
    for r in picks:
        pert = None if r["cell"] is None else (r["cell"], r["iteration"])
        t, _ = _trace(small, r["sfx"], r["fixed_point"], pert)
        assert t[-1, 0] == r["hpwl"] and t[-1, 1] == r["overflow"], r


def test_one_ulp_perturbation_grows_to_the_band(small):
    base, _ = _trace(small)
    pert, _ = _trace(small, perturb=(5000, 1))
    rel = np.abs(pert[:, 0] / base[:, 0] - 1)
    assert rel[:101].max() < 1e-9            # infinitesimal for 100 iterations
    assert rel[-1] > 1e-3                    # then a tenth of a percent or more
    assert abs(pert[-1, 1] - base[-1, 1]) > 1e-3


def test_fp64_band_is_wider_than_half_a_percent(band):
    """Final HPWL over the FP64 one-ulp ensemble: its standard deviation exceeds
    a quarter of a percent and its range a full percent, so 'within 0.5 % of the
    FP64 run' is inside the oracle's own noise on this workload."""
    s = band["fp64"]
    assert s["n"] == 32
    assert s["hpwl_std"] / s["hpwl_mean"] > 0.0025
    assert s["hpwl_max"] / s["hpwl_min"] - 1 > 0.01
    assert s["overflow_max"] - s["overflow_min"] > 0.005
    h = [r["hpwl"] for r in band["runs"] if r["sfx"] == "_f64"]
    # the unperturbed float-map run (the faithful oracle) is itself one draw
    assert min(h) <= h[0] <= max(h)
    f32 = band["f32_fixed_point"]
    # the f32 runs (reference-binary precision) fall in the same distribution
    assert abs(f32["hpwl_mean"] / s["hpwl_mean"] - 1) < 2 * s["hpwl_std"] / s["hpwl_mean"]


def test_f32_coordinates_part_at_step_three(small):
    """The f32 run agrees with the FP64 run to a few 1e-6 for steps 0-2 and
    leaves it at step 3, where the step size computed from two steps of a few
    f32 ulps each takes over."""
    d, g, _, td = small
    C = d.num_movable
    f64, _ = _trace(small, fixed_point=True, steps=4)
    P = O.Placer(d, g, target_density=td, sfx="_f32", fixed_point=True)
    f32, xs = [], []
    for _ in range(4):
        r = P.step()
        f32.append((r["hpwl"], r["overflow"], r["alpha"]))
        xs.append(P.positions()[0][:C].astype(np.float64))
    f32 = np.array(f32)
    rel = np.abs(f32[:, 0] / f64[:, 0] - 1)
    assert rel[:3].max() < 1e-5
    assert rel[3] > 1e-4
    # the step size after step 2 is off by far more than f32 rounding of scalars
    assert abs(f32[2, 2] / f64[2, 2] - 1) > 1e-4
    # steps 1 -> 2 move the cells by a few ulps of their f32 coordinates
    ulp = np.spacing(np.abs(xs[1]).astype(np.float32)).astype(np.float64)
    assert np.median(np.abs(xs[1] - xs[0]) / ulp) < 16
    # and step 3 by thousands
    assert np.median(np.abs(xs[3] - xs[2]) / ulp) > 1000


def test_large_band_is_wider_than_half_a_percent():
    """tests/golden/large_band.json (make_large_band.py: eight FP64 runs of the
    1M-cell, 1024^2, 800-iteration workload, one-ulp perturbations and float /
    fixed-point maps): the final HPWL spreads by several percent, so the metric's
    'within 0.5 % of the FP64 run' is inside the oracle's own noise there too."""
    with open(os.path.join(os.path.dirname(BAND), "large_band.json")) as f:
        b = json.load(f)
    s = b["fp64"]
    assert b["iterations"] == 800 and s["n"] == 8
    assert s["hpwl_std"] / s["hpwl_mean"] > 0.01
    assert s["hpwl_max"] / s["hpwl_min"] - 1 > 0.05
    assert s["overflow_max"] - s["overflow_min"] > 0.01
    runs = [r for r in b["runs"] if r["sfx"] == "_f64"]
    # every FP64 run agrees with the others to 1e-3 at iteration 50 (the runs
    # part later), so the spread is grown, not seeded
    h50 = np.array([r["trace"][0][1] for r in runs])
    assert np.all([r["trace"][0][0] == 50 for r in runs])
    assert h50.max() / h50.min() - 1 < 1e-3
