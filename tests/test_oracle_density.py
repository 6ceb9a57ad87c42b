"""Pins the oracle's density scatter and field gather (density.rs, eforce.rs)
by brute force geometry, conservation, order independence and the adjoint
identity between scatter and gather.  Mirrors the reference's
test_initial_density_map / test_density_map / test_density_map_parallel
(density.rs:333-514) with seeded inputs in place of the missing fixtures.
"""
import numpy as np
import pytest

import oracle_lib as O


def _brute_scatter(x, y, cw, ch, ox, oy, ratio, m, n, bw, bh):
    """sum over cells of ratio * |[x,x+cw] & bin_x| * |[y,y+ch] & bin_y|, bins
    [0, m-1) x [0, n-1) (the last row/column never receive movable density)."""
    out = np.zeros((m, n))
    edges_x = np.arange(m + 1) * bw
    edges_y = np.arange(n + 1) * bh
    for i in range(x.shape[0]):
        xi, yi = float(x[i]) + float(ox[i]), float(y[i]) + float(oy[i])
        px = np.clip(np.minimum(xi + float(cw[i]), edges_x[1:]) - np.maximum(xi, edges_x[:-1]), 0, None)
        py = np.clip(np.minimum(yi + float(ch[i]), edges_y[1:]) - np.maximum(yi, edges_y[:-1]), 0, None)
        px[m - 1] = 0
        py[n - 1] = 0
        out += float(ratio[i]) * np.outer(px, py)
    return out


def _cells(rng, nc, die, bw, bh):
    w = rng.integers(1, 9, nc).astype(np.float32) * 200
    h = np.full(nc, 2000, np.float32)
    sq = np.float32(np.sqrt(np.float32(2.0)))
    cw = np.maximum(w, np.float32(bw) * sq).astype(np.float32)
    ch = np.maximum(h, np.float32(bh) * sq).astype(np.float32)
    ox = ((w - cw) * np.float32(0.5)).astype(np.float32)
    oy = ((h - ch) * np.float32(0.5)).astype(np.float32)
    ratio = ((w * h) / (cw * ch)).astype(np.float32)
    x = rng.uniform(0, die - w - 1).astype(np.float32)
    y = rng.uniform(0, die - h - 1).astype(np.float32)
    return x, y, cw, ch, ox, oy, ratio


@pytest.fixture(scope="module")
def case():
    rng = np.random.default_rng(1)
    m = n = 32
    die = 64_000.0
    bw = bh = die / m
    cells = _cells(rng, 300, die, bw, bh)
    return cells, m, n, bw, bh


def test_density_map_sequential_brute_force(case):
    cells, m, n, bw, bh = case
    got = O.density_map("sequential", *cells, m, n, bw, bh, np.zeros((m, n)))
    ref = _brute_scatter(*cells, m, n, bw, bh)
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


def test_density_map_parallel_matches_sequential(case):
    """density.rs:102-171 vs :174-237: same bins, only summation order differs."""
    cells, m, n, bw, bh = case
    init = np.random.default_rng(2).random((m, n)) * bw * bh
    a = O.density_map("sequential", *cells, m, n, bw, bh, init)
    b = O.density_map("parallel", *cells, m, n, bw, bh, init, nthreads=1)
    c = O.density_map("parallel", *cells, m, n, bw, bh, init, nthreads=4)
    assert np.array_equal(a, b)
    assert np.abs(a - c).max() <= 1e-9 * np.abs(a).max()


def test_density_mass_conservation(case):
    """Total movable mass = sum ratio * clipped area, and the clamp preserves the
    real cell area (epotential.rs:94): ratio*cw*ch = w*h."""
    cells, m, n, bw, bh = case
    x, y, cw, ch, ox, oy, ratio = [a.astype(np.float64) for a in cells]
    got = O.density_map("sequential", *cells, m, n, bw, bh, np.zeros((m, n)))
    lim_x, lim_y = (m - 1) * bw, (n - 1) * bh
    ax = np.clip(np.minimum(x + ox + cw, lim_x) - np.maximum(x + ox, 0), 0, None)
    ay = np.clip(np.minimum(y + oy + ch, lim_y) - np.maximum(y + oy, 0), 0, None)
    assert np.isclose(got.sum(), float((ratio * ax * ay).sum()), rtol=1e-9)


def test_fixed_point_matches_float_and_is_order_free(case):
    cells, m, n, bw, bh = case
    x, y, cw, ch, ox, oy, ratio = cells
    scale = O.fx_scale(float((cw * ch * ratio).sum()) * 1.01)
    q = O.density_map_fx64(*cells, m, n, bw, bh, scale, sfx="_f64")
    ref = O.density_map("sequential", *cells, m, n, bw, bh, np.zeros((m, n)))
    # each contribution is rounded once: at most 9 per cell-bin pair... bound by count/2
    assert np.abs(q / scale - ref).max() <= x.shape[0] * 16 * 0.5 / scale
    perm = np.random.default_rng(3).permutation(x.shape[0])
    q2 = O.density_map_fx64(*[a[perm] for a in cells], m, n, bw, bh, scale, sfx="_f64")
    assert np.array_equal(q, q2)


def test_fx_scale_is_largest_safe_power_of_two():
    for bound in (1.0, 3.7e6, 1.8e12, 5.5e13):
        s = O.fx_scale(bound)
        assert s * bound <= 2.0 ** 62 < 2 * s * bound
        assert np.log2(s) == int(np.log2(s))


def test_initial_density_map_brute_force():
    """density.rs:26-99: overlap with [c - b/2, c + b/2] of each bin centre, bins
    [floor, ceil + 1) capped at M-1; per-thread maps merged in order."""
    rng = np.random.default_rng(4)
    m = n = 32
    die = 64_000.0
    bw = bh = die / m
    nf = 12
    x = rng.uniform(0, die * 0.7, nf)
    y = rng.uniform(0, die * 0.7, nf)
    cw = rng.uniform(1000, 15000, nf)
    ch = rng.uniform(1000, 15000, nf)
    got = O.initial_density_map(x, y, cw, ch, m, n, bw, bh, die, die)
    ref = _brute_scatter(x, y, cw, ch, np.zeros(nf), np.zeros(nf), np.ones(nf), m, n, bw, bh)
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()
    got4 = O.initial_density_map(x, y, cw, ch, m, n, bw, bh, die, die, nthreads=4)
    assert np.abs(got4 - got).max() <= 1e-9 * np.abs(ref).max()


def test_electric_force_brute_force_and_adjoint(case):
    """eforce.rs:15-91: g_i = -ratio_i sum_b ox oy field_b.  For a cell away from
    the last row/column this equals minus the inner product of its own scatter
    footprint (density.rs:220-227) with the field."""
    cells, m, n, bw, bh = case
    x, y, cw, ch, ox, oy, ratio = cells
    rng = np.random.default_rng(5)
    fx = rng.standard_normal((m, n))
    fy = rng.standard_normal((m, n))
    gx, gy = O.electric_force(fx, fy, *cells, bw, bh)
    checked = 0
    for i in range(x.shape[0]):
        xi, yi = float(x[i] + ox[i]), float(y[i] + oy[i])
        if xi + cw[i] >= (m - 1) * bw or yi + ch[i] >= (n - 1) * bh or xi < 0 or yi < 0:
            continue
        one = [a[i:i + 1] for a in cells]
        foot = _brute_scatter(*one, m, n, bw, bh)
        assert np.isclose(gx[i], -(foot * fx).sum(), rtol=1e-9, atol=1e-9)
        assert np.isclose(gy[i], -(foot * fy).sum(), rtol=1e-9, atol=1e-9)
        checked += 1
    assert checked > 100
    # a plausible mistake (field transposed) is caught
    gxt, _ = O.electric_force(fx.T.copy(), fy, *cells, bw, bh)
    assert np.abs(gxt - gx).max() > 1e-3 * np.abs(gx).max()
