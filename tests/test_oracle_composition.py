"""Pins of the oracle's compositional parts: how PotentialComputation::new sets
up the clamped cells (epotential.rs:83-103), how compute_potential wires the
density map, overflow, operators, transforms and energy together
(epotential.rs:271-346), and the Barzilai-Borwein step size and momentum of
Nesterov::step_adaptive at iterations >= 1 (nesterov.rs:60-100,148-158).

Each check recomputes the quantity from the oracle's exported state by an
independent route (hand values, brute force, dense closed forms, numpy), so a
plausible slip in the oracle -- a swapped operator, an overflow on the scaled
map, the ratio of a fixed cell left at w h / (cw ch), the step size as
sqrt(den/num) -- fails one of them.
"""
import math

import numpy as np
import pytest
import scipy.fft

import oracle_lib as O
from erplacer.netlist import Design

F32_SQRT2 = 1.41421353816986083984375     # std::f32::consts::SQRT_2 (epotential.rs:18)


def _tiny_design():
    """Two movable cells and two fixed ones on a 6400 x 6400 die, one two-pin net."""
    f = lambda *v: np.array(v, np.float32)          # noqa: E731
    i = lambda *v: np.array(v, np.int32)            # noqa: E731
    return Design(name="tiny", seed=0, die_w=6400.0, die_h=6400.0, num_movable=2,
                  x=f(0, 0, 1000, 4000), y=f(0, 0, 2000, 500),
                  w=f(50, 200, 1000, 20), h=f(300, 80, 1000, 20), num_pins=f(1, 1, 0, 0),
                  pin_offset_x=f(10, 20), pin_offset_y=f(5, 40), pin_node=i(0, 1),
                  net_start=i(0, 2), net_pins=i(0, 1), n_2nets=1,
                  node_pin_start=i(0, 1, 2, 2, 2), node_pins=i(0, 1),
                  movable_area=50.0 * 300 + 200 * 80, cell_utilization=0.8, params={})


def test_setup_hand_values():
    """epotential.rs:83-103 at 64 x 64 bins of 100 x 100 (sqrt(2) bin = 141.42135...
    with the f32 constant): hand-computed clamped sizes, offsets and ratios;
    fixed cells take the target density as ratio whatever their size."""
    s = O.Placer(_tiny_design(), 64, target_density=0.8, sfx="_f64").setup()
    assert s["bw"] == 100.0 and s["bh"] == 100.0
    c = 141.421353816986083984375                  # 100 * F32_SQRT2, exact in binary64
    exp = {
        "cw": [c, 200.0, 1000.0, c],
        "ch": [300.0, c, 1000.0, c],
        "ox": [-45.7106769084930419921875, 0.0, 0.0, -60.7106769084930419921875],
        "oy": [0.0, -30.7106769084930419921875, 0.0, -60.7106769084930419921875],
        "ratio": [0.3535533966440824, 0.5656854346305319, 0.8, 0.8],  # 50 / c, 80 / c
    }
    for k, v in exp.items():
        assert np.allclose(s[k], v, rtol=1e-13, atol=0), (k, s[k])
    assert math.isclose(c, 100.0 * F32_SQRT2, rel_tol=0)


@pytest.fixture(scope="module")
def placed():
    """One FP64 step (exact twiddles, so the closed forms apply to 1e-12) on a
    non-square grid with non-square bins: 64 x 128 bins over a square die."""
    from erplacer.netlist import generate
    d = generate(3000, 3150, 16, 6, 0.1, 0.7, 21)
    P = O.Placer(d, (64, 128), target_density=0.9, pi_mode=O.PI_EXACT)
    x0, y0 = P.positions()
    r = P.step()
    return d, P, r, x0, y0


def test_overflow_on_combined_unscaled_map(placed):
    """epotential.rs:271-277: the map the overflow sums is the fixed map x target
    density plus the movable scatter, before the 1/bin_area scaling, and the sum
    of excess over target x bin area is divided by the movable area."""
    d, P, r, x0, y0 = placed
    s = P.setup()
    C = d.num_movable
    A = s["bw"] * s["bh"]
    mov = O.density_map("sequential", x0[:C], y0[:C], s["cw"][:C], s["ch"][:C], s["ox"][:C],
                        s["oy"][:C], s["ratio"][:C], 64, 128, s["bw"], s["bh"],
                        np.zeros((64, 128)))
    init = P.map(O.MAP_INITIAL)
    D = P.map(O.MAP_DENSITY) * A                     # the oracle keeps the scaled map
    assert np.allclose(D, init + mov, rtol=1e-12, atol=1e-9 * A)
    ovf = 0.0
    for v in (init + mov).reshape(-1):               # brute force, bin by bin
        ovf += max(v - 0.9 * A, 0.0)
    ovf /= d.movable_area
    assert math.isclose(r["overflow"], ovf, rel_tol=1e-12)
    # the fixed map is really there and really scaled by the target density
    assert init.sum() > 0.1 * mov.sum()


def _dense_fields(rho, bw, bh):
    M, N = rho.shape
    a = scipy.fft.dctn(rho, type=2) / (M * N)
    wu = 2 * np.pi * np.arange(M) / M
    wv = 2 * np.pi * np.arange(N) / N * (bw / bh)
    den = wu[:, None] ** 2 + wv[None, :] ** 2
    den[0, 0] = 1.0
    inv = 1.0 / den
    inv[0, 0] = 0.0
    Wx = wu[:, None] * inv / 2
    Wy = wv[None, :] * inv / 2
    epsM = np.ones(M)
    epsM[0] = 0.5
    epsN = np.ones(N)
    epsN[0] = 0.5
    iM, iN = np.arange(M), np.arange(N)
    S_M = np.sin(np.pi * np.outer(2 * iM + 1, iM) / (2 * M))
    C_M = np.cos(np.pi * np.outer(2 * iM + 1, iM) / (2 * M))
    S_N = np.sin(np.pi * np.outer(2 * iN + 1, iN) / (2 * N))
    C_N = np.cos(np.pi * np.outer(2 * iN + 1, iN) / (2 * N))
    xi_x = 4 * S_M @ ((a * Wx) * epsN[None, :]) @ C_N.T
    xi_y = 4 * C_M @ (epsM[:, None] * (a * Wy)) @ S_N.T
    phi = scipy.fft.dctn(a * inv, type=3)
    return a, xi_x, xi_y, phi


def test_spectral_wiring_and_energy(placed):
    """epotential.rs:283-346 as dense products on the oracle's own scaled map:
    a = dctn(rho)/MN, xi_x = IDXST_x IDCT_y (a Wx), xi_y = IDCT_x IDXST_y (a Wy),
    phi = IDCT2 (a inv), energy = sum phi rho.  64 x 128 bins with bw != bh, so
    swapping Wx / Wy or the two inverses changes every map."""
    d, P, r, _, _ = placed
    s = P.setup()
    assert s["bw"] != s["bh"]
    rho = P.map(O.MAP_DENSITY)
    a, xi_x, xi_y, phi = _dense_fields(rho, s["bw"], s["bh"])
    rel = lambda u, v: np.linalg.norm(u - v) / np.linalg.norm(v)      # noqa: E731
    assert rel(P.map(O.MAP_AUV), a) < 1e-12
    assert rel(P.map(O.MAP_FIELD_X), xi_x) < 1e-11
    assert rel(P.map(O.MAP_FIELD_Y), xi_y) < 1e-11
    assert rel(P.map(O.MAP_POTENTIAL), phi) < 1e-11
    # the swapped wiring is far off
    _, sx, sy, _ = _dense_fields(rho, s["bh"], s["bw"])
    assert rel(P.map(O.MAP_FIELD_X), sx) > 1e-3
    assert math.isclose(r["energy"], float((phi * rho).sum()), rel_tol=1e-11)
    assert not math.isclose(r["energy"], float((phi * rho).sum() * s["bw"] * s["bh"]), rel_tol=1e-3)


def _precond_grad(d, gwl, gd, lam):
    """nesterov.rs:75-82: (g_wl + lam g_d) / min(num_pins + lam area, 1)."""
    C = d.num_movable
    area = d.w[:C].astype(np.float64) * d.h[:C]
    inv_p = 1.0 / np.minimum(d.num_pins[:C].astype(np.float64) + lam * area, 1.0)
    return (gwl[0] + lam * gd[0]) * inv_p, (gwl[1] + lam * gd[1]) * inv_p


def test_bb_step_and_momentum_from_exported_state():
    """Iterations 1-3 recomputed from the exported positions and gradients:
    alpha_k = sqrt(sum |dx|^2 / sum |dg|^2) over the preconditioned combined
    gradients (nesterov.rs:60-100); beta_k = (alpha_{k-1} - 1) / a_k with
    a_k = (1 + sqrt(4 alpha_{k-1}^2 + 1)) / 2 (nesterov.rs:148-158, alpha_{k-1}
    the previous step size); x_{k+1} = clamp(x_k + beta_k (x_k - x_{k-1})
    - alpha_k g_k, 0, W - w - 1) (nesterov.rs:127-132)."""
    from erplacer.netlist import generate
    d = generate(2500, 2600, 16, 4, 0.1, 0.7, 33)
    C = d.num_movable
    P = O.Placer(d, 64, target_density=1.0)
    xs = [np.stack(P.positions())[:, :C].astype(np.float64)]
    outs, gws, gds = [], [], []
    for _ in range(4):
        outs.append(P.step())
        xs.append(np.stack(P.positions())[:, :C].astype(np.float64))
        gws.append(P.cell_grads(O.GRAD_WL))
        gds.append(P.cell_grads(O.GRAD_DENSITY))
    # density weight each step used: iteration 0 scales dw0 = 8e-5 by the L1
    # ratio (nesterov.rs:47-55); later steps use the scheduler's value after
    # the previous step
    l1 = (np.abs(gws[0][0]).sum() + np.abs(gws[0][1]).sum()) / \
         (np.abs(gds[0][0]).sum() + np.abs(gds[0][1]).sum())
    lams = [8e-5 * l1] + [o["density_weight"] for o in outs[:-1]]
    g = [np.stack(_precond_grad(d, gws[k], gds[k], lams[k])) for k in range(4)]
    W = np.array([d.die_w, d.die_h])[:, None]
    sz = np.stack([d.w[:C], d.h[:C]]).astype(np.float64)
    alpha_prev = (1 + math.sqrt(5)) / 2             # iteration 0: momentum from alpha = 1
    assert math.isclose(outs[0]["alpha"], alpha_prev, rel_tol=1e-15) and outs[0]["beta"] == 0.0
    for k in (1, 2, 3):
        num = ((xs[k] - xs[k - 1]) ** 2).sum()
        den = ((g[k] - g[k - 1]) ** 2).sum()
        alpha = math.sqrt(num / den)
        a_k = (1 + math.sqrt(4 * alpha_prev ** 2 + 1)) / 2
        beta = (alpha_prev - 1) / a_k
        assert math.isclose(outs[k]["alpha"], alpha, rel_tol=1e-9), k
        assert not math.isclose(outs[k]["alpha"], math.sqrt(den / num), rel_tol=1e-2)
        assert math.isclose(outs[k]["beta"], beta, rel_tol=1e-9, abs_tol=1e-15), k
        nxt = np.clip(xs[k] + beta * (xs[k] - xs[k - 1]) - alpha * g[k], 0.0, W - sz - 1)
        assert np.allclose(xs[k + 1], nxt, rtol=1e-9, atol=1e-6 * d.die_w), k
        alpha_prev = alpha
    # the conflation makes beta negative whenever the previous BB step was < 1
    assert any(o["alpha"] < 1 for o in outs[1:3]) == any(o["beta"] < 0 for o in outs[2:4])
