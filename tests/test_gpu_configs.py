"""The benchmark configurations at their full sizes, stage by stage against the
FP64 oracle (placement.rs:180-228; stages epotential.rs:225-374, density.rs,
eforce.rs:15-91, hpwl.rs:139-343).

Each stage is checked on identical inputs, the GPU's own f32 positions, so the
figure measures the kernel and not the f32 coordinates:

* the movable density map, bit for bit against the CPU restatement of the
  fixed-point accumulation, with the scatter kernel the loop uses (owner or
  cell-major) and, on the same positions, the other one;
* the field and potential maps against FP64 transforms (dct2d.rs, idxst_idct.rs,
  idct_idxst.rs, idct2.rs) of the GPU's density map;
* the density gradient against the FP64 gather (eforce.rs) of those FP64 fields
  at the GPU's positions;
* the WA gradient and HPWL against the FP64 compute_hpwl at the GPU's positions
  and gamma.

Both at the centre start (iteration 0) and at the spread state the loop reaches
after the configuration's iteration count.  Step 0 is also compared end to end
with the FP64 oracle's own step 0 (its own f64 positions), where the f32
coordinates bound the agreement (tol_overlap below).
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL_SPECTRAL = 1e-5     # field / potential maps from identical density input


def tol_overlap(m):
    """f32 coordinates resolve an overlap of one bin to ulp(die)/bin ~ m * 6e-8."""
    return 1e-7 * m


def _fp64_fields(rho, m, n, bw, bh):
    a = O.dct2_fft2_forward(rho.astype(np.float64))
    wu, wv, inv = O.frequency_matrices(m, n, float(bw), float(bh))
    fx = O.idxst_idct_forward(a * wu)
    fy = O.idct_idxst_forward(a * wv)
    pot = O.idct2_fft2_forward(a * inv)
    return fx, fy, pot


def _stages(P, d, label, other_scatter=True):
    """Stage-by-stage parity at the context's current positions (see module doc)."""
    from erplacer import lib as L
    m, n = P.m, P.n
    C = d.num_movable
    P.set_keep_density(True)
    x, y = P.positions()
    gamma = P.info()["gamma"]
    h = P.compute_hpwl()
    ds = O.make_design_struct(d)
    rh, rgx, rgy, _, _ = O.compute_hpwl(ds, d, x.astype(np.float64), y.astype(np.float64), gamma)
    gx, gy = P.cell_grads(L.GRAD_WIRELENGTH)
    e_wa = max(rel_l2(gx, rgx), rel_l2(gy, rgy))
    e_h = abs(h / rh - 1)

    P.compute_potential()
    st = P.setup()
    bw, bh = np.float32(st["bw"]), np.float32(st["bh"])
    q_mov, _, scale = P.density_fixed()
    args = (x[:C], y[:C], st["cw"][:C], st["ch"][:C], st["ox"][:C], st["oy"][:C], st["ratio"][:C])
    q_ref = O.density_map_fx64(*args, m, n, bw, bh, scale, sfx="_f32")
    kind = P.scatter_kind()
    assert np.array_equal(q_mov, q_ref), f"{label}: scatter kind {kind}"
    if other_scatter and P.scatter_kind() == L.SCATTER_OWNER:
        P.set_scatter_kind(L.SCATTER_CELL_MAJOR)
        P.compute_potential()
        q2, _, _ = P.density_fixed()
        assert np.array_equal(q2, q_ref), f"{label}: cell-major scatter"
        P.set_scatter_kind(L.SCATTER_AUTO)
        P.compute_potential()

    rho = P.map(L.MAP_DENSITY)
    fx, fy, pot = _fp64_fields(rho, m, n, bw, bh)
    e_fx = rel_l2(P.map(L.MAP_FIELD_X), fx)
    e_fy = rel_l2(P.map(L.MAP_FIELD_Y), fy)
    e_pot = rel_l2(P.map(L.MAP_POTENTIAL), pot)
    dgx, dgy = P.cell_grads(L.GRAD_DENSITY)
    rdx, rdy = O.electric_force(fx, fy, *args, float(bw), float(bh))
    e_gd = max(rel_l2(dgx, rdx), rel_l2(dgy, rdy))
    # the gather alone: FP64 gather of the GPU's own f32 field
    gfx = P.map(L.MAP_FIELD_X).astype(np.float64)
    gfy = P.map(L.MAP_FIELD_Y).astype(np.float64)
    sdx, sdy = O.electric_force(gfx, gfy, *args, float(bw), float(bh))
    e_ga = max(rel_l2(dgx, sdx), rel_l2(dgy, sdy))
    # error relative to the sum of the terms' magnitudes (|overlap| |field|):
    # what f32 summation of terms that cancel can be held to
    adx, ady = O.electric_force(np.abs(fx), np.abs(fy), *args, float(bw), float(bh))
    e_gn = max(np.linalg.norm(dgx - rdx) / np.linalg.norm(adx),
               np.linalg.norm(dgy - rdy) / np.linalg.norm(ady))
    # how much the terms cancel: |gradient| / sum of |terms|
    cancel = max(np.linalg.norm(rdx) / np.linalg.norm(adx), np.linalg.norm(rdy) / np.linalg.norm(ady))
    # the gather against the f32 restatement (the reference binary's arithmetic)
    # on the GPU's f32 field
    f32x, f32y = O.electric_force(gfx.astype(np.float32), gfy.astype(np.float32), *args, bw, bh,
                                  sfx="_f32")
    e_f32 = max(rel_l2(dgx, f32x), rel_l2(dgy, f32y))
    e_f32n = max(np.linalg.norm(dgx - f32x) / np.linalg.norm(adx),
                 np.linalg.norm(dgy - f32y) / np.linalg.norm(ady))
    P.set_keep_density(False)
    r = dict(scatter=["cell-major", "owner"][kind], wa_grad=e_wa, hpwl=e_h, field_x=e_fx,
             field_y=e_fy, potential=e_pot, density_grad=e_gd, gather_only=e_ga,
             density_grad_normwise=float(e_gn), cancellation=float(cancel),
             gather_vs_f32=e_f32, gather_vs_f32_normwise=float(e_f32n))
    print(label, {k: (f"{v:.2e}" if isinstance(v, float) else v) for k, v in r.items()})
    return r


def _check(r, mfma):
    assert r["hpwl"] < 1e-6
    assert r["wa_grad"] < 1e-5
    assert r["field_x"] < TOL_SPECTRAL and r["field_y"] < TOL_SPECTRAL
    # The potential is discarded by Placement::step (placement.rs:185).  It
    # weights the spectrum by 1/w^2, so its error is that of the lowest
    # frequencies, which at a spread state are small beside the terms of their
    # sums; in the split-bf16 products (FP32 accumulation over K = L/2) that
    # error depends on the state reached: 5.8e-6 at the centre start, 2.1e-5
    # after 800 iterations at 1024^2.
    assert r["potential"] < (3 * TOL_SPECTRAL if mfma else TOL_SPECTRAL)
    # The gather reproduces the reference binary's f32 arithmetic: against the
    # f32 restatement on the GPU's own field the error is summation order only.
    # Near the centre start the terms of a cell's sum cancel to 1e-4..1e-5 of
    # their magnitudes, so the error is held relative to that magnitude there.
    assert r["gather_vs_f32_normwise"] < 1e-6
    if r["cancellation"] > 0.1:
        assert r["gather_vs_f32"] < 1e-6
        # against FP64 arithmetic on the same f32 inputs the f32 overlaps and
        # truncated bin ranges of eforce.rs:40-90 (unclamped, so a range one bin
        # wider adds a negative overlap) measure 4.9e-6 at 512^2 and up to
        # 2.3e-4 at 1024^2 and 2048^2
        assert r["density_grad"] < 5e-4
    assert r["density_grad_normwise"] < 5e-4


def _step0_end_to_end(P, d, g, td, label):
    """Step 0 against the FP64 oracle's own step 0 (f64 positions, float map)."""
    from erplacer import lib as L
    ref = O.Placer(d, g, target_density=td)
    P.reset()
    h = P.compute_hpwl()
    _, overflow = P.compute_potential()
    r = ref.step()
    e = {"hpwl": abs(h / r["hpwl"] - 1), "overflow": abs(overflow - r["overflow"])}
    for nm, a, b in (("field_x", L.MAP_FIELD_X, O.MAP_FIELD_X),
                     ("field_y", L.MAP_FIELD_Y, O.MAP_FIELD_Y)):
        e[nm] = rel_l2(P.map(a), ref.map(b))
    gx, gy = P.cell_grads(L.GRAD_WIRELENGTH)
    rx, ry = ref.cell_grads(O.GRAD_WL)
    e["wa_grad"] = max(rel_l2(gx, rx), rel_l2(gy, ry))
    gx, gy = P.cell_grads(L.GRAD_DENSITY)
    rx, ry = ref.cell_grads(O.GRAD_DENSITY)
    e["density_grad"] = max(rel_l2(gx, rx), rel_l2(gy, ry))
    # normwise: against the FP64 gather of |field| over the oracle's own geometry
    C = d.num_movable
    st = ref.setup()
    ox, oy = ref.positions()
    afx, afy = np.abs(ref.map(O.MAP_FIELD_X)), np.abs(ref.map(O.MAP_FIELD_Y))
    ax, ay = O.electric_force(afx, afy, ox[:C], oy[:C], st["cw"][:C], st["ch"][:C], st["ox"][:C],
                              st["oy"][:C], st["ratio"][:C], st["bw"], st["bh"])
    e["density_grad_normwise"] = max(np.linalg.norm(gx - rx) / np.linalg.norm(ax),
                                     np.linalg.norm(gy - ry) / np.linalg.norm(ay))
    print(label, "step 0 vs FP64 oracle step 0:", {k: f"{v:.2e}" for k, v in e.items()})
    m = g if isinstance(g, int) else g[0]
    assert e["hpwl"] < 1e-5 and e["wa_grad"] < 1e-5
    assert e["overflow"] < tol_overlap(m)
    assert e["field_x"] < tol_overlap(m) and e["field_y"] < tol_overlap(m)
    # every cell sits at the centre, where its terms cancel (rel-L2 is noise)
    assert e["density_grad_normwise"] < 1e-6
    del ref


@pytest.mark.parametrize("name,mode", [("medium", "fft"), ("large", "fft"), ("large", "mfma"),
                                       ("stress", "fft"), ("stress", "mfma")])
def test_config_stages(gpu_lib, name, mode):
    """cfg2 (200k cells, 512^2), cfg3 (1M, 1024^2), cfg5 (2M, 2048^2): every stage
    at step 0 and after the configuration's full iteration count."""
    from erplacer import lib as L
    from erplacer import workloads as W
    from erplacer.placement import Placement
    d, g, iters, td = W.make(name)
    sm = L.SPECTRAL_FFT if mode == "fft" else L.SPECTRAL_MFMA_BF16X3
    print(f"config {name} {mode}: {d.num_movable} cells, {g}^2, {iters} iterations", flush=True)
    with Placement(d, g, target_density=td, spectral_mode=sm) as P:
        _check(_stages(P, d, f"{name}/{mode} iteration 0"), mode == "mfma")
        info = P.run(iters)
        assert not info["nonfinite"] and info["iteration"] == iters
        _check(_stages(P, d, f"{name}/{mode} iteration {iters}"), mode == "mfma")
        _step0_end_to_end(P, d, g, td, f"{name}/{mode}")


def test_medium_spread_states_bit_exact(gpu_lib):
    """cfg2: the sorted cell-major scatter, bit for bit, at iterations 50, 200
    and 300 (the sort starts from the previous iteration's order each time)."""
    from erplacer import lib as L
    from erplacer import workloads as W
    from erplacer.placement import Placement
    d, g, _, td = W.make("medium")
    C = d.num_movable
    with Placement(d, g, target_density=td) as P:
        assert P.scatter_kind() == L.SCATTER_CELL_MAJOR     # 0.76 cells per bin
        done = 0
        for it in (50, 200, 300):
            P.run(it - done)
            done = it
            P.set_keep_density(True)
            x, y = P.positions()
            P.compute_potential()
            q, _, scale = P.density_fixed()
            st = P.setup()
            q_ref = O.density_map_fx64(x[:C], y[:C], st["cw"][:C], st["ch"][:C], st["ox"][:C],
                                       st["oy"][:C], st["ratio"][:C], g, g, st["bw"], st["bh"],
                                       scale, sfx="_f32")
            assert np.array_equal(q, q_ref), it
            P.set_keep_density(False)


@pytest.mark.parametrize("name,scale", [("small", 1.0), ("medium", 0.63)])
def test_owner_scatter_bit_exact(gpu_lib, name, scale):
    """The owner scatter (chosen for fewer than 0.75 cells per bin): the small
    workload (0.61 cells per bin at 128^2; at the centre start one tile
    neighbourhood holds all 10k cells, so its work is split in chunks) and a
    0.48-cells-per-bin case at 512^2; at the start and at spread states."""
    from erplacer import lib as L
    from erplacer import workloads as W
    from erplacer.placement import Placement
    d, g, iters, td = W.make(name, scale=scale)
    C = d.num_movable
    assert C < 0.75 * g * g
    with Placement(d, g, target_density=td) as P:
        assert P.scatter_kind() == L.SCATTER_OWNER
        done = 0
        for it in (0, 1, 50, 300):
            if it > done:
                P.run(it - done)
                done = it
            P.set_keep_density(True)
            x, y = P.positions()
            P.compute_potential()
            q, _, sc = P.density_fixed()
            st = P.setup()
            q_ref = O.density_map_fx64(x[:C], y[:C], st["cw"][:C], st["ch"][:C], st["ox"][:C],
                                       st["oy"][:C], st["ratio"][:C], g, g, st["bw"], st["bh"],
                                       sc, sfx="_f32")
            assert np.array_equal(q, q_ref), (name, it)
            P.set_keep_density(False)


def test_keep_density_toggle_leaves_trajectory(gpu_lib):
    """Keep mode on for one step and off again: the next graph replay starts
    from cleared accumulators, so the trajectory equals the one without the
    toggle bit for bit."""
    from erplacer import workloads as W
    from erplacer.placement import Placement
    d, g, _, td = W.make("small")
    with Placement(d, g, target_density=td) as A, Placement(d, g, target_density=td) as B:
        A.run(5)
        B.run(5)
        A.set_keep_density(True)
        A.step()
        A.set_keep_density(False)
        B.step()
        ia, ib = A.run(5), B.run(5)
        xa, ya = A.positions()
        xb, yb = B.positions()
        assert np.array_equal(xa, xb) and np.array_equal(ya, yb)
        assert ia["hpwl"] == ib["hpwl"] and ia["overflow"] == ib["overflow"]


def test_scatter_kind_rejected_where_cells_reach_far(gpu_lib):
    """Forcing the owner scatter where a movable cell spans more than two tiles
    past its own is refused; the cell-major scatter is always accepted, and
    AUTO returns to the choice made at creation."""
    from erplacer import lib as L
    from erplacer.netlist import generate
    from erplacer.placement import Placement
    d = generate(3000, 3150, 16, 4, 0.1, 0.7, 77)
    d.w[0] = np.float32(0.4 * d.die_w)        # one movable cell 200 bins wide at 512^2
    with Placement(d, 512, target_density=1.0) as P:
        assert P.scatter_kind() == L.SCATTER_CELL_MAJOR
        with pytest.raises(L.ErpError) as e:
            P.set_scatter_kind(L.SCATTER_OWNER)
        assert e.value.code == L.ERP_ERR_UNSUPPORTED
        with pytest.raises(L.ErpError):
            P.set_scatter_kind(7)
        P.set_scatter_kind(L.SCATTER_AUTO)
        assert P.scatter_kind() == L.SCATTER_CELL_MAJOR
    assert math.isfinite(d.die_w)
