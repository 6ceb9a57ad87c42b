"""LEF/DEF ingestion into the placement data contract (the step upstream of
``Placement::new``, main.rs:68-70).  reda-db is not vendored with the
reference, so its conventions are parity unpinned; these tests pin the reader
against hand-computed values of a small LEF/DEF pair and the orientation table
against the LEF/DEF definitions of the eight orientations as matrices.
"""
import numpy as np
import pytest

from erplacer import lefdef as LD

LEF = """
VERSION 5.8 ;
# a comment ; with a semicolon
UNITS DATABASE MICRONS 2000 ; END UNITS
SITE core SIZE 0.2 BY 2.0 ; END core
LAYER M1 TYPE ROUTING ; END M1
MACRO INV
  CLASS CORE ;
  ORIGIN 0 0 ;
  FOREIGN INV 0 0 ;
  SIZE 0.6 BY 2.0 ;
  SITE core ;
  PIN A DIRECTION INPUT ; USE SIGNAL ;
    PORT LAYER M1 ; RECT 0.1 0.5 0.2 0.9 ; END
  END A
  PIN Z DIRECTION OUTPUT ;
    PORT LAYER M1 ; RECT 0.4 1.0 0.5 1.6 ; RECT MASK 1 0.45 0.2 0.5 0.4 ; END
  END Z
  OBS LAYER M1 ; RECT 0 0 0.6 0.1 ; END
END INV
MACRO BLK
  CLASS BLOCK ;
  SIZE 10 BY 8 ;
  PIN P PORT LAYER M2 ; POLYGON 1 1 2 1 2 3 1 3 ; END END P
END BLK
END LIBRARY
"""

DEF = """
VERSION 5.8 ;
DESIGN t ;
UNITS DISTANCE MICRONS 1000 ;
DIEAREA ( 1000 2000 ) ( 41000 32000 ) ;
ROW r0 core 1000 2000 N DO 200 BY 1 STEP 200 0 ;
COMPONENTS 5 ;
- u1 INV + PLACED ( 5000 6000 ) N ;
- u2 INV + PLACED ( 7000 6000 ) FS ;
- b1 BLK + FIXED ( 20000 10000 ) W ;
- u3 INV + UNPLACED ;
- u4 INV + SOURCE DIST + PLACED ( 9000 8000 ) E ;
END COMPONENTS
PINS 1 ;
- in1 + NET n1 + DIRECTION INPUT + USE SIGNAL
  + LAYER M2 ( -50 0 ) ( 50 100 ) + PLACED ( 1000 15000 ) N ;
END PINS
SPECIALNETS 1 ;
- VDD ( * VDD ) + USE POWER ;
END SPECIALNETS
NETS 4 ;
- n1 ( PIN in1 ) ( u1 A ) ;
- n2 ( u1 Z ) ( u2 A ) ( b1 P ) + USE SIGNAL ;
- n3 ( u2 Z ) ( u4 A )
  + ROUTED M1 ( 100 200 ) ( 300 * ) ;
- nvdd ( * VDD ) ;
END NETS
END DESIGN
"""


def test_hand_design():
    d = LD.read_db(LEF, DEF)
    # movable u1 u2 u3 u4 (DEF order), fixed b1, IO pin in1
    assert d.num_movable == 4 and d.num_nodes == 6
    assert (d.die_w, d.die_h) == (40000.0, 30000.0)
    np.testing.assert_array_equal(d.x, [4000, 6000, 0, 8000, 19000, 0])
    np.testing.assert_array_equal(d.y, [4000, 4000, 0, 6000, 8000, 13000])
    np.testing.assert_array_equal(d.w, [600, 600, 600, 2000, 8000, 0])     # u4 (E), b1 (W) rotated
    np.testing.assert_array_equal(d.h, [2000, 2000, 2000, 600, 10000, 0])
    np.testing.assert_array_equal(d.num_pins, [2, 2, 0, 1, 1, 1])
    # nets: two-pin nets n1, n3 first, then n2; nvdd (only '*') dropped
    assert d.n_2nets == 2 and d.num_nets == 3
    np.testing.assert_array_equal(d.net_start, [0, 2, 4, 7])
    pins = [(int(d.pin_node[k]), float(d.pin_offset_x[k]), float(d.pin_offset_y[k]))
            for k in d.net_pins]
    assert pins == [
        (5, 0.0, 0.0), (0, 150.0, 700.0),                        # n1: in1, u1/A (N)
        (1, 450.0, 1100.0), (3, 700.0, 450.0),                   # n3: u2/Z (FS), u4/A (E)
        (0, 450.0, 900.0), (1, 150.0, 1300.0), (4, 6000.0, 1500.0),   # n2: u1/Z, u2/A, b1/P (W)
    ]
    assert d.movable_area == pytest.approx(4 * 600 * 2000)
    assert d.cell_utilization == pytest.approx(4.8e6 / (40000 * 30000 - 8000 * 10000))


_R90 = np.array([[0, -1], [1, 0]])
_MX = np.array([[1, 0], [0, -1]])
_MY = np.array([[-1, 0], [0, 1]])
_MATS = {"N": np.eye(2), "W": _R90, "S": _R90 @ _R90, "E": _R90 @ _R90 @ _R90,
         "FN": _MY, "FS": _MX, "FW": _R90 @ _MX, "FE": _R90 @ _MY}


@pytest.mark.parametrize("orient", sorted(_MATS))
def test_orientation_matches_matrix_definition(orient):
    """Offset = M p - (lower-left of M applied to the master box)."""
    w, h = 7.0, 3.0
    corners = np.array([[0, 0], [w, 0], [0, h], [w, h]]).T
    M = _MATS[orient]
    shift = (M @ corners).min(axis=1)
    rng = np.random.default_rng(len(orient))
    for px, py in rng.uniform(0, 1, (20, 2)) * [w, h]:
        expect = M @ np.array([px, py]) - shift
        got = LD.pin_offset(orient, px, py, w, h)
        assert np.allclose(got, expect)


def test_unknown_references_raise():
    with pytest.raises(ValueError, match="macro"):
        LD.read_db(LEF, DEF.replace("- u3 INV", "- u3 NAND9"))
    with pytest.raises(ValueError, match="no pin"):
        LD.read_db(LEF, DEF.replace("( u4 A )", "( u4 Q )"))
    with pytest.raises(ValueError, match="DIEAREA"):
        LD.read_db(LEF, DEF.replace("DIEAREA ( 1000 2000 ) ( 41000 32000 ) ;", ""))
    with pytest.raises(ValueError, match="orientation"):
        LD.pin_offset("NE", 0, 0, 1, 1)


def test_generated_def_round_trip():
    """A DEF written from random instances and pins reads back to the same arrays."""
    rng = np.random.default_rng(11)
    n_mov, n_fix, units = 40, 5, 2000
    widths = rng.integers(1, 9, n_mov + n_fix) * 0.2
    heights = np.where(np.arange(n_mov + n_fix) < n_mov, 2.0, rng.integers(2, 6, n_mov + n_fix) * 2.0)
    lef = ["UNITS DATABASE MICRONS 2000 ; END UNITS"]
    for i in range(n_mov + n_fix):
        lef.append(f"MACRO M{i} SIZE {widths[i]:.1f} BY {heights[i]:.1f} ;")
        for p in range(3):
            cx, cy = rng.uniform(0, widths[i]), rng.uniform(0, heights[i])
            lef.append(f"PIN P{p} PORT LAYER M1 ; RECT {cx - 0.01:.4f} {cy - 0.01:.4f} "
                       f"{cx + 0.01:.4f} {cy + 0.01:.4f} ; END END P{p}")
        lef.append(f"END M{i}")
    xs = rng.integers(0, 5000, n_mov + n_fix) * 20
    ys = rng.integers(0, 5000, n_mov + n_fix) * 20
    comps = [f"- c{i} M{i} + {'PLACED' if i < n_mov else 'FIXED'} ( {xs[i]} {ys[i]} ) N ;"
             for i in range(n_mov + n_fix)]
    nets = []
    for e in range(60):
        k = rng.integers(2, 5)
        members = rng.choice(n_mov + n_fix, k, replace=False)
        nets.append(f"- e{e} " + " ".join(f"( c{m} P{rng.integers(0, 3)} )" for m in members) + " ;")
    deff = "\n".join(["UNITS DISTANCE MICRONS 2000 ;", "DIEAREA ( 0 0 ) ( 100000 100000 ) ;",
                      f"COMPONENTS {len(comps)} ;", *comps, "END COMPONENTS",
                      f"NETS {len(nets)} ;", *nets, "END NETS"])
    d = LD.read_db("\n".join(lef), deff)
    assert d.num_movable == n_mov and d.num_nodes == n_mov + n_fix
    np.testing.assert_array_equal(d.x, xs.astype(np.float32))
    np.testing.assert_array_equal(d.y, ys.astype(np.float32))
    np.testing.assert_allclose(d.w, np.round(widths * units), rtol=0, atol=0)
    np.testing.assert_allclose(d.h, np.round(heights * units), rtol=0, atol=0)
    deg = np.diff(d.net_start)
    assert np.all(deg[:d.n_2nets] == 2) and np.all(deg[d.n_2nets:] > 2)
    assert int(d.num_pins.sum()) == d.num_pin_total
    assert np.all(d.pin_offset_x >= 0) and np.all(d.pin_offset_x <= d.w[d.pin_node])
    assert np.all(d.pin_offset_y >= 0) and np.all(d.pin_offset_y <= d.h[d.pin_node])


@pytest.mark.gpu
def test_lefdef_design_places(gpu_lib):
    from erplacer.placement import Placement
    d = LD.read_db(LEF, DEF)
    with Placement(d, 64) as P:
        info = P.run(30)
    assert info["iteration"] == 30
    assert np.isfinite(info["hpwl"]) and np.isfinite(info["overflow"])


def _generated_lefdef(n_mov=3000, n_fix=4, n_nets=3200, seed=23):
    """A LEF/DEF pair of standard cells on 0.2 x 2.0 um sites and a few blocks,
    movable cells at random legal sites, nets of 2-6 pins."""
    rng = np.random.default_rng(seed)
    lef = ["UNITS DATABASE MICRONS 2000 ; END UNITS"]
    for k in range(1, 9):
        lef.append(f"MACRO C{k} CLASS CORE ; SIZE {0.2 * k:.1f} BY 2.0 ;")
        for p in range(3):
            lef.append(f"PIN P{p} PORT LAYER M1 ; RECT {0.02 + 0.05 * p:.2f} 0.4 "
                       f"{0.06 + 0.05 * p:.2f} 0.8 ; END END P{p}")
        lef.append(f"END C{k}")
    lef.append("MACRO B CLASS BLOCK ; SIZE 30 BY 24 ; PIN P0 PORT LAYER M1 ; RECT 1 1 2 2 ; "
               "END END P0 PIN P1 PORT LAYER M1 ; RECT 10 20 11 21 ; END END P1 "
               "PIN P2 PORT LAYER M1 ; RECT 25 5 26 6 ; END END P2 END B")
    die = 110_000                                     # DEF units (1000 per micron)
    comps = []
    for i in range(n_mov):
        k = int(rng.integers(1, 9))
        comps.append(f"- c{i} C{k} + PLACED ( {int(rng.integers(0, 540)) * 200} "
                     f"{int(rng.integers(0, 54)) * 2000} ) N ;")
    for j in range(n_fix):
        comps.append(f"- b{j} B + FIXED ( {10000 + 50000 * (j % 2)} "
                     f"{10000 + 50000 * (j // 2)} ) N ;")
    names = [f"c{i}" for i in range(n_mov)] + [f"b{j}" for j in range(n_fix)]
    nets = []
    for e in range(n_nets):
        k = int(rng.choice([2, 2, 2, 3, 3, 4, 5, 6]))
        base = int(rng.integers(0, n_mov))
        members = {min(n_mov - 1, base + int(rng.integers(0, 40))) for _ in range(k)}
        if rng.random() < 0.02:
            members.add(n_mov + int(rng.integers(0, n_fix)))
        if len(members) < 2:
            members.add((base + 1) % n_mov)
        nets.append(f"- n{e} " + " ".join(f"( {names[m]} P{int(rng.integers(0, 3))} )"
                                          for m in sorted(members)) + " ;")
    deff = "\n".join(["UNITS DISTANCE MICRONS 1000 ;", f"DIEAREA ( 0 0 ) ( {die} {die} ) ;",
                      f"COMPONENTS {len(comps)} ;", *comps, "END COMPONENTS",
                      f"NETS {len(nets)} ;", *nets, "END NETS"])
    return "\n".join(lef), deff


@pytest.mark.gpu
def test_lefdef_design_tracks_oracle(gpu_lib):
    """A parsed LEF/DEF design (3000 standard cells, 4 blocks) placed on the GPU
    against the FP64 oracle on the same parsed design (main.rs:68-93): step 0 to
    rounding, the first ten steps within the tracking bounds of the generator's
    designs, and the density map bit for bit at step 0."""
    import math

    import oracle_lib as O
    from erplacer import lib as L
    from erplacer.placement import Placement
    lef, deff = _generated_lefdef()
    d = LD.read_db(lef, deff)
    assert d.num_movable == 3000 and d.num_nodes == 3004 and d.n_2nets > 0
    ref = O.Placer(d, 64)
    with Placement(d, 64) as P:
        P.set_keep_density(True)
        x0, _ = P.positions()
        rx0, _ = ref.positions()
        assert np.array_equal(x0.astype(np.float64), rx0)
        h = P.compute_hpwl()
        _, ovf = P.compute_potential()
        r = ref.step()
        assert math.isclose(h, r["hpwl"], rel_tol=1e-5) and abs(ovf - r["overflow"]) < 2e-5
        q, _, scale = P.density_fixed()
        st = P.setup()
        C = d.num_movable
        y0 = P.positions()[1]
        q_ref = O.density_map_fx64(x0[:C], y0[:C], st["cw"][:C], st["ch"][:C], st["ox"][:C],
                                   st["oy"][:C], st["ratio"][:C], 64, 64, st["bw"], st["bh"],
                                   scale, sfx="_f32")
        assert np.array_equal(q, q_ref)
        P.nesterov_step()
        P.scheduler_update()
        P.set_keep_density(False)
        for k in range(1, 10):
            a = P.step()
            b = ref.step()
            assert math.isclose(a["hpwl"], b["hpwl"], rel_tol=1e-3), k
            assert abs(a["overflow"] - b["overflow"]) < 1e-3, k
        assert L.SPECTRAL_FFT == 0
