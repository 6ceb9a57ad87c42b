"""Diagnostics parity: the progress line of placement.rs:213-223, the placement
drawing of drawing.rs:7-60 and the density heat map of epotential.rs:200-223.

The Rust formatting is pinned by the values Rust's documentation and its f32
shortest-representation rule fix (``{:e}`` of 42.0 is ``4.2e1``), and by the
round-trip/shortness property checked digit by digit.  The drawings are pinned
by hand-placed rectangles and by a per-pixel painter on random designs.
"""
import logging
import os
import struct
import zlib

import numpy as np
import pytest

from erplacer import diagnostics as D


# ---------------------------------------------------------------- formatting
@pytest.mark.parametrize("v,expect", [
    (42.0, "4.2e1"), (1234.5, "1.2345e3"), (0.1, "1e-1"), (1.0, "1e0"), (0.0, "0e0"),
    (-0.0, "-0e0"), (-1.5, "-1.5e0"), (123456789.0, "1.2345679e8"), (8e-5, "8e-5"),
    (1e10, "1e10"), (3.4028235e38, "3.4028235e38"), (1e-45, "1e-45"), (0.3, "3e-1"),
    (float("inf"), "inf"), (float("-inf"), "-inf"), (float("nan"), "NaN"),
])
def test_lower_exp_values(v, expect):
    assert D.rust_lower_exp(v) == expect


def test_lower_exp_width_and_zero_flag():
    assert D.rust_lower_exp(42.0, 8, True) == "0004.2e1"
    assert D.rust_lower_exp(-1.5, 8, True) == "-001.5e0"
    assert D.rust_lower_exp(42.0, 8, False) == "   4.2e1"
    assert D.rust_lower_exp(123456789.0, 8, True) == "1.2345679e8"   # longer than the width
    assert D.rust_lower_exp(float("nan"), 8, True) == "     NaN"


def test_lower_exp_is_shortest_round_trip():
    """The digits printed are the fewest that read back as the same f32."""
    rng = np.random.default_rng(5)
    vals = np.concatenate([rng.standard_normal(200) * 10.0 ** rng.integers(-30, 30, 200),
                           rng.random(100).astype(np.float32)]).astype(np.float32)
    for f in vals:
        s = D.rust_lower_exp(f)
        mant, ex = s.split("e")
        assert "+" not in ex and not ex.lstrip("-").startswith("0") or ex == "0"
        assert np.float32(float(s)) == f
        digits = mant.replace("-", "").replace(".", "")
        if len(digits) > 1:       # one digit fewer (either rounding) does not round-trip
            e10 = int(ex) - (len(digits) - 1)
            for cand in (int(digits) // 10, int(digits) // 10 + 1):
                shorter = cand * 10.0 ** (e10 + 1) * (-1 if f < 0 else 1)
                assert np.float32(shorter) != f


def test_log_line():
    assert (D.log_line(7, 42.0, 0.1, 8e-5, 4.0)
            == "[007] - HPWL 0004.2e1 - Overflow 00001e-1 - DW 00008e-5 - G 000004e0")
    assert D.log_line(1000, 1.0, 0.0, 1.0, 1.0).startswith("[1000] - HPWL 000001e0")
    assert D.SEPARATOR == "=" * 40 + "\n"


# ---------------------------------------------------------------- PNG
def test_png_round_trip_and_chunks(tmp_path):
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (7, 5, 3), dtype=np.uint8)
    p = os.path.join(tmp_path, "sub", "a.png")
    D.write_png(p, img)
    data = open(p, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, kinds = 8, []
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(kind + body) & 0xFFFFFFFF
        kinds.append(kind)
        pos += 12 + n
    assert kinds == [b"IHDR", b"IDAT", b"IEND"]
    assert np.array_equal(D.read_png(p), img)


# ---------------------------------------------------------------- density map
def test_density_map_orientation_and_colours():
    m = np.array([[0.0, 1.0, 2.0], [3.0, 4.0, 5.0]], np.float32)   # M=2 rows (x), N=3 (y)
    img = D.density_map_image(m)
    assert img.shape == (3, 2, 3)                  # height N, width M
    # bin (i, j) -> column i, row N-1-j
    assert tuple(img[2, 0]) == (0, 0, 255)         # (0,0) = min: blue
    assert tuple(img[0, 1]) == (255, 0, 0)         # (1,2) = max: red
    # (0,2): t = 2/5 = 0.4 -> red (0.4*255) as u8 = 102, blue (0.6*255) as u8 = 153
    t = np.float32(2.0) / (np.float32(5.0) + np.float32(1e-12))
    r = int(np.trunc(t * np.float32(255.0)))
    b = int(np.trunc((np.float32(1.0) - t) * np.float32(255.0)))
    assert tuple(img[0, 0]) == (r, 0, b)
    assert (r, b) == (102, 153)


def test_density_map_uniform_is_blue(tmp_path):
    p = D.draw_density_map(np.full((4, 4), 3.0, np.float32), 12, out_dir=str(tmp_path))
    assert p.endswith("density_map_00012.png")
    img = D.read_png(p)
    assert np.all(img == np.array([0, 0, 255], np.uint8))


# ---------------------------------------------------------------- placement drawing
def _paint_reference(x, y, w, h, nm, die_w, die_h):
    """Per-instance, per-pixel painter following drawing.rs:24-51 literally."""
    W, H = int(np.ceil(np.float32(die_w))), int(np.ceil(np.float32(die_h)))
    img = np.full((H, W, 3), 255, np.uint8)
    for i in range(len(x)):
        wi, hi = int(np.ceil(np.float32(w[i]))), int(np.ceil(np.float32(h[i])))
        xi = int(np.ceil(np.float32(x[i])))
        yi = H - int(np.ceil(np.float32(y[i]))) - hi
        col = (0, 0, 255) if i < nm else (255, 0, 0)
        if wi == 0 or hi == 0:
            continue
        for r in range(yi, yi + hi):
            for c in range(xi, xi + wi):
                if 0 <= r < H and 0 <= c < W:
                    img[r, c] = col
    return img


def test_placement_image_hand_cases():
    # die 10 x 6, instances 0-2 movable and 3-5 fixed: 0 at (1.2, 0.3) 2x1;
    # 1 at (0, 4) 3x2 (top-left corner); 2 overlapped by the later fixed 3 (red
    # on top); 4 a zero-width pin; 5 running off the right edge
    x = [1.2, 0.0, 5.0, 5.5, 4.0, 9.0]
    y = [0.3, 4.0, 2.0, 2.0, 1.0, 0.0]
    w = [2.0, 3.0, 2.0, 1.0, 0.0, 3.0]
    h = [1.0, 2.0, 1.0, 1.0, 1.0, 1.0]
    nm = 3
    img = D.placement_image(x, y, w, h, nm, 10.0, 6.0)
    assert img.shape == (6, 10, 3)
    assert tuple(img[4, 2]) == (0, 0, 255) and tuple(img[4, 3]) == (0, 0, 255)   # instance 0
    assert tuple(img[4, 1]) == (255, 255, 255)
    assert tuple(img[0, 0]) == (0, 0, 255)          # instance 1 is movable here (index < 3)
    assert tuple(img[3, 6]) == (255, 0, 0)          # instance 3 (fixed) over instance 2
    assert tuple(img[3, 5]) == (0, 0, 255)
    assert tuple(img[5, 9]) == (255, 0, 0)          # clipped at the right edge
    assert np.array_equal(img, _paint_reference(x, y, w, h, nm, 10.0, 6.0))


def test_placement_image_matches_painter_random(tmp_path):
    rng = np.random.default_rng(3)
    n, nm = 300, 240
    die_w, die_h = 61.5, 37.2
    x = rng.uniform(-3, die_w, n).astype(np.float32)
    y = rng.uniform(-3, die_h, n).astype(np.float32)
    w = rng.choice([0.0, 0.5, 1.0, 2.3, 7.9], n).astype(np.float32)
    h = rng.choice([0.0, 1.0, 1.7, 4.0], n).astype(np.float32)
    ref = _paint_reference(x, y, w, h, nm, die_w, die_h)
    got = D.placement_image(x, y, w, h, nm, die_w, die_h, max_pixels_per_pass=37)
    assert np.array_equal(got, ref)
    p = D.draw_placement(x, y, w, h, nm, die_w, die_h, 50, out_dir=str(tmp_path))
    assert p.endswith("solution_0000050.png")
    assert np.array_equal(D.read_png(p), ref)


# ---------------------------------------------------------------- logged loop (GPU)
@pytest.mark.gpu
def test_logged_run_matches_graph_run(gpu_lib, caplog):
    from erplacer import workloads as W
    from erplacer.placement import Placement
    d, g, iters, td = W.make("small")
    with Placement(d, g, target_density=td) as A, Placement(d, g, target_density=td) as B:
        ia = A.run(25)
        with caplog.at_level(logging.INFO, logger="erplacer"):
            ib = D.run_logged(B, 25)
        assert ia["hpwl"] == ib["hpwl"] and ia["overflow"] == ib["overflow"]
        assert ia["iteration"] == ib["iteration"] == 25
        xa, ya = A.positions()
        xb, yb = B.positions()
        assert np.array_equal(xa, xb) and np.array_equal(ya, yb)
    lines = [r.getMessage() for r in caplog.records if r.name == "erplacer"]
    heads = [l for l in lines if l.startswith("[")]
    assert [l[:5] for l in heads] == ["[000]", "[010]", "[020]"]
    assert lines.count(D.SEPARATOR) == 3
    for l in heads:
        parts = l.split(" - ")
        assert [p.split(" ")[0] for p in parts[1:]] == ["HPWL", "Overflow", "DW", "G"]
        for p in parts[1:]:
            v = p.split(" ")[1]
            assert len(v) >= 8 and np.isfinite(np.float32(float(v)))


@pytest.mark.gpu
def test_logged_lines_and_density_png_match_oracle(gpu_lib, caplog, tmp_path):
    """The GPU run's [000]/[010]/[020] lines against log_line of the FP64
    oracle's own scalars at the same iterations (placement.rs:213-223: HPWL and
    overflow of the step, the weight after the Nesterov step, the gamma the
    step used), field by field; and the step-0 density PNG of the GPU map
    against draw_density_map of the oracle's map (epotential.rs:200-223)."""
    import oracle_lib as O
    from erplacer import lib as L
    from erplacer import workloads as W
    from erplacer.placement import Placement
    d, g, _, td = W.make("small")
    ref = O.Placer(d, g, target_density=td)
    gamma = ref.setup()["gamma"]
    expect = {}
    for k in range(21):
        r = ref.step()
        if k % 10 == 0:
            # the weight is logged after the Nesterov step and before the
            # scheduler update; the first update leaves it unchanged (scheduler.rs:
            # 102-113), later ones multiply it, so take it from the step before
            dw = r["density_weight"] if k == 0 else prev_dw
            expect[k] = D.log_line(k, r["hpwl"], r["overflow"], dw, gamma)
        gamma = r["gamma"]
        prev_dw = r["density_weight"]
    with Placement(d, g, target_density=td) as P:
        with caplog.at_level(logging.INFO, logger="erplacer"):
            D.run_logged(P, 21)
    heads = [r.getMessage() for r in caplog.records
             if r.name == "erplacer" and r.getMessage().startswith("[")]
    assert len(heads) == 3
    for line in heads:
        k = int(line[1:4])
        want = expect[k]
        gv = [p.split(" ")[1] for p in line.split(" - ")[1:]]
        wv = [p.split(" ")[1] for p in want.split(" - ")[1:]]
        assert [p.split(" ")[0] for p in line.split(" - ")] == [p.split(" ")[0] for p in want.split(" - ")]
        for a, b in zip(gv, wv):
            assert len(a) >= 8 and len(b) >= 8
            assert abs(float(a) / float(b) - 1) < 5e-3, (line, want)
        if k == 0:
            # before the trajectories part, the f32 strings agree to 4 digits
            for a, b in zip(gv, wv):
                assert abs(float(a) / float(b) - 1) < 1e-4, (line, want)
    # step-0 density map PNG
    ref0 = O.Placer(d, g, target_density=td)
    ref0.step()
    with Placement(d, g, target_density=td) as P:
        P.compute_potential()
        gm = P.map(L.MAP_DENSITY)
    pa = D.draw_density_map(gm, 0, out_dir=str(tmp_path / "gpu"))
    pb = D.draw_density_map(ref0.map(O.MAP_DENSITY).astype(np.float32), 0, out_dir=str(tmp_path / "ref"))
    a, b = D.read_png(pa), D.read_png(pb)
    assert a.shape == b.shape == (g, g, 3)
    diff = np.abs(a.astype(int) - b.astype(int))
    assert diff.max() <= 1 and np.mean(diff.any(axis=-1)) < 1e-3
