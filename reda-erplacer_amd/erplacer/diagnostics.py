"""Diagnostics of the placement loop, printed and drawn as the reference does.

* ``log_line`` / ``step_logged`` / ``run_logged`` -- the progress line of
  ``Placement::step`` (placement.rs:213-223): every 10th iteration,
  ``[{:03}] - HPWL {:08e} - Overflow {:08e} - DW {:08e} - G {:08e}`` on the f32
  values, then a line of 40 ``=``.
* ``draw_placement`` -- drawing.rs:7-60: a die-sized RGB image, white
  background, movable instances blue and fixed ones red, drawn in instance
  order (a later instance covers an earlier one), zero-size instances skipped.
* ``draw_density_map`` -- epotential.rs:200-223: a blue-to-red heat map of a
  bin map normalised by its min and max.

PNG files are written with zlib only (8-bit RGB, no filtering).
"""
from __future__ import annotations

import logging
import math
import os
import struct
import zlib

import numpy as np

log = logging.getLogger("erplacer")


# ---------------------------------------------------------------- Rust formatting
def rust_lower_exp(v, width: int = 0, zero_pad: bool = False) -> str:
    """Rust's ``{:e}`` of an f32 (shortest digits that round-trip in f32,
    exponent without sign or leading zeros: 42.0 -> ``4.2e1``), right-aligned in
    ``width``; ``zero_pad`` (the ``0`` flag) pads with zeros after the sign."""
    f = np.float32(v)
    finite = bool(np.isfinite(f))
    if np.isnan(f):
        s = "NaN"
    elif not finite:
        s = "inf" if f > 0 else "-inf"
    else:
        t = np.format_float_scientific(f, unique=True, trim="-")
        mant, ex = t.split("e")
        s = f"{mant}e{int(ex)}"
    if len(s) < width:
        if zero_pad and finite:
            sign = "-" if s.startswith("-") else ""
            s = sign + "0" * (width - len(s)) + s[len(sign):]
        else:
            s = " " * (width - len(s)) + s
    return s


def log_line(iteration: int, hpwl, overflow, density_weight, gamma) -> str:
    """placement.rs:215-221 for one iteration."""
    e = lambda v: rust_lower_exp(v, 8, True)
    return (f"[{int(iteration):03d}] - HPWL {e(hpwl)} - Overflow {e(overflow)} "
            f"- DW {e(density_weight)} - G {e(gamma)}")


SEPARATOR = "=" * 40 + "\n"      # placement.rs:222, log::info!("{}\n", "=".repeat(40))


def step_logged(P, logger: logging.Logger | None = None) -> bool:
    """One ``Placement::step`` (placement.rs:180-228) phase by phase, logging
    before the scheduler update as the reference does; returns should_stop.
    Results are bit-identical to ``P.step()`` / ``P.run(1)``."""
    logger = logger or log
    before = P.info()
    iteration, gamma = int(before["iteration"]), before["gamma"]
    hpwl = P.compute_hpwl()
    _, overflow = P.compute_potential()
    P.nesterov_step()
    if iteration % 10 == 0:
        dw = P.info()["density_weight"]
        logger.info(log_line(iteration, hpwl, overflow, dw, gamma))
        logger.info(SEPARATOR)
    P.scheduler_update()
    return bool(P.info()["should_stop"])


def run_logged(P, iterations: int, logger: logging.Logger | None = None) -> dict:
    """``iterations`` steps with the reference's progress log: the logged
    iterations run phase by phase, the nine between them from the captured
    graph.  Returns the last step's info."""
    done = 0
    while done < iterations:
        it = int(P.info()["iteration"])
        if it % 10 == 0:
            step_logged(P, logger)
            done += 1
            continue
        chunk = min(10 - it % 10, iterations - done)
        P.run(chunk)
        done += chunk
    return P.info()


# ---------------------------------------------------------------- PNG
def write_png(path: str, rgb: np.ndarray) -> None:
    """8-bit RGB PNG of an [H, W, 3] uint8 array (filter type 0 on every row)."""
    rgb = np.ascontiguousarray(rgb, np.uint8)
    h, w, c = rgb.shape
    if c != 3:
        raise ValueError("rgb must be [H, W, 3]")
    raw = np.zeros((h, 1 + 3 * w), np.uint8)
    raw[:, 1:] = rgb.reshape(h, 3 * w)

    def chunk(kind: bytes, data: bytes) -> bytes:
        return (struct.pack(">I", len(data)) + kind + data
                + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF))

    png = (b"\x89PNG\r\n\x1a\n"
           + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0))
           + chunk(b"IDAT", zlib.compress(raw.tobytes(), 6))
           + chunk(b"IEND", b""))
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "wb") as f:
        f.write(png)


def read_png(path: str) -> np.ndarray:
    """Inverse of ``write_png`` (8-bit RGB, filter type 0 only)."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG file")
    pos, idat, w, h = 8, b"", 0, 0
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        if kind == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            if depth != 8 or ctype != 2:
                raise ValueError("only 8-bit RGB is read")
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 3 * w)
    if np.any(raw[:, 0] != 0):
        raise ValueError("only filter type 0 is read")
    return raw[:, 1:].reshape(h, w, 3).copy()


# ---------------------------------------------------------------- drawings
WHITE, BLUE, RED = (255, 255, 255), (0, 0, 255), (255, 0, 0)


def placement_image(x, y, w, h, num_movable: int, die_w: float, die_h: float,
                    max_pixels_per_pass: int = 1 << 23) -> np.ndarray:
    """The RGB raster of drawing.rs:7-55 ([height, width, 3] uint8).

    Image size ceil(die) (drawing.rs:12-13).  Instance i is the rectangle at
    column ceil(x_i), row height - ceil(y_i) - ceil(h_i) of size ceil(w_i) x
    ceil(h_i), clipped to the image; movable (i < num_movable) blue, fixed red;
    instances with a zero width or height are skipped (:44-46); a later
    instance paints over an earlier one (:29-51).
    """
    x, y, w, h = (np.asarray(a, np.float32) for a in (x, y, w, h))
    W, H = int(math.ceil(float(np.float32(die_w)))), int(math.ceil(float(np.float32(die_h))))
    if W <= 0 or H <= 0:
        raise ValueError("die area must be positive")
    if np.any(w < 0) or np.any(h < 0):
        raise ValueError("instance sizes must be non-negative (to_u32 would panic)")
    wi = np.ceil(w).astype(np.int64)
    hi = np.ceil(h).astype(np.int64)
    xi = np.ceil(x).astype(np.int64)
    yi = H - np.ceil(y).astype(np.int64) - hi
    x0, x1 = np.clip(xi, 0, W), np.clip(xi + wi, 0, W)
    y0, y1 = np.clip(yi, 0, H), np.clip(yi + hi, 0, H)
    keep = (wi > 0) & (hi > 0) & (x1 > x0) & (y1 > y0)
    idx = np.nonzero(keep)[0]
    owner = np.full(W * H, -1, np.int64)            # last instance covering each pixel
    area = (x1[idx] - x0[idx]) * (y1[idx] - y0[idx])
    s = 0
    while s < idx.size:                              # passes bounded in pixels
        csum = np.cumsum(area[s:])
        e = s + max(1, int(np.searchsorted(csum, max_pixels_per_pass, side="right")))
        r = idx[s:e]
        cnt = area[s:e]
        rid = np.repeat(r, cnt)
        k = np.arange(rid.size, dtype=np.int64) - np.repeat(np.cumsum(cnt) - cnt, cnt)
        wx = (x1 - x0)[rid]
        px = x0[rid] + k % wx
        py = y0[rid] + k // wx
        np.maximum.at(owner, py * W + px, rid)
        s = e
    img = np.empty((W * H, 3), np.uint8)
    img[:] = WHITE
    img[(owner >= 0) & (owner < num_movable)] = BLUE
    img[owner >= num_movable] = RED
    return img.reshape(H, W, 3)


def draw_placement(x, y, w, h, num_movable: int, die_w: float, die_h: float, iteration: int,
                   out_dir: str = "images") -> str:
    """drawing.rs:7-60: writes ``{out_dir}/solution_{iteration:07}.png``."""
    path = os.path.join(out_dir, f"solution_{int(iteration):07d}.png")
    write_png(path, placement_image(x, y, w, h, num_movable, die_w, die_h))
    return path


def draw_placement_of(P, iteration: int, out_dir: str = "images") -> str:
    """``draw_placement(db, &solution, iteration)`` for a ``Placement``: current
    positions, the design's (unclamped) sizes."""
    d = P.design
    x, y = P.positions()
    return draw_placement(x, y, d.w, d.h, d.num_movable, d.die_w, d.die_h, iteration, out_dir)


def density_map_image(m: np.ndarray) -> np.ndarray:
    """The RGB raster of epotential.rs:201-221 for a row-major [M, N] map
    (index i * N + j): width M, height N, bin (i, j) at column i, row N - 1 - j;
    t = clamp((v - min) / (max - min + 1e-12), 0, 1) in f32, red = (t * 255) as u8,
    blue = ((1 - t) * 255) as u8, green 0."""
    a = np.asarray(m, np.float32)
    if a.ndim != 2:
        raise ValueError("map must be [M, N]")
    lo = np.float32(np.min(a)) if a.size else np.float32(0)
    hi = np.float32(np.max(a)) if a.size else np.float32(0)
    den = np.float32(hi - lo) + np.float32(1e-12)
    t = np.clip((a - lo) / den, np.float32(0), np.float32(1)).astype(np.float32)

    def as_u8(v):                                   # Rust `as u8`: truncate, saturate, NaN -> 0
        v = np.nan_to_num(v, nan=0.0)
        return np.clip(np.trunc(v), 0, 255).astype(np.uint8)

    r = as_u8(t * np.float32(255.0))
    b = as_u8((np.float32(1.0) - t) * np.float32(255.0))
    rgb = np.stack([r, np.zeros_like(r), b], axis=-1)     # [M, N, 3] at (i, j)
    return np.ascontiguousarray(rgb.transpose(1, 0, 2)[::-1])   # [N, M, 3]: row N-1-j, column i


def draw_density_map(m: np.ndarray, index: int, out_dir: str = "density_map") -> str:
    """epotential.rs:201-223: writes ``{out_dir}/density_map_{index:05}.png``."""
    path = os.path.join(out_dir, f"density_map_{int(index):05d}.png")
    write_png(path, density_map_image(m))
    return path
