"""The named synthetic workloads (sizes, grids, iteration counts).

Each workload fixes the generator parameters, the bin grid and the iteration count.
The seed of workload ``k`` (in the order below) is ``20261203 + k``.
"""
from __future__ import annotations

import math

import numpy as np

from .netlist import generate, tail_alpha_for

SEED_BASE = 20261203

WORKLOADS = {
    # 10k cells, 128x128 bins, FP64 CPU case, 300 iterations
    "small": dict(index=0, cells=10_000, nets=10_500, cap=64, macros=20, macro_frac=0.10,
                  util=0.70, grid=128, iterations=300, target_density=1.0),
    # 200k cells, 512x512 bins, 600 iterations
    "medium": dict(index=1, cells=200_000, nets=210_000, cap=128, macros=100, macro_frac=0.10,
                   util=0.75, grid=512, iterations=600, target_density=1.0),
    # 1M cells, 1024x1024 bins, 800 iterations
    "large": dict(index=2, cells=1_000_000, nets=1_050_000, cap=256, macros=300,
                  macro_frac=0.10, util=0.80, grid=1024, iterations=800, target_density=1.0),
    # 2M cells, 2048x2048 bins, 1000 iterations, 2% of nets above 32 pins
    "stress": dict(index=4, cells=2_000_000, nets=2_100_000, cap=512, macros=1000,
                   macro_frac=0.25, util=0.85, grid=2048, iterations=1000, target_density=1.0,
                   frac_above_32=0.02, macro_aspect=(1 / 3, 3.0)),
}

SERVING = dict(index=3, requests=512, cells_lo=2_000, cells_hi=150_000, nets_ratio=1.05,
               util_lo=0.60, util_hi=0.85, grid_lo=64, grid_hi=512, cap=128,
               stop_overflow=0.1, target_density=1.0)


def make(name: str, scale: float = 1.0):
    """Design + grid + iterations for a named workload.

    ``scale`` < 1 shrinks cells, nets and macros proportionally (the grid is kept),
    which is how the parity tests get oracle-sized cases of the same shape.
    """
    p = WORKLOADS[name]
    cells = max(16, int(round(p["cells"] * scale)))
    nets = max(16, int(round(p["nets"] * scale)))
    macros = max(1, int(round(p["macros"] * math.sqrt(scale)))) if p["macros"] else 0
    alpha = 2.5
    if "frac_above_32" in p:
        alpha = tail_alpha_for(p["frac_above_32"], 32, 4, p["cap"], 0.15)
    d = generate(cells, nets, p["cap"], macros, p["macro_frac"], p["util"],
                 SEED_BASE + p["index"], tail_alpha=alpha,
                 macro_aspect=p.get("macro_aspect", (0.5, 2.0)), name=name)
    return d, p["grid"], p["iterations"], p["target_density"]


def serving_grid(cells: int) -> int:
    """next power of two of sqrt(cells)/2, clamped to [64, 512]."""
    g = 1 << max(0, math.ceil(math.log2(max(1.0, math.sqrt(cells) / 2))))
    return int(min(SERVING["grid_hi"], max(SERVING["grid_lo"], g)))


def serving_requests(n: int | None = None, seed_offset: int = 0, rate_per_s: float = 20.0):
    """The serving stream: (arrival_time_s, design, grid) per request, seeded."""
    s = SERVING
    n = s["requests"] if n is None else n
    rng = np.random.Generator(np.random.PCG64(SEED_BASE + s["index"] + seed_offset))
    gaps = rng.exponential(1.0 / rate_per_s, size=n)
    arrivals = np.cumsum(gaps) - gaps[0]
    out = []
    for i in range(n):
        cells = int(round(math.exp(rng.uniform(math.log(s["cells_lo"]), math.log(s["cells_hi"])))))
        util = float(rng.uniform(s["util_lo"], s["util_hi"]))
        nets = int(round(cells * s["nets_ratio"]))
        macros = int(min(100, max(4, round(20 * math.sqrt(cells / 10_000)))))
        d = generate(cells, nets, s["cap"], macros, 0.10, util,
                     SEED_BASE + 1000 * (s["index"] + 1) + i + seed_offset, name=f"req{i}")
        out.append((float(arrivals[i]), d, serving_grid(cells)))
    return out
