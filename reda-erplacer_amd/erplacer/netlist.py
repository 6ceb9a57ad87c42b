"""Seeded synthetic netlists in the reda-db data contract.

The reference reads LEF/DEF through reda-db into a structure-of-arrays ``DB<T>``
(main.rs:68-70).  The fields its placement loop touches are (placement.rs:52-134,
hpwl.rs:25-45,158-159, epotential.rs:69-95):

* ``instances.{coords.x/y, sizes.w/h, num_pins}`` with movable instances first,
  ``num_movable`` of them, fixed instances after (placement.rs:125);
* ``netlist.pins.{x,y}``: pin offsets from the owning instance origin;
* ``netlist.nets[i].pin_ids`` with nets ``[0, n_2nets)`` holding exactly two pins
  (hpwl.rs:173-181);
* ``netlist.macro_2_pins``: per-instance pin lists (hpwl.rs:57,103,127);
* ``diearea`` with origin 0 (density.rs:259-260), ``movable_area`` and
  ``cell_utilization`` (placement.rs:134).

This module produces exactly those fields from a seed, as CSR arrays instead of
``Vec<Vec<usize>>``.  Units are database units (DBU): a site is ``SITE_W`` DBU
wide and a row ``ROW_H`` DBU tall; every coordinate, size and pin offset is an
integer number of DBU, stored as float32 like the reference binary's ``T=f32``.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

SITE_W = 200.0
ROW_H = 2000.0


@dataclasses.dataclass
class Design:
    """One placement instance in the reda-db field layout (SoA, CSR)."""

    name: str
    seed: int
    die_w: float
    die_h: float
    num_movable: int
    x: np.ndarray            # float32 [num_nodes] DB coordinates (lower-left)
    y: np.ndarray
    w: np.ndarray            # float32 [num_nodes]
    h: np.ndarray
    num_pins: np.ndarray     # float32 [num_nodes]  (instances.num_pins as T)
    pin_offset_x: np.ndarray  # float32 [P]
    pin_offset_y: np.ndarray
    pin_node: np.ndarray     # int32 [P] owning instance of each pin
    net_start: np.ndarray    # int32 [E+1] CSR into net_pins
    net_pins: np.ndarray     # int32 [P] pin ids, net by net
    n_2nets: int
    node_pin_start: np.ndarray  # int32 [num_nodes+1] CSR (macro_2_pins)
    node_pins: np.ndarray       # int32 [P]
    movable_area: float
    cell_utilization: float
    params: dict

    @property
    def num_nodes(self) -> int:
        return int(self.x.shape[0])

    @property
    def num_nets(self) -> int:
        return int(self.net_start.shape[0] - 1)

    @property
    def num_pin_total(self) -> int:
        return int(self.pin_node.shape[0])

    def summary(self) -> dict:
        deg = np.diff(self.net_start)
        return {
            "name": self.name, "seed": self.seed,
            "cells": self.num_movable, "macros": self.num_nodes - self.num_movable,
            "nets": self.num_nets, "pins": self.num_pin_total, "n_2nets": self.n_2nets,
            "max_degree": int(deg.max()) if deg.size else 0,
            "die": [self.die_w, self.die_h],
            "movable_area": self.movable_area, "cell_utilization": self.cell_utilization,
        }


def from_nets(x, y, w, h, num_movable, nets, die_w, die_h, name="custom") -> Design:
    """A Design from instance arrays (movable first) and nets given as lists of
    (instance, pin offset x, pin offset y).  Two-pin nets are placed first (the
    contract's n_2nets prefix); pins are numbered net by net."""
    x, y, w, h = (np.asarray(a, np.float32) for a in (x, y, w, h))
    N = int(x.shape[0])
    order = sorted(range(len(nets)), key=lambda e: 0 if len(nets[e]) == 2 else 1)
    n2 = sum(1 for e in nets if len(e) == 2)
    pin_node, pox, poy, net_start = [], [], [], [0]
    for e in order:
        for node, ox, oy in nets[e]:
            pin_node.append(int(node))
            pox.append(float(ox))
            poy.append(float(oy))
        net_start.append(len(pin_node))
    pin_node = np.asarray(pin_node, np.int64)
    counts = np.bincount(pin_node, minlength=N) if pin_node.size else np.zeros(N, np.int64)
    node_pin_start = np.zeros(N + 1, np.int64)
    np.cumsum(counts, out=node_pin_start[1:])
    node_pins = np.argsort(pin_node, kind="stable")
    movable_area = float((w[:num_movable].astype(np.float64) * h[:num_movable]).sum())
    fixed_area = float((w[num_movable:].astype(np.float64) * h[num_movable:]).sum())
    return Design(
        name=name, seed=0, die_w=float(die_w), die_h=float(die_h), num_movable=int(num_movable),
        x=x, y=y, w=w, h=h, num_pins=counts.astype(np.float32),
        pin_offset_x=np.asarray(pox, np.float32), pin_offset_y=np.asarray(poy, np.float32),
        pin_node=pin_node.astype(np.int32), net_start=np.asarray(net_start, np.int32),
        net_pins=np.arange(pin_node.size, dtype=np.int32), n_2nets=n2,
        node_pin_start=node_pin_start.astype(np.int32), node_pins=node_pins.astype(np.int32),
        movable_area=movable_area,
        cell_utilization=movable_area / max(float(die_w) * float(die_h) - fixed_area, 1.0),
        params={},
    )


def _tail_degrees(rng, count, lo, cap, alpha):
    """Power-law degrees d in [lo, cap] with P(d) proportional to d**-alpha."""
    if count == 0:
        return np.zeros(0, np.int64)
    support = np.arange(lo, cap + 1, dtype=np.float64)
    p = support ** (-alpha)
    p /= p.sum()
    return rng.choice(np.arange(lo, cap + 1), size=count, p=p).astype(np.int64)


def tail_alpha_for(frac_above: float, threshold: int, lo: int, cap: int,
                   tail_share: float) -> float:
    """Exponent that puts ``frac_above`` of ALL nets above ``threshold`` pins,
    given that the power-law tail holds ``tail_share`` of the nets."""
    want = frac_above / tail_share
    lo_a, hi_a = 0.0, 6.0
    support = np.arange(lo, cap + 1, dtype=np.float64)
    for _ in range(80):
        mid = 0.5 * (lo_a + hi_a)
        p = support ** (-mid)
        share = p[support > threshold].sum() / p.sum()
        if share > want:
            lo_a = mid
        else:
            hi_a = mid
    return 0.5 * (lo_a + hi_a)


def _place_macros(rng, num_macros, die, total_area, aspect):
    """Non-overlapping fixed rectangles, snapped to the site/row grid."""
    if num_macros == 0:
        return np.zeros((0, 4))
    areas = rng.uniform(0.5, 1.5, num_macros)
    areas *= total_area / areas.sum()
    lo_a, hi_a = aspect
    ar = np.exp(rng.uniform(math.log(lo_a), math.log(hi_a), num_macros))
    rects = []
    # biggest first: random sequential placement with rejection
    for idx in np.argsort(-areas):
        a = areas[idx]
        for shrink in range(20):
            aa = a * (0.9 ** shrink)
            w = max(SITE_W, round(math.sqrt(aa * ar[idx]) / SITE_W) * SITE_W)
            h = max(ROW_H, round(aa / w / ROW_H) * ROW_H)
            w = min(w, die * 0.5)
            h = min(h, die * 0.5)
            placed = False
            for _ in range(200):
                x = math.floor(rng.uniform(0, die - w) / SITE_W) * SITE_W
                y = math.floor(rng.uniform(0, die - h) / ROW_H) * ROW_H
                ok = True
                for (rx, ry, rw, rh) in rects:
                    if x < rx + rw and rx < x + w and y < ry + rh and ry < y + h:
                        ok = False
                        break
                if ok:
                    rects.append((x, y, w, h))
                    placed = True
                    break
            if placed:
                break
    return np.asarray(rects, dtype=np.float64).reshape(-1, 4)


def generate(num_cells: int, num_nets: int, degree_cap: int, num_macros: int,
             macro_frac: float, utilization: float, seed: int, *,
             tail_alpha: float = 2.5, macro_aspect=(0.5, 2.0), locality: int = 64,
             macro_pin_prob: float = 0.05, name: str = "synthetic") -> Design:
    """Seeded netlist in the reda-db contract.

    Degrees: 60% two-pin, 25% three-pin, 15% power-law tail on [4, cap] with
    exponent ``tail_alpha``.  Cell widths are uniform in 1..8 sites, one row tall.
    The die is square and sized so that movable area / (die area - macro area)
    equals ``utilization``; macros cover ``macro_frac`` of the die.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    C, E, F = int(num_cells), int(num_nets), int(num_macros)

    # ---- cells
    wsites = rng.integers(1, 9, size=C)
    cw = wsites.astype(np.float64) * SITE_W
    ch = np.full(C, ROW_H)
    movable_area = float((cw * ch).sum())
    free_frac = 1.0 - macro_frac
    side = math.sqrt(movable_area / (utilization * free_frac))
    side = math.ceil(side / ROW_H) * ROW_H
    die = float(side)

    # ---- macros
    rects = _place_macros(rng, F, die, macro_frac * die * die, macro_aspect)
    F = rects.shape[0]
    fixed_area = float((rects[:, 2] * rects[:, 3]).sum()) if F else 0.0

    # DB coordinates of movable cells: a DEF-like random legal-ish spot (only the
    # initial pin positions of fixed cells and prev_coords use them).
    mx = np.floor(rng.uniform(0, die - cw) / SITE_W) * SITE_W
    my = np.floor(rng.uniform(0, die - ch) / ROW_H) * ROW_H
    x = np.concatenate([mx, rects[:, 0]]).astype(np.float32)
    y = np.concatenate([my, rects[:, 1]]).astype(np.float32)
    w = np.concatenate([cw, rects[:, 2]]).astype(np.float32)
    h = np.concatenate([ch, rects[:, 3]]).astype(np.float32)

    # ---- net degrees
    n2 = int(round(0.60 * E))
    n3 = int(round(0.25 * E))
    nt = E - n2 - n3
    tail = _tail_degrees(rng, nt, 4, degree_cap, tail_alpha)
    deg_multi = np.concatenate([np.full(n3, 3, np.int64), tail])
    rng.shuffle(deg_multi)
    degrees = np.concatenate([np.full(n2, 2, np.int64), deg_multi])
    degrees = np.minimum(degrees, C)  # a net cannot have more distinct cells than exist

    # ---- pin -> cell: anchor spread evenly over the cell order, the other pins
    # from a window around the anchor (netlist locality), distinct cells per net.
    order = rng.permutation(E)           # which net gets which anchor
    anchors = (order.astype(np.int64) * C) // max(E, 1)
    P = int(degrees.sum())
    net_start = np.zeros(E + 1, np.int64)
    np.cumsum(degrees, out=net_start[1:])
    pin_node = np.empty(P, np.int64)
    # window offsets, rejection of duplicates done per degree group vectorised
    for d in np.unique(degrees):
        idx = np.nonzero(degrees == d)[0]
        k = idx.shape[0]
        cells = np.empty((k, d), np.int64)
        cells[:, 0] = anchors[idx]
        win = max(locality, 2 * int(d))
        seen = [set([int(a)]) for a in cells[:, 0]] if d > 12 else None
        for j in range(1, int(d)):
            cand = (anchors[idx] + rng.integers(-win, win + 1, size=k)) % C
            # resolve duplicates against earlier columns by linear probing
            # (at most 64 steps of +1 per pin)
            if seen is None:
                for _ in range(64):
                    dup = np.zeros(k, bool)
                    for jj in range(j):
                        dup |= cand == cells[:, jj]
                    if not dup.any():
                        break
                    cand[dup] = (cand[dup] + 1) % C
            else:   # the same probing per net with a set (high degrees)
                for r in range(k):
                    c, s = int(cand[r]), seen[r]
                    for _ in range(64):
                        if c not in s:
                            break
                        c = (c + 1) % C
                    cand[r] = c
                    s.add(c)
            cells[:, j] = cand
        starts = net_start[idx]
        pos = starts[:, None] + np.arange(d)[None, :]
        pin_node[pos.ravel()] = cells.ravel()

    # some multi-pin nets also touch a macro pin (replace their last pin)
    if F and macro_pin_prob > 0 and E > n2:
        multi = np.arange(n2, E)
        pick = multi[rng.random(multi.shape[0]) < macro_pin_prob]
        last = net_start[pick + 1] - 1
        pin_node[last] = C + rng.integers(0, F, size=pick.shape[0])

    # pin offsets inside the owning instance, integer DBU
    pw = w[pin_node].astype(np.float64)
    ph = h[pin_node].astype(np.float64)
    pin_offset_x = np.floor(rng.random(P) * pw).astype(np.float32)
    pin_offset_y = np.floor(rng.random(P) * ph).astype(np.float32)

    # pins are numbered in net order, so net_pins is 0..P-1
    net_pins = np.arange(P, dtype=np.int64)

    # macro_2_pins CSR: pins of each instance in increasing pin id
    N = C + F
    counts = np.bincount(pin_node, minlength=N)
    node_pin_start = np.zeros(N + 1, np.int64)
    np.cumsum(counts, out=node_pin_start[1:])
    node_pins = np.argsort(pin_node, kind="stable")

    n_2nets = int((degrees[:n2] == 2).sum())
    assert n_2nets == n2 or C < 2

    params = dict(num_cells=C, num_nets=E, degree_cap=degree_cap, num_macros=F,
                  macro_frac=macro_frac, utilization=utilization, tail_alpha=tail_alpha,
                  macro_aspect=list(macro_aspect), locality=locality,
                  macro_pin_prob=macro_pin_prob, site_w=SITE_W, row_h=ROW_H)
    return Design(
        name=name, seed=seed, die_w=die, die_h=die, num_movable=C,
        x=x, y=y, w=w, h=h, num_pins=counts.astype(np.float32),
        pin_offset_x=pin_offset_x, pin_offset_y=pin_offset_y,
        pin_node=pin_node.astype(np.int32),
        net_start=net_start.astype(np.int32), net_pins=net_pins.astype(np.int32),
        n_2nets=n_2nets,
        node_pin_start=node_pin_start.astype(np.int32), node_pins=node_pins.astype(np.int32),
        movable_area=movable_area,
        cell_utilization=movable_area / (die * die - fixed_area),
        params=params,
    )
