"""LEF/DEF reader into the placement data contract (``netlist.Design``).

Stands for ``reda_db::read_db(lef, def)`` (main.rs:68-70), the step upstream of
``Placement::new``.  reda-db is not vendored with the reference, so its exact
conventions are unknown here; this reader follows the LEF/DEF 5.8 standard and
fills exactly the fields the placement loop reads (placement.rs:52-134,
hpwl.rs:25-45, epotential.rs:69-95):

* instances: movable components first (status PLACED or UNPLACED, in DEF
  order), then fixed ones (FIXED, COVER), then the design's IO pins as
  zero-size fixed instances at their placed location (drawing.rs:43-46 skips
  such zero-size instances when drawing);
* sizes: the master's SIZE in DEF units, width and height swapped for the
  rotated orientations (W, E, FW, FE);
* pins: one per net connection, at the centre of the bounding box of the
  master pin's shapes (all RECT and POLYGON points of all ports, shifted by the
  macro ORIGIN), transformed by the component orientation, as an offset from
  the placed lower-left corner; IO pins have offset 0;
* nets: connections in DEF order, two-pin nets first (the contract's
  ``n_2nets`` prefix); ``*`` connections and nets left with no pin are dropped;
* coordinates are shifted so that the DIEAREA bounding box starts at (0, 0)
  (the reference assumes a die origin of 0, density.rs:259-260); UNPLACED
  components sit at (0, 0) (movable cells start at the die centre anyway,
  placement.rs:129-130).

Units: LEF lengths are microns and are scaled by the DEF ``UNITS DISTANCE
MICRONS`` factor; sizes are rounded to whole DEF units and pin centres to half
units.  ``movable_area`` and ``cell_utilization`` are computed as in
``netlist.from_nets`` (movable area over die area minus fixed area).
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field

import numpy as np

from .netlist import Design, from_nets

_TOKEN = re.compile(r'"[^"]*"|;|\(|\)|[^\s;()]+')


def tokenize(text: str) -> list[str]:
    """Whitespace/';'/'(' ')' tokens of a LEF or DEF text, '#' comments removed."""
    text = re.sub(r"#[^\n]*", "", text)
    return _TOKEN.findall(text)


@dataclass
class Macro:
    name: str
    width: float = 0.0                 # microns
    height: float = 0.0
    origin: tuple = (0.0, 0.0)
    pins: dict = field(default_factory=dict)   # pin name -> (cx, cy) microns, bbox frame


def _skip_statement(tok, i):
    while i < len(tok) and tok[i] != ";":
        i += 1
    return i + 1


def parse_lef(text: str, macros: dict | None = None) -> dict:
    """MACRO definitions of a LEF text: name -> Macro (SIZE, ORIGIN, pin shape
    centres).  Layers, sites, vias and obstructions are skipped."""
    macros = {} if macros is None else macros
    tok = tokenize(text)
    i, n = 0, len(tok)
    while i < n:
        t = tok[i].upper()
        if t == "MACRO":
            name = tok[i + 1]
            m = Macro(name)
            i += 2
            while i < n and not (tok[i].upper() == "END" and i + 1 < n and tok[i + 1] == name):
                u = tok[i].upper()
                if u == "SIZE":                          # SIZE w BY h ;
                    m.width, m.height = float(tok[i + 1]), float(tok[i + 3])
                    i = _skip_statement(tok, i)
                elif u == "ORIGIN":
                    m.origin = (float(tok[i + 1]), float(tok[i + 2]))
                    i = _skip_statement(tok, i)
                elif u == "PIN":
                    pname = tok[i + 1]
                    i += 2
                    xs, ys = [], []
                    while i < n and not (tok[i].upper() == "END" and i + 1 < n and tok[i + 1] == pname):
                        u2 = tok[i].upper()
                        if u2 == "RECT":                 # RECT [MASK k] x1 y1 x2 y2 ;
                            j = i + 1
                            if tok[j].upper() == "MASK":
                                j += 2
                            x1, y1, x2, y2 = (float(v) for v in tok[j:j + 4])
                            xs += [x1, x2]
                            ys += [y1, y2]
                            i = _skip_statement(tok, i)
                        elif u2 == "POLYGON":            # POLYGON [MASK k] x y x y ... ;
                            j = i + 1
                            if tok[j].upper() == "MASK":
                                j += 2
                            k = _skip_statement(tok, i) - 1
                            vals = [float(v) for v in tok[j:k]]
                            xs += vals[0::2]
                            ys += vals[1::2]
                            i = k + 1
                        else:
                            i += 1
                    if xs:
                        cx = 0.5 * (min(xs) + max(xs))
                        cy = 0.5 * (min(ys) + max(ys))
                        m.pins[pname] = (cx, cy)
                    else:
                        m.pins[pname] = (0.0, 0.0)
                    i += 2
                elif u == "OBS":
                    while i < n and tok[i].upper() != "END":
                        i += 1
                    i += 1
                else:
                    i += 1
            macros[name] = m
            i += 2
        else:
            i += 1
    return macros


# DEF orientation -> point transform inside the placed bounding box:
# (px, py) in the master frame (size w x h) -> offset from the placed lower-left
_ORIENT = {
    "N": lambda px, py, w, h: (px, py),
    "S": lambda px, py, w, h: (w - px, h - py),
    "W": lambda px, py, w, h: (h - py, px),
    "E": lambda px, py, w, h: (py, w - px),
    "FN": lambda px, py, w, h: (w - px, py),
    "FS": lambda px, py, w, h: (px, h - py),
    "FW": lambda px, py, w, h: (py, px),
    "FE": lambda px, py, w, h: (h - py, w - px),
}
_ROTATED = {"W", "E", "FW", "FE"}


def pin_offset(orient: str, px: float, py: float, w: float, h: float):
    """Offset of a master-frame point after placing the master with a DEF
    orientation (LEF/DEF 5.8: W = R90, E = R270, FN = mirror about the y axis,
    FS = mirror about the x axis, FW = MX then R90, FE = MY then R90)."""
    try:
        return _ORIENT[orient.upper()](px, py, w, h)
    except KeyError:
        raise ValueError(f"unknown DEF orientation {orient!r}") from None


@dataclass
class _Comp:
    name: str
    macro: str
    status: str = "UNPLACED"
    x: float = 0.0
    y: float = 0.0
    orient: str = "N"


def _parse_placement(tok, i):
    """'( x y ) orient' starting at tok[i] == '('."""
    x, y = float(tok[i + 1]), float(tok[i + 2])
    return x, y, tok[i + 4], i + 5


def parse_def(text: str):
    """The sections of a DEF text the placement needs: units, die area,
    components, IO pins and nets (connections only)."""
    tok = tokenize(text)
    i, n = 0, len(tok)
    units, die, comps, iopins, nets = 1.0, None, [], [], []
    while i < n:
        t = tok[i].upper()
        if t == "UNITS" and i + 3 < n and tok[i + 1].upper() == "DISTANCE":
            units = float(tok[i + 3])
            i = _skip_statement(tok, i)
        elif t == "DIEAREA":
            k = _skip_statement(tok, i) - 1
            pts = [float(v) for v in tok[i + 1:k] if v not in ("(", ")")]
            xs, ys = pts[0::2], pts[1::2]
            die = (min(xs), min(ys), max(xs), max(ys))
            i = k + 1
        elif t == "COMPONENTS":
            i = _skip_statement(tok, i)
            while i < n and tok[i].upper() != "END":
                if tok[i] != "-":
                    i += 1
                    continue
                c = _Comp(tok[i + 1], tok[i + 2])
                i += 3
                while tok[i] != ";":
                    if tok[i] == "+" and tok[i + 1].upper() in ("PLACED", "FIXED", "COVER"):
                        c.status = tok[i + 1].upper()
                        c.x, c.y, c.orient, i = _parse_placement(tok, i + 2)
                    elif tok[i] == "+" and tok[i + 1].upper() == "UNPLACED":
                        c.status = "UNPLACED"
                        i += 2
                    else:
                        i += 1
                comps.append(c)
                i += 1
            i += 2                                   # END COMPONENTS
        elif t == "PINS":
            i = _skip_statement(tok, i)
            while i < n and tok[i].upper() != "END":
                if tok[i] != "-":
                    i += 1
                    continue
                name, x, y, placed = tok[i + 1], 0.0, 0.0, False
                i += 2
                while tok[i] != ";":
                    if (not placed and tok[i] == "+"
                            and tok[i + 1].upper() in ("PLACED", "FIXED", "COVER")):
                        x, y, _, i = _parse_placement(tok, i + 2)
                        placed = True
                    else:
                        i += 1
                iopins.append((name, x, y))
                i += 1
            i += 2
        elif t == "NETS" and i + 1 < n and tok[i - 1].upper() != "END":
            i = _skip_statement(tok, i)
            while i < n and tok[i].upper() != "END":
                if tok[i] != "-":
                    i += 1
                    continue
                name = tok[i + 1]
                i += 2
                conns = []
                while tok[i] not in (";", "+"):
                    if tok[i] == "(":
                        conns.append((tok[i + 1], tok[i + 2]))
                        i += 4
                    else:
                        i += 1
                i = _skip_statement(tok, i)
                nets.append((name, conns))
            i += 2
        elif t == "SPECIALNETS":
            while i < n and not (tok[i].upper() == "END" and tok[i + 1].upper() == "SPECIALNETS"):
                i += 1
            i += 2
        else:
            i += 1
    if die is None:
        raise ValueError("DEF has no DIEAREA")
    return units, die, comps, iopins, nets


def read_db(lef_texts, def_text: str, name: str = "lefdef") -> Design:
    """A ``Design`` from LEF texts (a string or a list of strings) and a DEF text."""
    if isinstance(lef_texts, str):
        lef_texts = [lef_texts]
    macros = {}
    for t in lef_texts:
        parse_lef(t, macros)
    units, (dx0, dy0, dx1, dy1), comps, iopins, nets = parse_def(def_text)

    def snap(v, step):
        return float(np.round(v / step) * step)

    movable = [c for c in comps if c.status in ("PLACED", "UNPLACED")]
    fixed = [c for c in comps if c.status in ("FIXED", "COVER")]
    order = movable + fixed
    index, x, y, w, h, orient = {}, [], [], [], [], []
    for k, c in enumerate(order):
        if c.macro not in macros:
            raise ValueError(f"component {c.name}: macro {c.macro!r} not in the LEF")
        m = macros[c.macro]
        mw, mh = snap(m.width * units, 1.0), snap(m.height * units, 1.0)
        if c.orient.upper() in _ROTATED:
            mw, mh = mh, mw
        index[("C", c.name)] = k
        placed = c.status != "UNPLACED"            # unplaced: the die's lower-left
        x.append(c.x - dx0 if placed else 0.0)
        y.append(c.y - dy0 if placed else 0.0)
        w.append(mw)
        h.append(mh)
        orient.append(c.orient)
    for p, (pname, px, py) in enumerate(iopins):
        index[("P", pname)] = len(order) + p
        x.append(px - dx0)
        y.append(py - dy0)
        w.append(0.0)
        h.append(0.0)
        orient.append("N")
    comp_of = {c.name: c for c in comps}
    net_list = []
    for _, conns in nets:
        pins = []
        for a, b in conns:
            if a == "*":
                continue
            if a.upper() == "PIN":
                pins.append((index[("P", b)], 0.0, 0.0))
                continue
            c = comp_of.get(a)
            if c is None:
                raise ValueError(f"net connection to unknown component {a!r}")
            m = macros[c.macro]
            if b not in m.pins:
                raise ValueError(f"macro {m.name} has no pin {b!r}")
            cx, cy = m.pins[b]
            cx, cy = (cx + m.origin[0]) * units, (cy + m.origin[1]) * units
            ox, oy = pin_offset(c.orient, cx, cy, snap(m.width * units, 1.0),
                                snap(m.height * units, 1.0))
            pins.append((index[("C", a)], snap(ox, 0.5), snap(oy, 0.5)))
        if pins:
            net_list.append(pins)
    return from_nets(np.asarray(x, np.float32), np.asarray(y, np.float32),
                     np.asarray(w, np.float32), np.asarray(h, np.float32), len(movable),
                     net_list, dx1 - dx0, dy1 - dy0, name=name)


def read_files(lef_paths, def_path: str) -> Design:
    """``read_db`` over files (LEF paths: one or a list)."""
    if isinstance(lef_paths, str):
        lef_paths = [lef_paths]
    lefs = [open(p).read() for p in lef_paths]
    with open(def_path) as f:
        return read_db(lefs, f.read(), name=def_path)
