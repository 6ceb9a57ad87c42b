"""Readers for the reference's text fixtures (utils.rs:160-230).

The reference's stage tests (dct2d.rs:351-477, idct2.rs, idct_idxst.rs,
idxst_idct.rs, density.rs:333-514) read DREAMPlace dumps with
``read_matrix_from_file`` (one matrix row per line, whitespace-separated reals,
empty lines ignored) and ``read_complex_matrix_from_file`` (pairs re im per
complex entry, every row the same number of pairs).  The dumps are not shipped
with the reference; a directory holding them is found through the environment
variable ``ERP_REFERENCE_FIXTURES`` (the reference's ``tests/`` directory).
"""
from __future__ import annotations

import os

import numpy as np

ENV = "ERP_REFERENCE_FIXTURES"


def read_matrix(path: str, dtype=np.float64):
    """utils.rs:202-230: returns (flat data, n_rows, n_cols of the first row)."""
    data, n_rows, n_cols = [], 0, 0
    with open(path) as f:
        for line in f:
            vals = line.split()
            if not vals:
                continue
            if n_cols == 0:
                n_cols = len(vals)
            data.extend(float(v) for v in vals)
            n_rows += 1
    return np.asarray(data, dtype=dtype), n_rows, n_cols


def read_complex_matrix(path: str, dtype=np.complex128):
    """utils.rs:150-199: (re, im) pairs; an odd count or a ragged row is an error."""
    data, n_rows, n_cols = [], 0, 0
    with open(path) as f:
        for line in f:
            vals = line.split()
            if not vals:
                continue
            if len(vals) % 2:
                raise ValueError(f"line {n_rows + 1} has an odd number of values, expected real/imag pairs")
            row = [complex(float(vals[k]), float(vals[k + 1])) for k in range(0, len(vals), 2)]
            if n_cols == 0:
                n_cols = len(row)
            elif n_cols != len(row):
                raise ValueError(f"inconsistent number of complex columns in line {n_rows + 1}")
            data.extend(row)
            n_rows += 1
    return np.asarray(data, dtype=dtype), n_rows, n_cols


def fixture(rel: str) -> str | None:
    """Path of a reference fixture (relative to its tests/ directory), or None."""
    root = os.environ.get(ENV)
    if not root:
        return None
    p = os.path.join(root, rel)
    return p if os.path.exists(p) else None
