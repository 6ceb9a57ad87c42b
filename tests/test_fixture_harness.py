"""The reference's fixture-driven stage tests, re-expressed against the C ABI.

The DREAMPlace dumps these tests read (dct2d.rs:351-477, idct2.rs:178-281,
idct_idxst.rs:174-278, idxst_idct.rs:193-297, density.rs:333-514) are not
shipped with the reference, so each test is skipped unless
ERP_REFERENCE_FIXTURES names a directory that holds them (the reference's
tests/); tolerances are the reference's own.  The readers themselves are
checked here on files written by the test.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from erplacer import fixtures as FX


def _need(*rels):
    paths = [FX.fixture(r) for r in rels]
    if not all(paths):
        pytest.skip(f"reference fixtures not supplied ({FX.ENV} unset or missing {rels[0]})")
    return paths


# ----------------------------------------------------------------- readers (CPU)
def test_read_matrix_round_trip(tmp_path):
    a = np.arange(12, dtype=np.float64).reshape(3, 4) / 7.0
    p = tmp_path / "m.txt"
    with open(p, "w") as f:
        f.write("\n")
        for row in a:
            f.write("  ".join(f"{v:.17g}" for v in row) + "\n\n")
    data, r, c = FX.read_matrix(str(p))
    assert (r, c) == (3, 4) and np.array_equal(data, a.reshape(-1))


def test_read_complex_matrix_round_trip_and_errors(tmp_path):
    z = (np.arange(6) + 1j * np.arange(6)[::-1]).reshape(2, 3)
    p = tmp_path / "c.txt"
    with open(p, "w") as f:
        for row in z:
            f.write(" ".join(f"{v.real:.17g} {v.imag:.17g}" for v in row) + "\n")
    data, r, c = FX.read_complex_matrix(str(p))
    assert (r, c) == (2, 3) and np.array_equal(data, z.reshape(-1))
    p.write_text("1 2 3\n")
    with pytest.raises(ValueError, match="odd number"):
        FX.read_complex_matrix(str(p))
    p.write_text("1 2 3 4\n1 2\n")
    with pytest.raises(ValueError, match="inconsistent"):
        FX.read_complex_matrix(str(p))


def test_fixture_lookup_without_environment(monkeypatch):
    monkeypatch.delenv(FX.ENV, raising=False)
    assert FX.fixture("fft2/expk_m.txt") is None


# ------------------------------------------------------- oracle stage (CPU)
def test_make_expk_512x512():
    """dct2d.rs:351-373 against the oracle's make_expk (f32 pi, utils.rs:232-249)."""
    pm, pn = _need("fft2/expk_m.txt", "fft2/expk_n.txt")
    for p in (pm, pn):
        ref, L, _ = FX.read_complex_matrix(p)
        got = O.make_expk(L)
        assert np.max(np.abs(got[:ref.size] - ref)) < 1e-6


# ------------------------------------------------------------ GPU stage tests
@pytest.mark.gpu
@pytest.mark.parametrize("name,inp,out,tol", [
    ("dct2_fft2_forward", "fft2/dct2/dct2_input_512x512.txt", "fft2/dct2/dct2_post_512x512.txt", 1e-5),
    ("idct2_fft2_forward", "fft2/idct2/idct2_input_512x512.txt", "fft2/idct2/idct2_post_512x512.txt", 1e-2),
    ("idct_idxst_forward", "fft2/idct_idxst/idct_idxst_input_512x512.txt",
     "fft2/idct_idxst/idct_idxst_post_512x512.txt", 1e-3),
    ("idxst_idct_forward", "fft2/idxst_idct/idxst_idct_input_512x512.txt",
     "fft2/idxst_idct/idxst_idct_post_512x512.txt", 1e-3),
])
def test_transform_512x512(gpu_lib, name, inp, out, tol):
    from erplacer import placement as P
    pi, po = _need(inp, out)
    x, m, n = FX.read_matrix(pi)
    ref, _, _ = FX.read_matrix(po)
    got = getattr(P, name)(x.reshape(m, n).astype(np.float32)).reshape(-1)
    assert np.max(np.abs(got - ref)) < tol


def _scale_for(mass):
    e = math.frexp((1 << 62) / (mass * 1.01 + 1.0))[1]
    return math.ldexp(1.0, e - 1)


@pytest.mark.gpu
def test_density_map_72090():
    """density.rs:393-450 (sequential) / :452-514 (parallel): both compare to the
    same dump with tolerance 1e-1."""
    from erplacer import placement as P
    names = ["density_x", "density_y", "density_node_size_x", "density_node_size_y",
             "density_offset_x", "density_offset_y", "density_ratios", "density_out",
             "initial_density_map"]
    paths = _need(*[f"density/{k}.txt" for k in names])
    arr = {k: FX.read_matrix(p)[0] for k, p in zip(names, paths)}
    ref, m, n = FX.read_matrix(paths[7])
    bin_w = (1.7396e+06 - 6000.) / 200. / 512.
    bin_h = (1.1724e+06 - 6000.) / 200. / 512.
    mass = float(np.sum(arr["density_ratios"] * arr["density_node_size_x"] * arr["density_node_size_y"])
                 + np.sum(np.abs(arr["initial_density_map"])) * bin_w * bin_h)
    got, _ = P.compute_density_map(arr["density_x"], arr["density_y"], arr["density_node_size_x"],
                                   arr["density_node_size_y"], arr["density_offset_x"],
                                   arr["density_offset_y"], arr["density_ratios"], m, n, bin_w, bin_h,
                                   _scale_for(mass), initial=arr["initial_density_map"].reshape(m, n))
    assert np.max(np.abs(got.reshape(-1) - ref)) < 1e-1


@pytest.mark.gpu
def test_initial_density_map_4_nodes():
    """density.rs:334-390 (tolerance 1e-2); bin centres from the dump's grid."""
    from erplacer import placement as P
    names = ["initial_density_x", "initial_density_y", "initial_density_node_size_x",
             "initial_density_node_size_y", "initial_density_out"]
    paths = _need(*[f"density/{k}.txt" for k in names])
    x, y, sx, sy = (FX.read_matrix(p)[0] for p in paths[:4])
    ref, m, n = FX.read_matrix(paths[4])
    bin_w = (1.7396e+06 - 6000.) / 200. / 512.
    bin_h = (1.1724e+06 - 6000.) / 200. / 512.
    mass = float(np.sum(sx * sy)) + 1.0
    got, _ = P.compute_initial_density_map(x, y, sx, sy, m, n, bin_w, bin_h, bin_w * m, bin_h * n,
                                           _scale_for(mass))
    assert np.max(np.abs(got.reshape(-1) - ref)) < 1e-2
