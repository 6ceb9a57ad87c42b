"""Pins the oracle's spectral restatement (dct2d.rs, idct2.rs, idct_idxst.rs,
idxst_idct.rs, irfft2d.rs, utils.rs:232-249, frequencies.rs) to closed forms:
scipy's DCT-II/III, dense sine/cosine products, numpy's FFTs, and hand-derived
special values.  Stage tests follow the reference's own per-stage tests
(dct2d.rs:351-477 etc.), with the DREAMPlace fixtures replaced by these pins.
"""
import math

import numpy as np
import pytest
import scipy.fft as sf

import oracle_lib as O

SIZES = [(8, 16), (16, 8), (32, 32), (64, 128)]


def _mats(m, n):
    i = np.arange(m)[:, None]
    u = np.arange(m)[None, :]
    j = np.arange(n)[:, None]
    v = np.arange(n)[None, :]
    CM = np.cos(np.pi * (2 * i + 1) * u / (2 * m))
    SM = np.sin(np.pi * (2 * i + 1) * u / (2 * m))
    CN = np.cos(np.pi * (2 * j + 1) * v / (2 * n))
    SN = np.sin(np.pi * (2 * j + 1) * v / (2 * n))
    em = np.ones(m); em[0] = 0.5
    en = np.ones(n); en[0] = 0.5
    return CM, SM, CN, SN, em, en


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_make_expk_closed_form():
    """utils.rs:232-249: e^{-i pi k / 2n}; the reference uses f32 pi (utils.rs:3)."""
    n = 512
    k = np.arange(n)
    exact = O.make_expk(n, O.PI_EXACT)
    assert np.abs(exact - np.exp(-1j * np.pi * k / (2 * n))).max() < 1e-15
    f32pi = O.make_expk(n, O.PI_F32)
    pi32 = float(np.float32(np.pi))
    assert np.abs(f32pi - np.exp(-1j * pi32 * k / (2 * n))).max() < 1e-15
    # the f32-pi twiddles differ from exact ones by ~|pi - pi32|, well inside 1e-6 (dct2d.rs:366)
    assert 1e-9 < np.abs(f32pi - exact).max() < 1e-6


def test_frequency_matrices_closed_form():
    """frequencies.rs:4-72 with hand values: (0,0) -> 0; (u,0) -> 1/wu^2 etc."""
    m, n, bw, bh = 16, 32, 3.0, 2.0
    wuh, wvh, inv = O.frequency_matrices(m, n, bw, bh)
    wu = 2 * np.pi * np.arange(m) / m
    wv = 2 * np.pi * np.arange(n) / n * (bw / bh)
    d = wu[:, None] ** 2 + wv[None, :] ** 2
    d[0, 0] = 1.0
    ref_inv = 1.0 / d
    ref_inv[0, 0] = 0.0
    assert _rel(inv, ref_inv) < 1e-14
    assert inv[0, 0] == 0 and wuh[0, 0] == 0 and wvh[0, 0] == 0
    assert math.isclose(inv[1, 0], 1.0 / (2 * np.pi / m) ** 2, rel_tol=1e-14)
    assert math.isclose(wuh[1, 0], 0.5 / (2 * np.pi / m), rel_tol=1e-14)
    assert math.isclose(wvh[0, 1], 0.5 / (2 * np.pi / n * bw / bh), rel_tol=1e-14)


@pytest.mark.parametrize("n", [2, 8, 64, 1024])
def test_fft_core_against_numpy(n):
    rng = np.random.default_rng(n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.abs(O.fft(z) - np.fft.fft(z)).max() < 1e-12 * n
    assert np.abs(O.fft(z, inverse=True) - np.fft.ifft(z) * n).max() < 1e-12 * n


def test_dct2d_preprocess_index_map():
    """dct2d.rs:47-81: rows and columns reordered v[p]=x[2p], v[L-1-p]=x[2p+1]."""
    m, n = 8, 16
    x = np.arange(m * n, dtype=np.float64).reshape(m, n)
    got = O.dct2d_preprocess(x)
    def perm(L):
        p = np.empty(L, int)
        p[: L // 2] = np.arange(0, L, 2)
        p[L // 2:] = np.arange(L - 1, 0, -2)
        return p
    assert np.array_equal(got, x[perm(m)][:, perm(n)])


@pytest.mark.parametrize("m,n", SIZES)
def test_dct2_fft2_forward(m, n):
    """dct2d.rs:256-279 = scipy dctn type 2 / (m n)."""
    x = np.random.default_rng(m * n).standard_normal((m, n))
    got = O.dct2_fft2_forward(x, pi_mode=O.PI_EXACT)
    assert _rel(got, sf.dctn(x, type=2) / (m * n)) < 1e-13


@pytest.mark.parametrize("m,n", SIZES)
def test_idct2_fft2_forward(m, n):
    """idct2.rs:144-166 = scipy dctn type 3 = 4 C_M (eps eps X) C_N^T."""
    x = np.random.default_rng(m + n).standard_normal((m, n))
    got = O.idct2_fft2_forward(x, pi_mode=O.PI_EXACT)
    assert _rel(got, sf.dctn(x, type=3)) < 1e-12


@pytest.mark.parametrize("m,n", SIZES)
def test_idxst_idct_forward(m, n):
    """idxst_idct.rs:156-180 = 4 S_M (X diag eps_v) C_N^T (sine along x)."""
    x = np.random.default_rng(3 * m + n).standard_normal((m, n))
    CM, SM, CN, SN, em, en = _mats(m, n)
    got = O.idxst_idct_forward(x, pi_mode=O.PI_EXACT)
    assert _rel(got, 4 * SM @ (x * en[None, :]) @ CN.T) < 1e-12


@pytest.mark.parametrize("m,n", SIZES)
def test_idct_idxst_forward(m, n):
    """idct_idxst.rs:137-161 = 4 C_M (diag eps_u X) S_N^T (sine along y)."""
    x = np.random.default_rng(5 * m + n).standard_normal((m, n))
    CM, SM, CN, SN, em, en = _mats(m, n)
    got = O.idct_idxst_forward(x, pi_mode=O.PI_EXACT)
    assert _rel(got, 4 * CM @ (em[:, None] * x) @ SN.T) < 1e-12


def test_f32_pi_twiddles_stay_within_reference_tolerance():
    """With the reference's f32 pi the transforms move by < 1e-6 relative."""
    x = np.random.default_rng(1).standard_normal((64, 64))
    for name in ("dct2_fft2_forward", "idct2_fft2_forward", "idct_idxst_forward", "idxst_idct_forward"):
        a = getattr(O, name)(x, pi_mode=O.PI_F32)
        b = getattr(O, name)(x, pi_mode=O.PI_EXACT)
        assert 0 < _rel(a, b) < 1e-6, name


def test_idct2_inverts_dct2():
    x = np.random.default_rng(2).standard_normal((32, 64))
    y = O.idct2_fft2_forward(O.dct2_fft2_forward(x, pi_mode=O.PI_EXACT), pi_mode=O.PI_EXACT)
    assert _rel(y, 4 * x) < 1e-13


def test_irfft_2d_matches_numpy():
    """irfft2d.rs:7-83 = numpy.fft.irfft2 (Hermitian rebuild, Nyquist imag dropped)."""
    m, n = 16, 32
    rng = np.random.default_rng(4)
    spec = rng.standard_normal((m, n // 2 + 1)) + 1j * rng.standard_normal((m, n // 2 + 1))
    got = O.irfft_2d(spec, n)
    assert _rel(got, np.fft.irfft2(spec, s=(m, n))) < 1e-13


def test_transforms_threaded_equal_serial():
    """Row/column parallel FFTs (dct2d.rs:294,318; irfft2d.rs:22,48) change nothing."""
    x = np.random.default_rng(6).standard_normal((128, 64))
    for name in ("dct2_fft2_forward", "idct_idxst_forward"):
        a = getattr(O, name)(x, nthreads=1)
        b = getattr(O, name)(x, nthreads=4)
        assert np.array_equal(a, b)


def test_plausible_mistakes_are_caught():
    """A transposed operand or a dropped eps_0 moves the result far outside the
    tolerances used above (so the pins above are not vacuous)."""
    m, n = 16, 16
    x = np.random.default_rng(8).standard_normal((m, n))
    CM, SM, CN, SN, em, en = _mats(m, n)
    got = O.idct_idxst_forward(x, pi_mode=O.PI_EXACT)
    assert _rel(got, 4 * CM @ x @ SN.T) > 1e-3             # eps_u dropped
    assert _rel(got, 4 * SM @ (x * en[None]) @ CN.T) > 1e-3  # sine on the wrong axis
