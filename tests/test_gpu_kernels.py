"""Per-kernel parity of the HIP leaf functions against the CPU oracle.

Mirrors the reference's stage tests (dct2d.rs:351-477, idct2.rs:178-281,
idct_idxst.rs:174-278, idxst_idct.rs:193-297, density.rs:333-514): same
functions, fixtures replaced by seeded inputs checked against the FP64 oracle
(and, for the fixed-point density map, bit for bit against the oracle's CPU
restatement of the fixed-point accumulation).
"""
import numpy as np
import pytest

import oracle_lib as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu

# FP32 transforms vs FP64: the relative L2 bound the field and gradient are held to.
TOL_SPECTRAL = 1e-5


def tol_overlap(m):
    """f32 coordinates (the reference binary's precision, main.rs:69) resolve an
    overlap of ~1 bin to ~ulp(die)/bin = m * 6e-8 relative; contributions of
    clustered cells share that error, so map-level comparisons against FP64
    overlaps use m * 1e-7.  Bit-exactness is checked against the f32 restatement."""
    return 1e-7 * m


def _map(rng, m, n):
    # density-like input: positive with structure plus noise
    i = np.linspace(0, 1, m)[:, None]
    j = np.linspace(0, 1, n)[None, :]
    return (np.exp(-((i - 0.4) ** 2 + (j - 0.6) ** 2) * 8) + 0.1 * rng.random((m, n))).astype(np.float32)


@pytest.mark.parametrize("m,n", [(64, 64), (128, 256), (256, 128), (512, 512), (1024, 1024)])
def test_dct2_fft2_forward(gpu_lib, m, n):
    from erplacer import placement as P
    x = _map(np.random.default_rng(m + n), m, n)
    got = P.dct2_fft2_forward(x)
    ref = O.dct2_fft2_forward(x.astype(np.float64))
    assert rel_l2(got, ref) < TOL_SPECTRAL


@pytest.mark.parametrize("name", ["idct2_fft2_forward", "idct_idxst_forward", "idxst_idct_forward"])
@pytest.mark.parametrize("m,n", [(64, 64), (128, 256), (512, 512), (2048, 2048)])
def test_inverse_transforms(gpu_lib, name, m, n):
    from erplacer import placement as P
    rng = np.random.default_rng(7 * m + n)
    x = (rng.standard_normal((m, n)) / (1 + np.add.outer(np.arange(m), np.arange(n)))).astype(np.float32)
    got = getattr(P, name)(x)
    ref = getattr(O, name)(x.astype(np.float64))
    assert rel_l2(got, ref) < TOL_SPECTRAL, name


# Split-BF16 (hi*hi + hi*lo + lo*hi) MFMA products with FP32 accumulation:
# ~2^-17 relative per term; at 2048^2 the relative L2 error stays below 1e-5.
TOL_MFMA = 1e-5


@pytest.mark.parametrize("m,n", [(128, 128), (128, 256), (256, 128), (512, 512), (1024, 1024),
                                 (2048, 2048)])
def test_dct2_mfma(gpu_lib, m, n):
    from erplacer import lib as L
    from erplacer import placement as P
    x = _map(np.random.default_rng(m + n), m, n)
    got = P.dct2_fft2_forward(x, mode=L.SPECTRAL_MFMA_BF16X3)
    ref = O.dct2_fft2_forward(x.astype(np.float64))
    err = rel_l2(got, ref)
    print(f"mfma dct2 {m}x{n}: rel L2 {err:.2e}")
    assert err < TOL_MFMA


@pytest.mark.parametrize("name", ["idct2_fft2_forward", "idct_idxst_forward", "idxst_idct_forward"])
@pytest.mark.parametrize("m,n", [(128, 128), (256, 512), (512, 256), (1024, 1024), (2048, 2048)])
def test_inverse_mfma(gpu_lib, name, m, n):
    from erplacer import lib as L
    from erplacer import placement as P
    rng = np.random.default_rng(7 * m + n)
    x = (rng.standard_normal((m, n)) / (1 + np.add.outer(np.arange(m), np.arange(n)))).astype(np.float32)
    got = getattr(P, name)(x, mode=L.SPECTRAL_MFMA_BF16X3)
    ref = getattr(O, name)(x.astype(np.float64))
    err = rel_l2(got, ref)
    print(f"mfma {name} {m}x{n}: rel L2 {err:.2e}")
    assert err < TOL_MFMA, name


def test_mfma_rejects_small_grids(gpu_lib):
    from erplacer import lib as L
    from erplacer import placement as P
    with pytest.raises(L.ErpError):
        P.dct2_fft2_forward(np.zeros((64, 64), np.float32), mode=L.SPECTRAL_MFMA_BF16X3)


def test_idct2_of_dct2_is_4x(gpu_lib):
    from erplacer import placement as P
    x = _map(np.random.default_rng(3), 256, 256)
    y = P.idct2_fft2_forward(P.dct2_fft2_forward(x))
    assert rel_l2(y, 4.0 * x) < TOL_SPECTRAL


def _cells(rng, nc, die, bw, bh, clustered=False):
    w = rng.integers(1, 9, nc).astype(np.float32) * 200
    h = np.full(nc, 2000, np.float32)
    sq = np.float32(np.sqrt(np.float32(2.0)))
    cw = np.maximum(w, np.float32(bw) * sq).astype(np.float32)
    ch = np.maximum(h, np.float32(bh) * sq).astype(np.float32)
    ox = ((w - cw) * np.float32(0.5)).astype(np.float32)
    oy = ((h - ch) * np.float32(0.5)).astype(np.float32)
    ratio = ((w * h) / (cw * ch)).astype(np.float32)
    if clustered:
        x = (np.float32(die / 2) - w / np.float32(2)).astype(np.float32)
        y = (np.float32(die / 2) - h / np.float32(2)).astype(np.float32)
    else:
        x = rng.uniform(0, die - w - 1).astype(np.float32)
        y = rng.uniform(0, die - h - 1).astype(np.float32)
    return x, y, cw, ch, ox, oy, ratio


@pytest.mark.parametrize("clustered", [False, True])
@pytest.mark.parametrize("m", [64, 512])
def test_density_map(gpu_lib, m, clustered):
    """Movable scatter: int64 accumulators bit-identical to the CPU fixed-point
    restatement; float map within FP32 rounding of the FP64 sequential oracle
    (density.rs:174-237)."""
    from erplacer import placement as P
    rng = np.random.default_rng(m)
    die = 400_000.0
    bw = bh = np.float32(die) / np.float32(m)
    nc = 20_000
    x, y, cw, ch, ox, oy, ratio = _cells(rng, nc, die, bw, bh, clustered)
    scale = O.fx_scale(float((cw * ch * ratio).sum()) * 1.01)
    initial = (rng.random((m, m)) * float(bw * bh) * 0.3).astype(np.float32)
    got, q = P.compute_density_map(x, y, cw, ch, ox, oy, ratio, m, m, bw, bh, scale, initial)
    q_ref = O.density_map_fx64(x, y, cw, ch, ox, oy, ratio, m, m, bw, bh, scale, sfx="_f32")
    assert np.array_equal(q, q_ref)
    ref = O.density_map("sequential", x, y, cw, ch, ox, oy, ratio, m, m, float(bw), float(bh),
                        initial.astype(np.float64))
    assert rel_l2(got, ref) < tol_overlap(m)
    # the last row / column never receive movable density (density.rs:209,218)
    mov = got - initial
    assert np.all(mov[m - 1, :] == 0) and np.all(mov[:, m - 1] == 0)


def test_density_map_order_independent(gpu_lib):
    """Permuting the cells leaves the int64 map bit-identical (the reference's
    Mutex path, density.rs:158, is order-dependent)."""
    from erplacer import placement as P
    rng = np.random.default_rng(11)
    m, die = 128, 300_000.0
    bw = np.float32(die) / np.float32(m)
    arrs = _cells(rng, 30_000, die, bw, bw)
    scale = O.fx_scale(float((arrs[2] * arrs[3] * arrs[6]).sum()) * 1.01)
    _, q1 = P.compute_density_map(*arrs, m, m, bw, bw, scale)
    perm = rng.permutation(arrs[0].shape[0])
    _, q2 = P.compute_density_map(*[a[perm] for a in arrs], m, m, bw, bw, scale)
    assert np.array_equal(q1, q2)


def test_initial_density_map(gpu_lib):
    """Fixed cells against bin centres (density.rs:26-99)."""
    from erplacer import placement as P
    rng = np.random.default_rng(5)
    m, die = 256, 500_000.0
    bw = np.float32(die) / np.float32(m)
    nf = 40
    x = rng.uniform(0, die * 0.8, nf).astype(np.float32)
    y = rng.uniform(0, die * 0.8, nf).astype(np.float32)
    cw = rng.uniform(1000, die * 0.15, nf).astype(np.float32)
    ch = rng.uniform(1000, die * 0.15, nf).astype(np.float32)
    scale = O.fx_scale(float((cw * ch).sum()) * 1.01)
    got, q = P.compute_initial_density_map(x, y, cw, ch, m, m, bw, bw, die, die, scale)
    q_ref = O.initial_density_fx64(x, y, cw, ch, m, m, bw, bw, die, die, scale, sfx="_f32")
    assert np.array_equal(q, q_ref)
    ref = O.initial_density_map(x, y, cw, ch, m, m, float(bw), float(bw), die, die)
    assert rel_l2(got, ref) < tol_overlap(m)


def test_compute_electric_force(gpu_lib):
    """eforce.rs:15-91 against the FP64 oracle on a realistic field."""
    from erplacer import placement as P
    rng = np.random.default_rng(9)
    m, die = 256, 400_000.0
    bw = np.float32(die) / np.float32(m)
    rho = _map(rng, m, m).astype(np.float64)
    a = O.dct2_fft2_forward(rho)
    wu, wv, inv = O.frequency_matrices(m, m, float(bw), float(bw))
    fx = O.idxst_idct_forward(a * wu).astype(np.float32)
    fy = O.idct_idxst_forward(a * wv).astype(np.float32)
    x, y, cw, ch, ox, oy, ratio = _cells(rng, 50_000, die, bw, bw)
    gx, gy = P.compute_electric_force(fx, fy, x, y, cw, ch, ox, oy, ratio, bw, bw)
    rx, ry = O.electric_force(fx.astype(np.float64), fy.astype(np.float64), x, y, cw, ch, ox, oy,
                              ratio, float(bw), float(bw))
    assert rel_l2(gx, rx) < max(1e-5, tol_overlap(m)) and rel_l2(gy, ry) < max(1e-5, tol_overlap(m))
