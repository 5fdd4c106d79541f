"""Pins of the oracle's transforms (the checker of the HIP transforms).

Known answers restate the reference's transform tests (tests/test_xyzfft.cpp:77-125,
201-274, 323-373) and add a brute-force evaluation of the truncated series.
"""
import math

import numpy as np
import pytest

from cdns import config
from helpers import band_spectrum, maxrel, ref_transform_config
import pyoracle

PI = math.pi


def dct1_definition(x):
    """REDFT00: Y_k = x_0 + (-1)^k x_{n-1} + 2 sum_{j=1}^{n-2} x_j cos(pi j k / (n-1))."""
    n = len(x)
    ny = n - 1
    k = np.arange(n)
    j = np.arange(1, ny)
    return x[0] + ((-1.0) ** k) * x[ny] + 2 * (np.cos(PI * np.outer(k, j) / ny) @ x[1:ny])


def y_spectrum_expected(my, NY, NYS):
    f = np.cos(my * PI * np.arange(NY + 1) / NY)
    Y = dct1_definition(f)
    return np.array([Y[k] / (2 * NY) if k == 0 else Y[k] / NY for k in range(NYS + 1)])


def x_peak(kxx, kx, NXH, NXS):
    if kx == 0:
        return 1.0 if kxx == 0 else 0.0
    if kx > NXS:
        return 0.0
    return 0.5 if kxx == kx else 0.0


def z_peak(izz, kz, NZH, NZS):
    if kz == 0:
        return 1.0 if izz == 0 else 0.0
    if abs(kz) > NZS:
        return 0.0
    return 0.5 if izz in (kz, -kz) else 0.0


def x2_peak(kxx, kx, NXH, NXS):
    if kx == 0:
        return 1.0 if kxx == 0 else 0.0
    if kx > NXS:
        return 0.0
    return (1.0 if kx == NXH else 0.5) if kxx == kx else 0.0


def z2_peak(izz, kz, NZH, NZS):
    if kz == 0:
        return 1.0 if izz == 0 else 0.0
    if abs(kz) > NZS or izz not in (kz, -kz):
        return 0.0
    return 1.0 if abs(kz) == NZH else 0.5


def _mode_field(nz, ny, nx, kx, my, kz):
    iz = np.arange(nz)[:, None, None]
    iy = np.arange(ny + 1)[None, :, None]
    ix = np.arange(nx)[None, None, :]
    return np.ascontiguousarray(np.cos(2 * PI * kx * ix / nx) * np.cos(my * PI * iy / ny) * np.cos(2 * PI * kz * iz / nz))


def test_dct1_definition_matches_even_fft():
    x = np.random.default_rng(1).standard_normal(17)
    ext = np.concatenate([x, x[-2:0:-1]])
    assert np.allclose(dct1_definition(x), np.fft.fft(ext).real[:17], atol=1e-12)


@pytest.mark.parametrize("cfgi,modes", [((16, 30, 16), [(16, 3, 16), (0, 0, 0), (5, 7, -9), (16, 30, 16)]),
                                        ((10, 20, 10), [(3, 3, -4), (10, 20, 10), (0, 2, 0)])])
def test_coarse_forward_analytic(cfgi, modes):
    """F1 (tests/test_xyzfft.cpp:201-245), 1e-12."""
    cfg = ref_transform_config(*cfgi)
    o = pyoracle.Oracle(cfg)
    NXS, NYS, NZS = cfgi
    for kx, my, kz in modes:
        P = _mode_field(32, 32, 32, kx, my, kz)
        U = o.scalar()
        o.xyzfft("F", P, U)
        Uy = y_spectrum_expected(my, 32, NYS)
        E = np.zeros_like(U)
        for izz in range(-NZS, NZS + 1):
            for kxx in range(NXS + 1):
                E[izz + NZS, :, kxx] = x_peak(kxx, kx, 16, NXS) * Uy * z_peak(izz, kz, 16, NZS)
        assert np.abs(U - E).max() < 1e-12, (kx, my, kz)


def test_fine_forward_analytic():
    """F5a (tests/test_xyzfft.cpp:331-356): no Nyquist halving in the fine transform, 1e-12."""
    cfg = ref_transform_config(16, 30, 16)
    o = pyoracle.Oracle(cfg)
    for kz in (3, -9, 16):
        for kx in (5, 16):
            P = _mode_field(32, 32, 32, kx, 3, kz)
            U = o.scalar()
            o.xyzfft2("F", P, U)
            Uy = y_spectrum_expected(3, 32, 30)
            E = np.zeros_like(U)
            for izz in range(-16, 17):
                for kxx in range(17):
                    E[izz + 16, :, kxx] = x2_peak(kxx, kx, 16, 16) * Uy * z2_peak(izz, kz, 16, 16)
            assert np.abs(U - E).max() < 1e-12, (kx, kz)


@pytest.mark.parametrize("cfgi", [(16, 30, 16), (10, 20, 10)])
def test_coarse_and_fine_roundtrips(cfgi):
    """F2/F3/F5b/F5c (tests/test_xyzfft.cpp:250-273, 358-372), 1e-12."""
    cfg = ref_transform_config(*cfgi)
    o = pyoracle.Oracle(cfg)
    for fwd, bwd, phys in ((o.xyzfft, o.xyzfft, o.coarse_phys), (o.xyzfft2, o.xyzfft2, o.fine_phys)):
        U0 = band_spectrum(cfg, 0xC0DE)
        P = phys()
        bwd("B", P, U0.copy())
        U1 = o.scalar()
        fwd("F", P, U1)
        assert np.abs(U1 - U0).max() < 1e-12
        P2 = phys()
        bwd("B", P2, U1.copy())
        assert np.abs(P2 - P).max() < 1e-12


def _series_on_product_grid(cfg, U):
    """Direct evaluation of sum U T_j(y_l) e^{i(kx x + kz z)} (real part, Hermitian in kx) on
    the product grid, by dense sums."""
    NX2, NY2, NZ2, NXS, NYS, NZS = (cfg[k] for k in ("NX2", "NY2", "NZ2", "NXS", "NYS", "NZS"))
    l = np.arange(NY2 + 1)
    T = np.cos(PI * np.outer(l, np.arange(NYS + 1)) / NY2)
    Y = np.einsum("lj,zjx->zlx", T, U)
    iz = np.arange(NZ2)
    kz = np.arange(-NZS, NZS + 1)
    Ez = np.exp(2j * PI * np.outer(iz, kz) / NZ2)
    Z = np.einsum("ik,klx->ilx", Ez, Y)
    ix = np.arange(NX2)
    kx = np.arange(NXS + 1)
    Ex = np.exp(2j * PI * np.outer(kx, ix) / NX2)
    w = np.where(kx == 0, 1.0, 2.0)
    Z = Z.copy()
    Z[..., 0] = Z[..., 0].real
    return np.einsum("ilk,kx->ilx", Z * w, Ex).real


@pytest.mark.parametrize("name", ["channel_32", "ref32"])
def test_fine_backward_is_exact_series_evaluation(name):
    """The unnormalised fine synthesis equals the dense evaluation of the truncated series."""
    cfg = config.channel_32() if name == "channel_32" else ref_transform_config(10, 20, 10)
    o = pyoracle.Oracle(cfg)
    U = band_spectrum(cfg, 77)
    P = o.fine_phys()
    o.xyzfft2("B", P, U.copy())
    assert maxrel(P, _series_on_product_grid(cfg, U)) < 1e-13


def test_strided_component_integrity():
    """Backward reads only its component; forward writes only its component
    (tests/test_xyzfft.cpp:375-407)."""
    cfg = config.channel_32()
    o = pyoracle.Oracle(cfg)
    src = band_spectrum(cfg, 0xAFF1)
    guard = -7.0 + 3.0j
    vin = np.full(src.shape + (3,), guard, np.complex128)
    vin[..., 1] = src
    P = o.fine_phys()
    o.xyzfft2("B", P, vin, stride=3, offset=1)
    P0 = o.fine_phys()
    o.xyzfft2("B", P0, src.copy())
    assert np.array_equal(P, P0)
    out = np.full(src.shape + (3,), guard, np.complex128)
    o.xyzfft2("F", P, out, stride=3, offset=2)
    assert (out[..., :2] == guard).all()
    assert np.abs(out[..., 2] - src).max() < 1e-12
    with pytest.raises(ValueError):
        o.xyzfft2("F", P, out, stride=0)
