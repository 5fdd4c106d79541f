"""Pins of the oracle's spectral operators (tests/test_vector_ops.cpp, test_nonlinear_ops.cpp)."""
import numpy as np
import pytest

from cdns import config
from helpers import stage3_config
import pyoracle


@pytest.fixture(scope="module")
def orc():
    return pyoracle.Oracle(stage3_config())


def test_chebyshev_derivative_and_flow_rate(orc):
    """T4' = 8 T3 + 8 T1 and Q(2 T0 + 3 T2) = 1 (tests/test_vector_ops.cpp:34-47)."""
    f = orc.scalar()
    f[16, 4, 0] = 1.0
    d = orc.derivative(1, f)
    assert d[16, 3, 0].real == 8.0 and d[16, 1, 0].real == 8.0
    assert np.abs(np.delete(d.ravel(), [np.ravel_multi_index((16, 3, 0), d.shape),
                                        np.ravel_multi_index((16, 1, 0), d.shape)])).max() == 0.0
    f[...] = 0
    f[16, 0, 0] = 2.0
    f[16, 2, 0] = 3.0
    assert orc.flow_rate(f) == 1.0


def test_fourier_derivatives(orc):
    t = orc.tables()
    f = orc.scalar()
    f[-2 + 16, 3, 2] = 0.25 - 0.75j
    d = orc.derivative(0, f)
    assert abs(d[14, 3, 2] - 1j * t["WAVX"][2] * f[14, 3, 2]) < 1e-14
    d = orc.derivative(2, f)
    assert abs(d[14, 3, 2] - 1j * t["WAVZ"][14] * f[14, 3, 2]) < 1e-14


def test_chebyshev_derivative_against_numpy():
    """D f equals numpy's Chebyshev series derivative."""
    cfg = config.channel_32()
    o = pyoracle.Oracle(cfg)
    rng = np.random.default_rng(4)
    f = o.scalar()
    f[...] = rng.standard_normal(f.shape) + 1j * rng.standard_normal(f.shape)
    d = o.derivative(1, f)
    ref = np.polynomial.chebyshev.chebder(np.moveaxis(f, 1, 0), axis=0)
    ref = np.concatenate([ref, np.zeros((1,) + ref.shape[1:])], axis=0)
    assert np.abs(d - np.moveaxis(ref, 0, 1)).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("nys,nxs", [(4, 7), (7, 8), (30, 16)])
def test_curl_matches_independent_derivatives(nys, nxs):
    """tests/test_vector_ops.cpp:66-103, 1e-12."""
    cfg = stage3_config()
    cfg["NYS"], cfg["NXS"] = nys, nxs
    o = pyoracle.Oracle(cfg)
    z = np.arange(-16, 17)[:, None, None]
    y = np.arange(nys + 1)[None, :, None]
    x = np.arange(nxs + 1)[None, None, :]
    v = o.vector()
    for c in range(3):
        v[..., c] = np.sin(.3 * x + .7 * y + c) + 1j * np.cos(.2 * z + .9 * y + c)
    orig = v.copy()
    w = o.curl(v)
    D = [[o.derivative(a, np.ascontiguousarray(v[..., c])) for a in range(3)] for c in range(3)]
    for c in range(3):
        j, k = (c + 1) % 3, (c + 2) % 3
        assert np.abs(w[..., c] - (D[k][j] - D[j][k])).max() < 1e-12
    assert np.array_equal(v, orig)


def test_div_curl_vanishes(orc):
    """tests/test_vector_ops.cpp:49-64."""
    v = orc.vector()
    for iz in range(-2, 3):
        for iy in range(5):
            for ix in range(17):
                for k in range(3):
                    v[iz + 16, iy, ix, k] = 0.01 * (1 + ix + iy + k) + 0.02j * iz
    assert np.abs(orc.divergence(orc.curl(v))).max() < 1e-12


def test_gradient_and_laplacian_of_T2_mode(orc):
    """f = T2(y) e^{i(kx x + kz z)}: T2' = 4 T1, T2'' = 4 T0 (tests/test_vector_ops.cpp:105-139)."""
    t = orc.tables()
    f = orc.scalar()
    amp = 2.0 - 1.0j
    f[14, 2, 1] = amp
    g = orc.gradient(f)
    E = np.zeros_like(g)
    E[14, 2, 1, 0] = 1j * t["WAVX"][1] * amp
    E[14, 1, 1, 1] = 4.0 * amp
    E[14, 2, 1, 2] = 1j * t["WAVZ"][14] * amp
    assert np.abs(g - E).max() < 1e-13
    lap = orc.divergence(g)
    L = np.zeros_like(f)
    L[14, 0, 1] = 4.0 * amp
    L[14, 2, 1] = -(t["WAVX"][1] ** 2 + t["WAVZ"][14] ** 2) * amp
    assert np.abs(lap - L).max() < 1e-12


def test_cross_product_constant_vectors():
    """(1,2,3) x (4,3,2) = (-5,10,-5); a x a = 0; antisymmetry (tests/test_nonlinear_ops.cpp:9-33)."""
    o = pyoracle.Oracle(config.channel_32())
    a, b = o.vector(), o.vector()
    for c in range(3):
        a[16, 0, 0, c] = c + 1.0
        b[16, 0, 0, c] = 4.0 - c
    ab, ba, aa = o.cross_product(a, b), o.cross_product(b, a), o.cross_product(a, a)
    assert np.abs(ab[16, 0, 0] - np.array([-5.0, 10.0, -5.0])).max() < 1e-12
    ab[16, 0, 0] = 0
    assert np.abs(ab).max() < 1e-13
    assert np.abs(aa).max() < 1e-13
    assert np.abs(o.cross_product(a, b) + ba).max() < 1e-13


def test_cross_product_single_modes_closed_form():
    """u = (cos(kx x), 0, 0), w = (0, 0, T1(y)): u x w = (0, -cos(kx x) y, 0); the product of two
    Fourier modes lands on the sum/difference wavenumbers."""
    cfg = config.channel_32()
    o = pyoracle.Oracle(cfg)
    a, b = o.vector(), o.vector()
    a[16, 0, 3, 0] = 0.5  # cos(3 alpha x) = (e^{i} + e^{-i})/2, only kx >= 0 stored
    b[16, 1, 0, 2] = 1.0  # T1(y) = y
    n = o.cross_product(a, b)
    E = o.vector()
    E[16, 1, 3, 1] = -0.5
    assert np.abs(n - E).max() < 1e-13


def test_nonlinear_term_of_zero_and_uniform_velocity():
    """tests/test_timestepper.cpp:7-19, 56-74."""
    o = pyoracle.Oracle(stage3_config())
    v = o.vector()
    nl, bc = o.nonlinear_term(v)
    assert np.abs(nl).max() == 0 and np.abs(bc).max() == 0
    for c in range(3):
        v[16, 0, 0, c] = c + 1.0
    nl, bc = o.nonlinear_term(v)
    assert np.abs(nl).max() < 1e-12 and np.abs(bc).max() < 1e-12
