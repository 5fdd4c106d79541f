"""Statistics and diagnostics of the oracle: plane moments (statistics.cpp:54-90), Clenshaw-Curtis
weights (diagnostics.cpp:17-33) and Diagnostics::evaluate (diagnostics.cpp:39-75), pinned by the
reference's own known answers (tests/test_statistics.cpp:49-80, tests/test_diagnostics.cpp:11-55)
and by quadrature identities."""
import math

import numpy as np

from cdns import config
import pyoracle


def _coords(cfg):
    """Mesh coordinates (mesh.cpp:94-116): XC = (i+1) dx, ZC = (k+1) dz, YC = cos(pi j / NY)."""
    NX, NY, NZ = cfg["NX"], cfg["NY"], cfg["NZ"]
    XC = (np.arange(NX) + 1) * (cfg["XL"] / NX)
    ZC = (np.arange(NZ) + 1) * (cfg["ZL"] / NZ)
    YC = np.cos(np.pi * np.arange(NY + 1) / NY)
    YC[NY // 2] = 0.0
    YC[NY // 2 + 1:] = -YC[: NY // 2][::-1]
    return XC, YC, ZC


def analytic_fields(cfg):
    """The fields of tests/test_diagnostics.cpp:23-35: u = y cos(kx x), v = sin(kx x), w = 2 cos(kz z),
    raw pressure 5 + |u|^2/2, in the physical layout [iz][iy][ix]."""
    XC, YC, ZC = _coords(cfg)
    kx, kz = 2 * math.pi / cfg["XL"], 2 * math.pi / cfg["ZL"]
    Z, Y, X = np.meshgrid(ZC, YC, XC, indexing="ij")
    u = Y * np.cos(kx * X)
    v = np.sin(kx * X)
    w = 2 * np.cos(kz * Z)
    p = 5 + 0.5 * (u * u + v * v + w * w)
    return [np.ascontiguousarray(a) for a in (u, v, w, p)], kx, kz


def test_plane_moments_of_uniform_samples():
    """Uniform v = (1+2s, 2+2s, 3+2s), corrected pressure 4+2s (tests/test_statistics.cpp:49-62):
    every plane sum divided by the plane size is the closed-form moment."""
    cfg = config.reference_example()
    o = pyoracle.Oracle(cfg)
    shape = o.coarse_phys().shape
    n = cfg["NX"] * cfg["NZ"]
    for s in range(2):
        u, v, w = (np.full(shape, 1.0 + c + 2 * s) for c in range(3))
        pc = 4.0 + 2 * s
        p = np.full(shape, pc + 0.5 * (u[0, 0, 0] ** 2 + v[0, 0, 0] ** 2 + w[0, 0, 0] ** 2))
        sums, fin = o.plane_moments(u, v, w, p)
        a, b, c = 1.0 + 2 * s, 2.0 + 2 * s, 3.0 + 2 * s
        want = [a, b, c, pc, a * a, b * b, c * c, a * b, b * c, c * a, pc * a, pc * b, pc * c, pc * pc,
                a ** 3, b ** 3, c ** 3, a ** 4, b ** 4, c ** 4]
        assert fin.all()
        np.testing.assert_allclose(sums / n, np.tile(want, (shape[1], 1)), rtol=1e-13)


def test_clenshaw_curtis_weights_integrate_polynomials():
    """Weights / 2 over [-1, 1]: they integrate 1 to 1 and y^2 to 1/3, y^4 to 1/5, odd powers to 0
    (exact for the polynomial degrees the rule covers)."""
    for NY in (16, 32, 33):
        cfg = dict(config.reference_example(), NY=NY, NYS=min(NY - 2, 30))
        o = pyoracle.Oracle(cfg)
        wt = o.cc_weights()
        y = np.cos(np.pi * np.arange(NY + 1) / NY)
        for k, want in ((0, 1.0), (1, 0.0), (2, 1 / 3), (3, 0.0), (4, 1 / 5), (8, 1 / 9)):
            assert abs(np.sum(wt * y ** k) - want) < 1e-14, (NY, k)


def test_diagnostics_analytic_known_answer():
    """tests/test_diagnostics.cpp:11-55: EKT = n/2, TKE = 4/3, divergence L2 = sqrt(kx^2/6 + 2 kz^2),
    plane mean of the corrected pressure = 5; a non-finite spectral flag or pressure propagates."""
    cfg = config.reference_example()
    o = pyoracle.Oracle(cfg)
    (u, v, w, p), kx, kz = analytic_fields(cfg)
    spec = o.vector()
    for c, f in enumerate((u, v, w)):
        o.xyzfft("F", f, spec, stride=3, offset=c)
    sums, fin = o.plane_moments(u, v, w, p)
    d = o.diagnostics(sums, fin, spec, True)
    n = u.size
    assert d["finite"]
    assert abs(d["ekt"] - 0.5 * n) < 1e-8 * n
    assert abs(d["tke"] - 4.0 / 3.0) < 1e-12
    assert abs(d["divergence_l2"] - math.sqrt(kx * kx / 6 + 2 * kz * kz)) < 1e-11
    assert np.all(np.abs(sums[:, 3] / (cfg["NX"] * cfg["NZ"]) - 5) < 1e-13)
    assert not o.diagnostics(sums, fin, spec, False)["finite"]
    p_bad = p.copy()
    p_bad.flat[1] = np.nan
    sums_bad, fin_bad = o.plane_moments(u, v, w, p_bad, higher=False)
    assert not fin_bad.all()
    assert not o.diagnostics(sums_bad, fin_bad, spec, True)["finite"]
