"""Shared test helpers: configurations, seeded inputs and comparisons."""
import math

import numpy as np

from cdns import config

PI = math.pi


def ref_transform_config(nxs=16, nys=30, nzs=16):
    """The reference's 3-D transform test mesh (tests/test_xyzfft.cpp:30-38): 32^3 with product grid 32^3."""
    return config.make(32, 32, 32, nxs, nys, nzs, 32, 32, 32, 2 * PI, 2 * PI, 1.0, 1e-3)


def stage3_config():
    """examples/config.toml of the reference (tests/stage3_fixture.hpp:5-7)."""
    return config.reference_example()


def band_spectrum(cfg, seed):
    """Band-limited realisable spectrum (tests/test_xyzfft.cpp:127-162 construction, own RNG):
    kx=0 (and kx=NXH when full) columns Hermitian in kz; full-z Nyquist slots equal."""
    NXS, NYS, NZS = cfg["NXS"], cfg["NYS"], cfg["NZS"]
    NXH, NZH = cfg["NX2"] // 2, cfg["NZ2"] // 2
    rng = np.random.default_rng(seed)
    s = np.zeros((2 * NZS + 1, NYS + 1, NXS + 1), np.complex128)
    fullz, fullx = NZS == NZH, NXS == NXH
    for iy in range(NYS + 1):
        for kx in range(NXS + 1):
            if kx == 0 or (fullx and kx == NXH):
                for kz in range(NZS + 1):
                    if kz == 0:
                        s[NZS, iy, kx] = rng.uniform(-1, 1)
                    elif fullz and kz == NZH:
                        r = rng.uniform(-1, 1)
                        s[NZS + kz, iy, kx] = r
                        s[NZS - kz, iy, kx] = r
                    else:
                        v = rng.uniform(-1, 1) + 1j * rng.uniform(-1, 1)
                        s[NZS + kz, iy, kx] = v
                        s[NZS - kz, iy, kx] = np.conj(v)
            else:
                s[:, iy, kx] = rng.uniform(-1, 1, 2 * NZS + 1) + 1j * rng.uniform(-1, 1, 2 * NZS + 1)
                if fullz:
                    s[NZS - NZH, iy, kx] = s[NZS + NZH, iy, kx]
    return s


def random_vector(shape_like, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    return scale * (rng.standard_normal(shape_like.shape) + 1j * rng.standard_normal(shape_like.shape))


def maxrel(a, b):
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))
