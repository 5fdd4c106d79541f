"""Seeded initial velocity fields, as the reference's initial_bc namespace builds them.

  DisturbanceEngine / centered_uniform   src/sim/initial_bc.hpp:13, initial_bc.cpp:15-18
  disturb_velocity                       src/sim/initial_bc.cpp:20-55
  initialize_velocity                    src/sim/initial_bc.cpp:66-95
  velocity_boundary                      src/sim/initial_bc.cpp:97-103
  Mesh::YC                               src/core/mesh.cpp:100-110

The engine is the C++ standard library's std::mt19937_64 inside the backend library (host
code, no device work).  The transforms are the coarse-grid Transforms::channel_xyzfft of a
`Channel` (on the GPU); any object with the same `xyzfft(direction, phys, spec, stride, offset)`
and `cfg` serves.  Physical arrays are [iz][iy][ix] (NZ x (NY+1) x NX); spectral velocity is
[(iz+NZS)][iy][kx][c] complex.
"""
import ctypes
import math

import numpy as np

from . import lib


class DisturbanceEngine:
    """std::mt19937_64 seeded with `seed` (the reference's DisturbanceEngine(seed))."""

    def __init__(self, seed):
        h = ctypes.c_void_p()
        if lib().cdns_engine_create(ctypes.c_ulonglong(int(seed) & 0xFFFFFFFFFFFFFFFF), ctypes.byref(h)) != 0:
            raise RuntimeError("cdns_engine_create failed")
        self._h = h

    def close(self):
        if self._h:
            lib().cdns_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _draw(self, kind, n, dtype):
        out = np.empty(int(n), dtype)
        if lib().cdns_engine_draw(self._h, kind, ctypes.c_void_p(out.ctypes.data), ctypes.c_ulonglong(out.size)) != 0:
            raise RuntimeError("cdns_engine_draw failed")
        return out

    def raw(self, n):
        """The next n outputs of engine()."""
        return self._draw(0, n, np.uint64)

    def __call__(self):
        return int(self.raw(1)[0])

    def centered_uniform(self, n):
        """The next n values of centered_uniform(engine): (engine() >> 11) 2^-53 - 0.5."""
        return self._draw(1, n, np.float64)


def wall_normal_points(NY):
    """Mesh::YC of a serial run: cos(pi j / NY) for j < NY/2, mirrored with the opposite sign,
    and 0 at NY/2 (mesh.cpp:100-110; the points the loop does not reach stay 0)."""
    pi = math.acos(-1.0)
    yc = np.zeros(NY + 1)
    for j in range(NY // 2):
        yc[j] = math.cos(pi * float(j) / float(NY))
        yc[NY - j] = -yc[j]
    yc[NY // 2] = 0.0
    return yc


def _centreline(cfg):
    return cfg["Re"] / 2.0 if cfg["INDEX_GRAD_PRE"] == 1 else 1.5


def _spectral_vector(cfg):
    return np.zeros((2 * cfg["NZS"] + 1, cfg["NYS"] + 1, cfg["NXS"] + 1, 3), np.complex128)


def disturb_velocity(transforms, velocity, amplitude, engine, index_disturb_z=None):
    """Adds amplitude-scaled centred uniform noise in physical space, in place.

    Each component is synthesised on the coarse grid; every physical point (iz, iy, ix), in that
    order, draws three values (u, v, w; w times INDEX_disturb_Z) scaled by
    uc (1 - YC^2) amplitude; the result is analysed back into `velocity`.  amplitude == 0 leaves
    the field and the engine untouched."""
    if amplitude == 0.0:
        return velocity
    cfg = transforms.cfg
    dz = cfg.get("INDEX_disturb_Z", 1) if index_disturb_z is None else index_disturb_z
    NX, NY, NZ = cfg["NX"], cfg["NY"], cfg["NZ"]
    phys = np.zeros((3, NZ, NY + 1, NX))
    for c in range(3):
        transforms.xyzfft("B", phys[c], velocity, 3, c)
    yc = wall_normal_points(NY)
    uc = _centreline(cfg)
    scale = (uc * (1.0 - yc * yc) * amplitude)[None, :, None]
    draws = engine.centered_uniform(3 * NZ * (NY + 1) * NX).reshape(NZ, NY + 1, NX, 3)
    phys[0] += scale * draws[..., 0]
    phys[1] += scale * draws[..., 1]
    phys[2] += (scale * draws[..., 2]) * float(dz)
    for c in range(3):
        transforms.xyzfft("F", phys[c], velocity, 3, c)
    return velocity


def initialize_velocity(transforms, engine=None, eps_init=None, index_disturb_z=None):
    """The laminar profile (uc cos(theta) (1 - y^2), V_Microblow0, uc sin(theta) (1 - y^2)) analysed
    on the coarse grid, then disturb_velocity with EPS_init.  Returns the spectral velocity."""
    cfg = transforms.cfg
    eps = cfg.get("EPS_init", 0.0) if eps_init is None else eps_init
    NX, NY, NZ = cfg["NX"], cfg["NY"], cfg["NZ"]
    th = cfg["THETA"] * math.acos(-1.0) / 180.0
    uc = _centreline(cfg)
    scale = (uc * math.cos(th), cfg["V_Microblow0"], uc * math.sin(th))
    yc = wall_normal_points(NY)
    velocity = _spectral_vector(cfg)
    phys = np.zeros((NZ, NY + 1, NX))
    for c in range(3):
        profile = np.ones(NY + 1) if c == 1 else 1.0 - yc * yc
        phys[...] = (scale[c] * profile)[None, :, None]
        transforms.xyzfft("F", phys, velocity, 3, c)
    if eps != 0.0:
        if engine is None:
            raise ValueError("a disturbance needs an engine")
        disturb_velocity(transforms, velocity, eps, engine, index_disturb_z)
    return velocity


def velocity_boundary(cfg):
    """[(iz+NZS)][wall][kx][c]: V_Microblow0 on the v component of the mean mode at both walls."""
    b = np.zeros((2 * cfg["NZS"] + 1, 2, cfg["NXS"] + 1, 3), np.complex128)
    b[cfg["NZS"], :, 0, 1] = cfg["V_Microblow0"]
    return b
