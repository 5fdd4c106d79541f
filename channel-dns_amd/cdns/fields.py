"""Seeded initial velocity fields and field comparisons.

Fields are numpy complex128 arrays in the reference's spectral layout
`[(iz+NZS)][iy][kx][c]` (src/solver/fields.hpp:17-22): Chebyshev coefficients
in y, Fourier coefficients in x (kx >= 0 only) and z (-NZS..NZS).

The perturbation is divergence-free by construction: u' = curl(A) with
A = (1 - y^2)^2 * phi, which also makes u' and its wall-normal derivative of A
vanish at both walls, so u' satisfies no-slip exactly.
"""
import math

import numpy as np


def _cheb_mul_poly(c, p):
    """Chebyshev-series product along axis 1: c (.., n, ..) times the short series p."""
    n = c.shape[1]
    out = np.zeros_like(c)
    for a in range(len(p)):
        if p[a] == 0:
            continue
        for b in range(n):
            for k in (a + b, abs(a - b)):
                if k < n:
                    out[:, k] += 0.5 * p[a] * c[:, b]
    return out


def cheb_dy(c):
    """Chebyshev derivative coefficients along axis 1 (recurrence of vector_ops.cpp:17-25)."""
    n = c.shape[1] - 1
    d = np.zeros_like(c)
    if n == 0:
        return d
    d[:, n - 1] = 2 * n * c[:, n]
    for y in range(n - 2, -1, -1):
        d[:, y] = d[:, y + 2] + 2 * (y + 1) * c[:, y + 1]
    d[:, 0] *= 0.5
    return d


def _wavenumbers(cfg):
    kx = np.arange(cfg["NXS"] + 1) * 2 * math.pi / cfg["XL"]
    kz = np.arange(-cfg["NZS"], cfg["NZS"] + 1) * 2 * math.pi / cfg["ZL"]
    return kx, kz


def physical_values(cfg, u):
    """Evaluate a spectral scalar field [kz][iy][kx] on the coarse grid NZ x (NY+1) x NX."""
    NX, NY, NZ, NXS, NYS, NZS = (cfg[k] for k in ("NX", "NY", "NZ", "NXS", "NYS", "NZS"))
    l = np.arange(NY + 1)
    T = np.cos(np.pi * np.outer(l, np.arange(NYS + 1)) / NY)  # (NY+1, NYS+1)
    Y = np.einsum("lj,zjx->zlx", T, u)  # (nzk, NY+1, nkx)
    Z = np.zeros((NZ, NY + 1, NXS + 1), np.complex128)
    for kz in range(-NZS, NZS + 1):
        Z[kz % NZ] += Y[kz + NZS]
    Z = np.fft.ifft(Z, axis=0) * NZ
    X = np.zeros((NZ, NY + 1, NX // 2 + 1), np.complex128)
    X[..., : NXS + 1] = Z
    return np.fft.irfft(X, n=NX, axis=2) * NX


def _hermitian_kx0(phi, NZS):
    """Make the kx=0 column of a [kz][iy][kx] array Hermitian in kz (real field)."""
    col = phi[:, :, 0]
    sym = 0.5 * (col + np.conj(col[::-1]))
    phi[:, :, 0] = sym
    phi[NZS, :, 0] = phi[NZS, :, 0].real
    return phi


def mean_profile(cfg, kind):
    """Chebyshev coefficients (NYS+1) of the streamwise mean velocity.

    'laminar': u = (Re/2)(1 - y^2), the centreline of dp/dx=-1 (initial_bc.cpp:71).
    'log': Reichardt's law of the wall in both halves, interpolated on the
           NYS+1 Gauss-Lobatto points (exactly zero at the walls).
    """
    NYS, Re = cfg["NYS"], cfg["Re"]
    uc = Re / 2.0 if cfg["INDEX_GRAD_PRE"] == 1 else 1.5
    c = np.zeros(NYS + 1)
    if kind == "laminar":
        c[0], c[2] = uc / 2.0, -uc / 2.0
        return c, uc
    y = np.cos(np.pi * np.arange(NYS + 1) / NYS)
    yp = (1.0 - np.abs(y)) * Re
    kappa = 0.41
    up = np.log(1 + kappa * yp) / kappa + 7.8 * (1 - np.exp(-yp / 11.0) - yp / 11.0 * np.exp(-yp / 3.0))
    up[0] = up[-1] = 0.0
    # DCT-I interpolation on the NYS+1 points.
    j = np.arange(NYS + 1)
    w = np.ones(NYS + 1)
    w[0] = w[-1] = 0.5
    M = np.cos(np.pi * np.outer(j, j) / NYS)
    c = (2.0 / NYS) * (M @ (w * up))
    c[0] *= 0.5
    c[-1] *= 0.5
    return c, float(up.max())


def seeded_velocity(cfg, seed=1234, mean="laminar", spectrum="band", kmax=4, amplitude=0.1, kcut=None):
    """Initial spectral velocity [(iz+NZS)][iy][kx][c] = mean profile + curl(A) perturbation.

    spectrum='band': phi has random modes with |kx|,|kz|,j-index <= kmax; the
        perturbation is scaled so that max |u'| on the coarse grid = amplitude*centreline.
    spectrum='k53':  phi has random phases and |u'_k| ~ |k|^(-11/6) (a k^-5/3 energy
        spectrum) over the retained modes below 2/3 of the cut-off, or with kx, |kz| and the
        Chebyshev index all <= kcut when kcut is given; scaled to an rms of amplitude*centreline
        on the coarse grid.
    """
    NXS, NYS, NZS = cfg["NXS"], cfg["NYS"], cfg["NZS"]
    nzk, nj, nkx = 2 * NZS + 1, NYS + 1, NXS + 1
    rng = np.random.default_rng(seed)
    kx, kz = _wavenumbers(cfg)
    prof, uc = mean_profile(cfg, mean)
    A = []
    jmax = NYS - 4
    for comp in range(3):
        phi = np.zeros((nzk, nj, nkx), np.complex128)
        if spectrum == "band":
            for izi in range(nzk):
                for ix in range(nkx):
                    if abs(izi - NZS) > kmax or ix > kmax:
                        continue
                    m = min(kmax, jmax) + 1
                    phi[izi, :m, ix] = rng.standard_normal(m) + 1j * rng.standard_normal(m)
        else:
            KX = kx[None, None, :]
            KZ = kz[:, None, None]
            KY = (np.pi / 2.0) * np.arange(nj)[None, :, None]
            kk = np.sqrt(KX ** 2 + KZ ** 2 + KY ** 2)
            kk[kk == 0] = 1.0
            cz, cx, cy = (2 * NZS) // 3, (2 * NXS) // 3, min(jmax, (2 * NYS) // 3)
            if kcut is not None:
                cz, cx, cy = min(cz, kcut), min(cx, kcut), min(cy, kcut)
            mask = (np.abs(np.arange(-NZS, NZS + 1))[:, None, None] <= cz) & \
                   (np.arange(nkx)[None, None, :] <= cx) & (np.arange(nj)[None, :, None] <= cy)
            amp = kk ** (-17.0 / 6.0) * mask
            phase = rng.uniform(0, 2 * np.pi, size=amp.shape)
            phi = amp * np.exp(1j * phase) * rng.standard_normal(amp.shape)
        phi = _hermitian_kx0(phi, NZS)
        phi[NZS, :, 0] = 0.0  # no mean potential
        a = _cheb_mul_poly(phi, [3.0 / 8.0, 0.0, -0.5, 0.0, 1.0 / 8.0])
        A.append(a)
    ikx = 1j * kx[None, None, :]
    ikz = 1j * kz[:, None, None]
    u = cheb_dy(A[2]) - ikz * A[1]
    v = ikz * A[0] - ikx * A[2]
    w = ikx * A[1] - cheb_dy(A[0])
    pert = np.stack([u, v, w], axis=-1)
    phys = [physical_values(cfg, pert[..., c]) for c in range(3)]
    if spectrum == "band":
        scale_now = max(np.abs(p).max() for p in phys)
    else:
        scale_now = math.sqrt(sum(float(np.mean(p ** 2)) for p in phys) / 3.0)
    if scale_now > 0:
        pert *= amplitude * uc / scale_now
    vel = pert
    vel[NZS, :, 0, 0] += prof
    return np.ascontiguousarray(vel)


def rel_l2(a, b):
    """||a - b|| / ||b|| over all entries."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def rel_l2_components(a, b):
    return [rel_l2(a[..., c], b[..., c]) for c in range(a.shape[-1])]


# ------------------------------------------------------------ flow diagnostics --
def mean_streamwise(cfg, vel):
    """Chebyshev coefficients of the plane-averaged streamwise velocity (mode kx = kz = 0)."""
    return np.real(vel[cfg["NZS"], :, 0, 0]).copy()


def wall_shear(cfg, vel):
    """Mean wall shear d<u>/dy at y = +1 and y = -1 from the Chebyshev coefficients:
    T_j'(1) = j^2 and T_j'(-1) = (-1)^(j+1) j^2 (doc/algorithm.md:174)."""
    c = mean_streamwise(cfg, vel)
    j = np.arange(c.size, dtype=np.float64)
    return float(np.sum(j * j * c)), float(np.sum((-1.0) ** (j + 1) * j * j * c))


def divergence_l2(cfg, vel):
    """L2 norm of the spectral divergence i kx u + dv/dy + i kz w over all retained modes."""
    kx, kz = _wavenumbers(cfg)
    u, v, w = vel[..., 0], vel[..., 1], vel[..., 2]
    dv = cheb_dy(v)  # axis 1 is the Chebyshev index
    div = 1j * kx[None, None, :] * u + dv + 1j * kz[:, None, None] * w
    return float(np.linalg.norm(div.ravel()))


def precision_report(cfg, got, ref):
    """Errors of a state against a reference state: per-component velocity rel-L2, the mean
    profile (rel-L2 of its coefficients), both wall shears (relative) and the divergence of each."""
    sg, sr = wall_shear(cfg, got), wall_shear(cfg, ref)
    return {
        "velocity_rel_l2": rel_l2_components(got, ref),
        "mean_profile_rel_l2": rel_l2(mean_streamwise(cfg, got), mean_streamwise(cfg, ref)),
        "wall_shear_rel": [abs(a - b) / abs(b) for a, b in zip(sg, sr)],
        "divergence_l2": [divergence_l2(cfg, got), divergence_l2(cfg, ref)],
    }
