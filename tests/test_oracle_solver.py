"""Pins of the oracle's Chebyshev-tau solver, influence matrix and sigma correction.

Besides the reference's own checks (tests/test_eqsolve.cpp), each piece is pinned against
mathematics: the tau recurrence against a dense solve of the same tau system, the tau system
against calculus (manufactured polynomials), and the fully corrected coupled solve against a
dense solve of the discrete Stokes problem it must satisfy (momentum for j <= NYS-2, velocity
wall conditions, and an exactly divergence-free velocity).
"""
import numpy as np
import pytest

from cdns import config
from helpers import stage3_config
import pyoracle


def small_config(nys=12, nxs=4, nzs=3, a_v=(1.0, 1.0, 1.0), b_v=(0.0, 0.0, 0.0)):
    cfg = config.make(16, 16, 16, nxs, nys, nzs, 24, 24, 24, 4 * np.pi, 2 * np.pi, 180.0, 1e-3,
                      a_V=a_v, b_V=b_v)
    return cfg


def cheb_D(n):
    """Chebyshev derivative matrix acting on coefficients 0..n."""
    D = np.zeros((n + 1, n + 1))
    for k in range(n + 1):
        e = np.zeros(n + 1)
        e[k] = 1.0
        d = np.polynomial.chebyshev.chebder(e)
        D[: len(d), k] = d
    return D


def dense_tau(n, w, a, b, rhs, lo, up, WL, WD, WR):
    """The tau system solve_tau_column eliminates (eqsolve_pv.cpp:16-60), assembled densely."""
    A = np.zeros((n + 1, n + 1))
    r = np.zeros(n + 1, complex)

    def f(j):
        return (a + b * j * j) if not (a == 0 and b == 0) else ((1 - j % 2) * j * j + j % 2)

    for j in range(0, n + 1, 2):
        A[0, j] = f(j)
    for j in range(1, n + 1, 2):
        A[1, j] = f(j)
    r[0], r[1] = (lo + up) / 2, (lo - up) / 2
    for j in range(2, n + 1):
        A[j, j] = 1 + WD[j] * w
        A[j, j - 2] += WL[j] * w
        if j + 2 <= n:
            A[j, j + 2] += WR[j] * w
        r[j] = WL[j] * rhs[j - 2] + WD[j] * rhs[j] + (WR[j] * rhs[j + 2] if j + 2 <= n else 0)
    return np.linalg.solve(A, r)


@pytest.mark.parametrize("ab", [(1.0, 0.0), (0.0, 1.0), (1.0, 0.3), (0.0, 0.0)])
def test_tau_recurrence_matches_dense_solve(ab):
    cfg = config.channel_32()
    o = pyoracle.Oracle(cfg)
    g = 270000.0
    o.setup(g)
    t = o.tables()
    wn = o.table("wn_v")
    rng = np.random.default_rng(1)
    rhs = o.scalar()
    rhs[...] = rng.standard_normal(rhs.shape) + 1j * rng.standard_normal(rhs.shape)
    bc = o.boundary()
    bc[...] = rng.standard_normal(bc.shape) + 1j * rng.standard_normal(bc.shape)
    out = o.solve_tau(rhs, g, ab[0], ab[1], bc, which=1)
    n = cfg["NYS"]
    for izi in (0, 5, 16, 20, 32):
        for kx in (0, 1, 7, 16):
            if izi == 16 and kx == 0 and ab == (0.0, 0.0):
                continue  # not a pure-Neumann problem unless g == 0
            d = dense_tau(n, wn[izi, kx], ab[0], ab[1], rhs[izi, :, kx], bc[izi, 0, kx], bc[izi, 1, kx],
                          t["WL"], t["WD"], t["WR"])
            assert np.abs(d - out[izi, :, kx]).max() < 1e-13 * np.abs(d).max()


def test_tau_recovers_manufactured_polynomials():
    """For f of degree <= NYS with a f(+-1) + b f'(+-1) = h, r = f'' + w f, the tau solve returns f."""
    cfg = small_config(nys=14)
    o = pyoracle.Oracle(cfg)
    g = 5000.0
    o.setup(g)
    wn = o.table("wn_v")
    n = cfg["NYS"]
    D = cheb_D(n)
    rng = np.random.default_rng(3)
    for a, b in ((1.0, 0.0), (0.7, 0.2)):
        F = o.scalar()
        F[...] = rng.standard_normal(F.shape) + 1j * rng.standard_normal(F.shape)
        R = np.einsum("jk,zkx->zjx", D @ D, F) + wn[:, None, :] * F
        bc = o.boundary()
        Tp = np.ones(n + 1)
        Tm = (-1.0) ** np.arange(n + 1)
        j2 = np.arange(n + 1) ** 2
        bc[:, 0, :] = np.einsum("j,zjx->zx", a * Tp + b * j2, F)            # y = +1, d/dn = d/dy
        bc[:, 1, :] = np.einsum("j,zjx->zx", a * Tm + b * Tm * j2, F)       # y = -1, d/dn = -d/dy
        out = o.solve_tau(R, g, a, b, bc, which=1)
        assert np.abs(out - F).max() < 1e-10 * np.abs(F).max()


def test_setup_closed_forms_and_no_slip_unit_responses():
    """wn = -(kx^2 + kz^2 + g); finite ad; unit pressure-gradient responses satisfy no-slip and
    setup is reproducible (tests/test_eqsolve.cpp:57-75, 112-136)."""
    cfg = stage3_config()
    o = pyoracle.Oracle(cfg)
    t = o.tables()
    o.setup(150.0)
    wn = o.table("wn_v")
    assert np.array_equal(wn, -(t["WAVX"][None, :] ** 2 + t["WAVZ"][:, None] ** 2 + 150.0))
    assert np.isfinite(o.table("ad_v")).all()
    q = o.table("q_unit")
    assert np.isfinite(q).all() and (q != 0).all()
    first = o.table("u_dpdx")
    for name in ("u_dpdx", "w_dpdz"):
        r = o.table(name)
        assert abs(r.sum()) < 1e-12 and abs((r * (-1.0) ** np.arange(r.size)).sum()) < 1e-12
    o.setup(300.0)
    assert not np.array_equal(o.table("u_dpdx"), first)
    o.setup(150.0)
    assert np.abs(o.table("u_dpdx") - first).max() <= 1e-15


def test_zero_input_gives_zero_wall_divergence():
    """tests/test_eqsolve.cpp:77-94."""
    cfg = stage3_config()
    o = pyoracle.Oracle(cfg)
    o.setup(150.0)
    P, V = o.solve_modes(o.scalar(), o.boundary(), o.vector(), o.boundary(3), correct_sigma=False)
    assert np.isfinite(P).all()
    assert np.abs(o.divergence_boundary(V)).max() < 1e-10


def _rand(o, seed):
    rng = np.random.default_rng(seed)
    rp = o.scalar()
    rp[...] = (rng.standard_normal(rp.shape) + 1j * rng.standard_normal(rp.shape)) * 1e-3
    rv = o.vector()
    rv[...] = (rng.standard_normal(rv.shape) + 1j * rng.standard_normal(rv.shape)) * 1e-3
    bp = o.boundary()
    bp[...] = (rng.standard_normal(bp.shape) + 1j * rng.standard_normal(bp.shape)) * 1e-3
    bv = o.boundary(3)
    bv[...] = (rng.standard_normal(bv.shape) + 1j * rng.standard_normal(bv.shape)) * 1e-3
    return rp, rv, bp, bv


@pytest.mark.parametrize("nxs", [7, 8, 10, 16])
@pytest.mark.parametrize("neumann_v", [False, True])
def test_fused_solve_matches_staged_composition(nxs, neumann_v):
    """solve_modes == solve_coupled_base + influence + sigma (tests/test_eqsolve.cpp:138-184)."""
    cfg = stage3_config()
    cfg["NXS"] = nxs
    if neumann_v:
        cfg["a_V"][1], cfg["b_V"][1] = 0.0, 1.0
    o = pyoracle.Oracle(cfg)
    o.setup(150.0)
    rp, rv, bp, bv = _rand(o, nxs)
    for correct in (False, True):
        for fuse in (False, True):
            ref_rhs = o.divergence(rv) if fuse else rp
            P1, V1 = o.solve_staged(ref_rhs, bp, rv, bv, correct)
            P2, V2 = o.solve_modes(None if fuse else rp, bp, rv, bv, correct)
            assert np.abs(P2 - P1).max() <= 1e-10 * max(1.0, np.abs(P1).max())
            assert np.abs(V2 - V1).max() <= 1e-10 * max(1.0, np.abs(V1).max())


def dense_coupled(n, kx, kz, w, Re, R, pbc_unused, vbc, aV, bV):
    """The discrete Stokes problem the corrected solve satisfies, for one (kx,kz) != (0,0):
    (D^2 + w) u_c - Re (d_c P - R_c) = 0 for j <= n-2 (3(n-1) rows), a u + b du/dn = h at both
    walls (6 rows), i kx u + D v + i kz w = 0 for all j (n+1 rows).  Unknowns [P, u, v, w]."""
    N = n + 1
    D = cheb_D(n)
    D2 = D @ D
    I = np.eye(N)
    A = np.zeros((4 * N, 4 * N), complex)
    b = np.zeros(4 * N, complex)
    grad = [1j * kx * I, D, 1j * kz * I]
    row = 0
    for c in range(3):
        blk = D2 + w * I
        for j in range(n - 1):
            A[row, N * (1 + c): N * (2 + c)] = blk[j]
            A[row, 0:N] = -Re * grad[c][j]
            b[row] = -Re * R[c][j]
            row += 1
    jj = np.arange(N)
    Tp, Tm = np.ones(N), (-1.0) ** jj
    for c in range(3):
        A[row, N * (1 + c): N * (2 + c)] = aV[c] * Tp + bV[c] * jj ** 2
        b[row] = vbc[0][c]
        row += 1
        A[row, N * (1 + c): N * (2 + c)] = aV[c] * Tm + bV[c] * Tm * jj ** 2
        b[row] = vbc[1][c]
        row += 1
    for j in range(N):
        A[row, N * 1 + j] = 1j * kx
        A[row, N * 2: N * 3] = D[j]
        A[row, N * 3 + j] = 1j * kz
        row += 1
    assert row == 4 * N
    x = np.linalg.solve(A, b)
    return x[:N], np.stack([x[N:2 * N], x[2 * N:3 * N], x[3 * N:]], axis=-1)


# (a Neumann condition on v is implied by continuity at the wall, which makes this system singular)
@pytest.mark.parametrize("aV,bV", [((1.0, 1.0, 1.0), (0.0, 0.0, 0.0)), ((1.0, 1.0, 1.0), (0.1, 0.0, 0.2))])
def test_corrected_solve_is_the_discrete_stokes_solution(aV, bV):
    """Influence matrix + B problem + sigma produce the unique (P, u) of the dense Stokes system."""
    cfg = small_config(nys=12, a_v=aV, b_v=bV)
    o = pyoracle.Oracle(cfg)
    g = 2.0e4
    o.setup(g)
    t = o.tables()
    rp, rv, bp, bv = _rand(o, 9)
    P, V = o.solve_modes(None, bp, rv, bv, correct_sigma=True)
    n = cfg["NYS"]
    wn = o.table("wn_v")
    checked = 0
    for izi in range(o.nzk):
        for kx in range(o.nkx):
            if izi == cfg["NZS"] and kx == 0:
                continue
            R = [rv[izi, :, kx, c] for c in range(3)]
            vb = [[bv[izi, 0, kx, c] for c in range(3)], [bv[izi, 1, kx, c] for c in range(3)]]
            Pd, Vd = dense_coupled(n, t["WAVX"][kx], t["WAVZ"][izi], wn[izi, kx], cfg["Re"], R, None, vb, aV, bV)
            scale = max(np.abs(Vd).max(), 1e-300)
            assert np.abs(V[izi, :, kx, :] - Vd).max() < 1e-9 * scale, (izi, kx)
            assert np.abs(P[izi, :, kx] - Pd).max() < 1e-9 * max(np.abs(Pd).max(), 1e-300), (izi, kx)
            checked += 1
    assert checked == o.nzk * o.nkx - 1


def test_sigma_correction_is_needed_for_exact_divergence():
    """Without sigma the velocity is divergence-free only at the walls; with it, everywhere."""
    cfg = config.channel_32()
    o = pyoracle.Oracle(cfg)
    o.setup(1.5e3 * 180)
    rp, rv, bp, bv = _rand(o, 2)
    bv[...] = 0
    root = (cfg["NZS"], slice(None), 0)
    P, V = o.solve_modes(None, bp, rv, bv, correct_sigma=False)
    div = o.divergence(V)
    div[root] = 0
    db = o.divergence_boundary(V)
    db[cfg["NZS"], :, 0] = 0
    assert np.abs(db).max() < 1e-15
    assert np.abs(div).max() > 1e-8
    P, V = o.solve_modes(None, bp, rv, bv, correct_sigma=True)
    div = o.divergence(V)
    div[root] = 0
    assert np.abs(div).max() < 1e-14
