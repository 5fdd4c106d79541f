"""CPU pins of the oracle's tangent-linear and discrete adjoint of the step (no reference
counterpart exists).  Each transposed piece passes a dot-product test against the forward
piece; the tangent-linear step matches central finite differences of the forward step; the
adjoint step passes the dot-product test against the tangent-linear step; and horizon
gradients (with checkpointing) match central finite differences (eps = 1e-6) to 1e-6."""
import math

import numpy as np
import pytest

from cdns import config, fields
from helpers import ref_transform_config
import pyoracle

STATE = ["velocity", "velocity_prev", "nonlinear_prev", "pressure_bc_prev"]


def rdot(a, b):
    return float(np.sum(np.real(np.conj(a) * b)))


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.fixture(scope="module", params=["ref32_full", "channel_32", "stage3"])
def pair(request):
    cfg = {"ref32_full": ref_transform_config(16, 30, 16), "channel_32": config.channel_32(),
           "stage3": config.reference_example()}[request.param]
    o = pyoracle.Oracle(cfg)
    return request.param, cfg, o, pyoracle.Adjoint(o)


def test_transposed_transforms_and_curl(pair):
    name, cfg, o, A = pair
    rng = np.random.default_rng(11)
    u = crandn(rng, o.scalar().shape)  # arbitrary, not Hermitian: every array entry is an input
    lam = rng.standard_normal(o.fine_phys().shape)
    p = o.fine_phys()
    o.xyzfft2("B", p, u.copy())
    lhs, rhs = float(np.sum(p * lam)), rdot(u, A.fine_backward_T(lam))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)
    s = o.scalar()
    o.xyzfft2("F", lam.copy(), s)
    g = crandn(rng, s.shape)
    lhs, rhs = rdot(g, s), float(np.sum(lam * A.fine_forward_T(g)))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)
    v, lw = crandn(rng, o.vector().shape), crandn(rng, o.vector().shape)
    lhs, rhs = rdot(lw, o.curl(v)), rdot(A.curl_T(lw), v)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


def test_transposed_solve(pair):
    """solve_modes (fused RHS, influence and sigma corrections) is linear in (RHS, pressure
    wall values) once the velocity wall values are zero; its transpose passes the dot test."""
    name, cfg, o, A = pair
    rng = np.random.default_rng(12)
    o.setup(1.5 / cfg["DT"] * cfg["Re"])
    r, bc = crandn(rng, o.vector().shape), crandn(rng, o.boundary().shape)
    _, V = o.solve_modes(None, bc, r, o.boundary(3), True)
    lv = crandn(rng, V.shape)
    lr, lbc = A.solve_T(lv)
    lhs, rhs = rdot(lv, V), rdot(lr, r) + rdot(lbc, bc)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("name", ["channel_32", "stage3"])
def test_tangent_linear_and_adjoint_step(name):
    """Laminar-mean field after 4 steps, with a body force; iterations 5 (BDF2), 1 and 2 (the
    first-order start and the switch); stage3 = constant flux with a blowing wall."""
    cfg = config.channel_32() if name == "channel_32" else config.reference_example()
    rng = np.random.default_rng(13)
    o = pyoracle.Oracle(cfg)
    A = pyoracle.Adjoint(o)
    f = 0.01 * crandn(rng, o.vector().shape)
    o.set_body_force(f)
    o.set_state("velocity", fields.seeded_velocity(cfg, seed=5))
    for it in range(1, 5):
        o.step(it)
    X = [o.get_state(n) for n in STATE]
    dX = [crandn(rng, a.shape) * np.abs(a).max() for a in X]
    df = 0.01 * crandn(rng, f.shape)
    for it in (5, 1, 2):
        dXo = A.tl_step(it, X, dX, df)
        eps = 1e-6
        outs = []
        for sgn in (1.0, -1.0):
            for n, a, d in zip(STATE, X, dX):
                o.set_state(n, a + sgn * eps * d)
            o.set_body_force(f + sgn * eps * df)
            if it >= 3:
                o.setup(pyoracle.Oracle.step_coefficients(2, cfg["DT"], cfg["Re"])[4])
            o.step(it)
            outs.append([o.get_state(n) for n in STATE])
        o.set_body_force(f)
        for a, p, m in zip(dXo, *outs):
            fd = (p - m) / (2 * eps)
            assert np.abs(a - fd).max() <= 1e-8 * np.abs(fd).max(), (it, np.abs(a - fd).max() / np.abs(fd).max())
        L = [crandn(rng, a.shape) for a in X]
        Lo, lf = A.adj_step(it, X, L)
        lhs = sum(rdot(l, d) for l, d in zip(L, dXo))
        rhs = sum(rdot(l, d) for l, d in zip(Lo, dX)) + rdot(lf, df)
        assert abs(lhs - rhs) <= 1e-11 * abs(lhs), (it, lhs, rhs)


def test_objectives_closed_forms():
    """KE of u = c T0 on the mean mode is c^2 XL ZL (int_{-1}^{1} dy = 2, times 1/2); KE of
    u = c T2 is (1/2) c^2 XL ZL (14/15); the shear objective of the laminar profile is Re."""
    cfg = config.channel_32()
    o = pyoracle.Oracle(cfg)
    A = pyoracle.Adjoint(o)
    u = o.vector()
    u[cfg["NZS"], 0, 0, 0] = 3.0
    J, g = A.objective(0, u)
    assert math.isclose(J, 9.0 * cfg["XL"] * cfg["ZL"], rel_tol=1e-14)
    u[:] = 0
    u[cfg["NZS"], 2, 0, 2] = 2.0
    J, _ = A.objective(0, u)
    assert math.isclose(J, 0.5 * 4.0 * cfg["XL"] * cfg["ZL"] * (14.0 / 15.0), rel_tol=1e-14)
    lam = fields.seeded_velocity(cfg, seed=1, mean="laminar", kmax=4, amplitude=0.0)
    J, _ = A.objective(1, lam)
    top, bottom = fields.wall_shear(cfg, lam)
    assert math.isclose(J, 0.5 * (bottom - top), rel_tol=1e-13)
    assert math.isclose(J, cfg["Re"], rel_tol=1e-13)


@pytest.mark.parametrize("objective", [0, 1])
def test_horizon_gradient_vs_finite_differences(objective):
    """10 steps of 32x33x32 with checkpoints every 4 (segments 4, 4, 2), gradients of the
    terminal objective w.r.t. the initial velocity and a body force, central differences with
    eps = 1e-6 along seeded directions scaled to the field."""
    cfg = config.channel_32()
    rng = np.random.default_rng(14 + objective)
    o = pyoracle.Oracle(cfg)
    A = pyoracle.Adjoint(o)
    u0 = fields.seeded_velocity(cfg, seed=5, mean="log", spectrum="k53", amplitude=0.1, kcut=8)
    f = 0.01 * fields.seeded_velocity(cfg, seed=7, mean="laminar", kmax=4, amplitude=1.0)
    f[cfg["NZS"], :, 0, :] = 0
    nsteps, kc = 10, 4
    J, gu, gf = A.run(u0, f, nsteps, kc, objective)

    def J_of(u, ff):
        o.set_body_force(ff)
        for n, z in zip(STATE[1:], (o.vector(), o.vector(), o.boundary())):
            o.set_state(n, z)
        o.set_state("velocity", u)
        for it in range(1, nsteps + 1):
            o.step(it)
        return A.objective(objective, o.get_state("velocity"))[0]

    assert math.isclose(J, J_of(u0, f), rel_tol=1e-14)
    # directions long enough that the round-off floor of the differences, about
    # |J| 1e-16 / eps, stays below 1e-6 of the directional derivative
    scale = np.abs(u0).max()
    eps = 1e-6
    for _ in range(2):
        v, w = crandn(rng, u0.shape) * scale, crandn(rng, u0.shape) * 10 * scale
        fd_u = (J_of(u0 + eps * v, f) - J_of(u0 - eps * v, f)) / (2 * eps)
        fd_f = (J_of(u0, f + eps * w) - J_of(u0, f - eps * w)) / (2 * eps)
        assert abs(fd_u - rdot(gu, v)) <= 1e-6 * abs(fd_u), (fd_u, rdot(gu, v))
        assert abs(fd_f - rdot(gf, w)) <= 1e-6 * abs(fd_f), (fd_f, rdot(gf, w))
