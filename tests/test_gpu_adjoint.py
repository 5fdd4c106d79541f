"""The HIP discrete adjoint (no reference counterpart) against the oracle's adjoint, which the
CPU suite pins by dot-product and finite-difference tests; plus finite differences of the GPU
forward path itself and the torch.autograd bridge."""
import math

import numpy as np
import pytest

from cdns import config, fields
import pyoracle

pytestmark = pytest.mark.gpu

STATE = ["velocity", "velocity_prev", "nonlinear_prev", "pressure_bc_prev"]


def rdot(a, b):
    return float(np.sum(np.real(np.conj(a) * b)))


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _config(name):
    if name == "channel_32":
        return config.channel_32()
    if name == "long_columns":
        # NYS = 150: four entries of each parity chain per lane pair in the adjoint solve
        return config.make(16, 160, 16, 8, 150, 8, 24, 240, 24, 2 * math.pi, math.pi, 180.0, 5e-4)
    return config.reference_example()


@pytest.mark.parametrize("name", ["channel_32", "stage3", "long_columns"])
def test_adjoint_step_matches_oracle(gpu, name):
    """One transposed step (iterations 5, 1, 2) on the same state and adjoint inputs."""
    cfg = _config(name)
    rng = np.random.default_rng(21)
    o = pyoracle.Oracle(cfg)
    A = pyoracle.Adjoint(o)
    o.set_state("velocity", fields.seeded_velocity(cfg, seed=5))
    for it in range(1, 5):
        o.step(it)
    X = [o.get_state(n) for n in STATE]
    dev = gpu.Channel(cfg)
    for it in (5, 1, 2):
        L = [crandn(rng, a.shape) for a in X]
        lf0 = crandn(rng, X[0].shape)
        Lr, lfr = A.adj_step(it, X, L, lf0)
        Lg, lfg = dev.adjoint_step(it, X[0], L, lf0)
        # errors relative to the larger of the output and the incoming adjoint of the same slot:
        # the adjoint of B_prev is round-off (the influence-matrix correction makes the new
        # velocity independent of the provisional pressure wall value), and the solve's
        # transposed RHS carries the 1e-10 conditioning of the forward solve_modes parity
        errs = [float(np.abs(g - r).max() / max(np.abs(r).max(), np.abs(i).max()))
                for g, r, i in zip(Lg + [lfg], Lr + [lfr], L + [lf0])]
        print(f"{name} it={it} adjoint step errors (u, u_prev, N_prev, B_prev, f): {errs}; "
              f"max |adj B_prev| {np.abs(Lr[3]).max():.2e}")
        assert max(errs) < 1e-10, errs


def test_adjoint_run_matches_oracle(gpu):
    """10 steps of 32x33x32 with a body force, checkpoints every 4: J and both gradients."""
    cfg = config.channel_32()
    o = pyoracle.Oracle(cfg)
    A = pyoracle.Adjoint(o)
    u0 = fields.seeded_velocity(cfg, seed=5, mean="log", spectrum="k53", amplitude=0.1, kcut=8)
    f = 0.01 * fields.seeded_velocity(cfg, seed=7, mean="laminar", kmax=4, amplitude=1.0)
    f[cfg["NZS"], :, 0, :] = 0
    dev = gpu.Channel(cfg)
    for objective in (0, 1):
        Jr, gur, gfr = A.run(u0, f, 10, 4, objective)
        Jg, gug, gfg = dev.adjoint_run(u0, f, 10, 4, objective)
        print(f"objective {objective}: J {Jg} vs {Jr}; grad u0 rel {rel(gug, gur)}, grad f rel {rel(gfg, gfr)}")
        assert math.isclose(Jg, Jr, rel_tol=1e-11)
        assert rel(gug, gur) < 1e-10 and rel(gfg, gfr) < 1e-10


def _J_forward(dev, u, f, nsteps, objective):
    from cdns.autograd import forward_objective

    return forward_objective(dev, u, f, nsteps, objective)


def _smooth_direction(cfg, seed, scale):
    """A seeded divergence-free band-limited field (|k| <= 4, vanishing at the walls) without
    mean mode: at dt = 1e-3 on this grid, directions with energy at the highest kx, kz grow by
    ~1.4 per step under the explicit advection (the oracle behaves the same), so finite
    differences along them measure that instability, not the gradient."""
    v = fields.seeded_velocity(cfg, seed=seed, mean="laminar", kmax=4, amplitude=1.0)
    v[cfg["NZS"], :, 0, :] = 0
    return v * (scale / np.abs(v).max())


def _central_differences(Jof, eps):
    """The two-point and the four-point central differences of J along a direction."""
    p1, m1, p2, m2 = Jof(eps), Jof(-eps), Jof(2 * eps), Jof(-2 * eps)
    return (p1 - m1) / (2 * eps), (8 * (p1 - m1) - (p2 - m2)) / (12 * eps)


def _fd_disagreement(fd, adj, J, eps):
    """|fd - adj| relative to the directional derivative itself (the larger of the two), floored
    at the absolute round-off of the central difference, about |J| 1e-16 / eps."""
    return abs(fd - adj) / max(abs(fd), abs(adj), abs(J) * 1e-16 / eps)


def test_adjoint_gradient_c4_horizon_vs_gpu_finite_differences(gpu):
    """The adjoint workload (64x65x64, product 96x97x96, 2pi x 2 x pi): 200 steps with a
    checkpoint every 20, gradients of the terminal kinetic energy w.r.t. the initial velocity and
    a body force (seeded smooth, amplitude 0.01); central differences (eps = 1e-6) of the GPU
    forward path in 16 seeded random directions (every retained coefficient, both parts, of the
    size of the largest velocity coefficient; 10 times that for the force), each within 1e-6 of
    the directional derivative.  The difference is the four-point central one, (8 (J(e) - J(-e)) -
    (J(2e) - J(-2e))) / 12e, whose own error is O(eps^4): along some random directions the
    two-point one is off by up to 4e-6 from its O(eps^2) term alone (printed beside).  dt =
    2.5e-4: with this k^-5/3 field the 200-step horizon overflows at dt = 1e-3 (the explicitly
    advected highest modes grow), and at 5e-4 their growth makes the forward map so curved along
    a random direction that both differences are off by 1e-4 to 1e-3 at this eps (measured on the
    MI355X)."""
    cfg = dict(config.channel_64(), DT=2.5e-4)
    u0 = fields.seeded_velocity(cfg, seed=64, mean="log", spectrum="k53", amplitude=0.1)
    f = 0.01 * fields.seeded_velocity(cfg, seed=65, mean="laminar", kmax=4, amplitude=1.0)
    f[cfg["NZS"], :, 0, :] = 0
    dev = gpu.Channel(cfg)
    nsteps, kc = 200, 20
    J, gu, gf = dev.adjoint_run(u0, f, nsteps, kc, 0)
    assert math.isclose(J, _J_forward(dev, u0, f, nsteps, 0), rel_tol=1e-13)
    rng = np.random.default_rng(2004)
    scale = np.abs(u0).max()
    eps = 1e-6
    worst = 0.0
    worst2 = 0.0
    for trial in range(16):
        v = crandn(rng, u0.shape) * scale
        w = crandn(rng, u0.shape) * (10 * scale)
        d_u = _central_differences(lambda e: _J_forward(dev, u0 + e * v, f, nsteps, 0), eps)
        d_f = _central_differences(lambda e: _J_forward(dev, u0, f + e * w, nsteps, 0), eps)
        eu = _fd_disagreement(d_u[1], rdot(gu, v), J, eps)
        ef = _fd_disagreement(d_f[1], rdot(gf, w), J, eps)
        worst = max(worst, eu, ef)
        worst2 = max(worst2, _fd_disagreement(d_u[0], rdot(gu, v), J, eps), _fd_disagreement(d_f[0], rdot(gf, w), J, eps))
        print(f"direction {trial}: u0 fd {d_u[1]:.10e} adj {rdot(gu, v):.10e} err {eu:.2e}; "
              f"f fd {d_f[1]:.10e} adj {rdot(gf, w):.10e} err {ef:.2e}", flush=True)
    print(f"worst disagreement over 16 directions: {worst:.2e} (two-point difference: {worst2:.2e})")
    assert worst < 1e-6


def test_adjoint_gradient_smooth_directions_dt_1e3(gpu):
    """The same grid at its nominal dt = 1e-3 over 40 steps (checkpoint every 20), along seeded
    smooth directions, with the four-point central difference at eps = 1e-5: at this dt the
    rounding of the forward map is amplified by the growing highest modes, so at eps = 1e-6 the
    differences carry about 1.3e-6 of noise (two- and four-point alike, measured on the MI355X),
    ten times less at 1e-5, where the O(eps^4) term is still far below."""
    cfg = config.channel_64()
    u0 = fields.seeded_velocity(cfg, seed=64, mean="log", spectrum="k53", amplitude=0.1, kcut=8)
    f = 0.01 * fields.seeded_velocity(cfg, seed=65, mean="laminar", kmax=4, amplitude=1.0)
    f[cfg["NZS"], :, 0, :] = 0
    dev = gpu.Channel(cfg)
    nsteps, kc = 40, 20
    J, gu, gf = dev.adjoint_run(u0, f, nsteps, kc, 0)
    scale = np.abs(u0).max()
    eps = 1e-5
    worst = 0.0
    for trial in range(4):
        v = _smooth_direction(cfg, 1000 + trial, scale)
        w = _smooth_direction(cfg, 2000 + trial, 10 * scale)
        fd_u = _central_differences(lambda e: _J_forward(dev, u0 + e * v, f, nsteps, 0), eps)[1]
        fd_f = _central_differences(lambda e: _J_forward(dev, u0, f + e * w, nsteps, 0), eps)[1]
        worst = max(worst, _fd_disagreement(fd_u, rdot(gu, v), J, eps), _fd_disagreement(fd_f, rdot(gf, w), J, eps))
    print(f"worst disagreement over 4 smooth directions: {worst:.2e}")
    assert worst < 1e-6


def _step_map(dev, it, X, f):
    """The GPU forward step `it` as a map of the state (u, u_prev, N_prev, B_prev) and the body
    force f to the state after it."""
    dev.set_body_force(f)
    for n, x in zip(STATE, X):
        dev.set_state(n, x)
    dev.step(it)
    return [dev.get_state(n) for n in STATE]


@pytest.mark.parametrize("it", [1, 2, 5])
def test_adjoint_step_dot_product(gpu, it):
    """<J dX, L> = <dX, J^T L> for one step of 32x33x32, with J dX the tangent of the GPU forward
    step by central differences (exact up to round-off: the step is quadratic in u and linear in
    the history and the force) and J^T L from cdns_adjoint_step; and the same identity with the
    oracle's tangent-linear step."""
    cfg = config.channel_32()
    rng = np.random.default_rng(40 + it)
    o = pyoracle.Oracle(cfg)
    o.set_state("velocity", fields.seeded_velocity(cfg, seed=5))
    for k in range(1, it):
        o.step(k)
    X = [o.get_state(n) for n in STATE]
    f0 = 0.01 * fields.seeded_velocity(cfg, seed=7, mean="laminar", kmax=4, amplitude=1.0)
    dev = gpu.Channel(cfg)
    dev.run(2, it_begin=1)  # solver set up
    dX = [crandn(rng, x.shape) * max(np.abs(x).max(), 1.0) for x in X]
    df = crandn(rng, f0.shape) * np.abs(f0).max()
    L = [crandn(rng, x.shape) for x in X]
    h = 1e-4
    plus = _step_map(dev, it, [x + h * d for x, d in zip(X, dX)], f0 + h * df)
    minus = _step_map(dev, it, [x - h * d for x, d in zip(X, dX)], f0 - h * df)
    JdX = [(p - m) / (2 * h) for p, m in zip(plus, minus)]
    dev.set_body_force(f0)
    LT, lf = dev.adjoint_step(it, X[0], L, np.zeros_like(f0))
    lhs = sum(rdot(a, b) for a, b in zip(JdX, L))
    rhs = sum(rdot(a, b) for a, b in zip(dX, LT)) + rdot(df, lf)
    terms = sum(np.linalg.norm(a) * np.linalg.norm(b) for a, b in zip(JdX, L))
    print(f"it={it}: <J dX, L> {lhs:.12e}  <dX, J^T L> {rhs:.12e}  rel {abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) <= 1e-9 * max(abs(lhs), 1e-3 * terms)
    # the oracle's tangent-linear step (pinned against central differences on the CPU)
    A = pyoracle.Adjoint(o)
    o.set_body_force(f0)
    tl = A.tl_step(it, X, dX, df)
    lhs_o = sum(rdot(a, b) for a, b in zip(tl, L))
    assert abs(lhs_o - rhs) <= 1e-10 * max(abs(lhs_o), 1e-3 * terms), (lhs_o, rhs)


def test_torch_autograd_bridge(gpu):
    import torch

    from cdns.autograd import terminal_objective

    cfg = config.channel_32()
    u0 = fields.seeded_velocity(cfg, seed=5, mean="log", spectrum="k53", amplitude=0.1, kcut=8)
    f = 0.01 * fields.seeded_velocity(cfg, seed=7, mean="laminar", kmax=4, amplitude=1.0)
    dev = gpu.Channel(cfg)
    ut = torch.tensor(u0.view(np.float64).reshape(u0.shape + (2,)), device="cuda", requires_grad=True)
    ft = torch.tensor(f.view(np.float64).reshape(f.shape + (2,)), device="cuda", requires_grad=True)
    J = terminal_objective(dev, ut, ft, nsteps=8, k_ckpt=3, objective=1)
    gu, gf = torch.autograd.grad(J, (ut, ft))
    Jr, gur, gfr = dev.adjoint_run(u0, f, 8, 3, 1)
    assert math.isclose(J.item(), Jr, rel_tol=1e-13)
    gu_c = gu.detach().cpu().numpy().view(np.complex128)[..., 0]
    gf_c = gf.detach().cpu().numpy().view(np.complex128)[..., 0]
    assert rel(gu_c, gur) < 1e-14 and rel(gf_c, gfr) < 1e-14


def test_retau395_mixed_500_steps_and_wall_shear_adjoint(gpu):
    """The Re_tau=395 case (256x193x192, product 384x289x288): 500 steps in the mixed form against
    the FP64 form (velocity rel-L2, mean profile, wall shear, divergence), and the FP64 gradient of
    the terminal mean wall shear through the same 500 steps with a checkpoint every 25, against
    central differences (eps = 1e-6) of the GPU forward path in two seeded directions.  dt = 5e-5
    and the perturbation band-limited to |k| <= 16 (the cost of a step does not depend on either):
    the explicitly advected highest kx modes are unstable at larger steps (at 5e-4 omega dt ~ 1.4;
    at 1e-4 the full-band FP64 field overflows between steps 200 and 300, and the band-limited
    mixed form, whose FP32 rounding seeds those modes, departs from FP64 between steps 200 and 500,
    measured on the MI355X), so 500 steps there measure that growth, not the precision."""
    cfg = dict(config.retau395(), DT=5e-5)
    nsteps, kc = 500, 25
    u0 = fields.seeded_velocity(cfg, seed=395, mean="log", spectrum="k53", amplitude=0.1, kcut=16)
    dev = gpu.Channel(cfg)
    dev.set_state("velocity", u0)
    dev.run(nsteps, it_begin=1)
    f64 = dev.get_state("velocity")
    mx = gpu.Channel(cfg, precision=gpu.MIXED)
    mx.set_state("velocity", u0)
    mx.run(nsteps, it_begin=1)
    mixed = mx.get_state("velocity")
    mx.close()
    rep = fields.precision_report(cfg, mixed, f64)
    print(f"retau395 mixed vs FP64 after {nsteps} steps: {rep}", flush=True)
    assert np.isfinite(mixed).all() and np.isfinite(f64).all()
    assert max(rep["velocity_rel_l2"]) < 1e-5
    assert rep["mean_profile_rel_l2"] < 1e-6 and max(rep["wall_shear_rel"]) < 1e-6
    J, gu, _ = dev.adjoint_run(u0, None, nsteps, kc, 1)
    assert math.isclose(J, _J_forward(dev, u0, None, nsteps, 1), rel_tol=1e-12)
    assert np.isfinite(gu).all()
    scale = np.abs(u0).max()
    eps = 1e-6
    worst = 0.0
    for trial in range(2):
        v = _smooth_direction(cfg, 3950 + trial, scale)
        fd = (_J_forward(dev, u0 + eps * v, None, nsteps, 1) - _J_forward(dev, u0 - eps * v, None, nsteps, 1)) / (2 * eps)
        e = abs(fd - rdot(gu, v)) / (np.linalg.norm(gu) * np.linalg.norm(v))
        worst = max(worst, e)
        print(f"retau395 direction {trial}: fd {fd:.10e} adj {rdot(gu, v):.10e} err {e:.2e}", flush=True)
    assert worst < 1e-6
