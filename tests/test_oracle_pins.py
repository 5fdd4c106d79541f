"""Pins of the oracle's mean-mode pressure and nonlinear pressure wall condition by mathematics.

Mean-mode (kx = kz = 0) pressure.  The stored pressure is P = p + |u|^2/2 (rotational form,
doc/algorithm.md:25) with the gauge P_0 = 0 (eqsolve_pv.cpp:19-27, 51-54).  A plane-mean flow
U(y) e_x + W(y) e_z with no fluctuations has u x curl(u) = (0, U U' + W W', 0) and a uniform
mean v (zero, or the wall velocity), so the wall-normal momentum of the mean mode reduces to dP/dy = N_y = d/dy (U^2 +
W^2)/2: the pressure is the Chebyshev series of (U^2 + W^2)/2 with its T_0 coefficient removed.
This runs through the root-mode influence special case (eqsolve_pv.cpp:288-300), the sigma
denominator override (eqsolve_pv.cpp:476) and the gauge branch of the tau column; the expected
series comes from numpy's Chebyshev product, independent of the oracle.

Pressure wall condition (timestepper.cpp:26-43).  B(+1) = sum_j [i kx w_z/Re - i kz w_x/Re + N_y]_j
and B(-1) = sum_j (-1)^(j+1) [...]_j.  For a divergence-free u, i kx w_z - i kz w_x =
-(curl curl u)_y = lap v, so B(+1) is the value at y = +1 of lap(v)/Re + N_y and B(-1) minus its
value at y = -1.  lap v is formed here with numpy's Chebyshev derivative, not the oracle's, so a
dropped term, a sign error in either wall sum or a wrong wall parity fails the comparison.
"""
import numpy as np
from numpy.polynomial import chebyshev as C

from cdns import config, fields
import pyoracle


def _mean_only(cfg, U, W):
    """A plane-mean state: U(y), W(y) and the uniform wall-normal velocity V0 that the wall
    velocities impose on the mean mode (correct_zero_mode, timestepper.cpp:60-68)."""
    vel = np.zeros((2 * cfg["NZS"] + 1, cfg["NYS"] + 1, cfg["NXS"] + 1, 3), np.complex128)
    vel[cfg["NZS"], : len(U), 0, 0] = U
    vel[cfg["NZS"], 0, 0, 1] = cfg["V_Microblow0"]
    vel[cfg["NZS"], : len(W), 0, 2] = W
    return vel


def _expected_mean_pressure(nj, U, W):
    half = 0.5 * (C.chebadd(C.chebmul(U, U), C.chebmul(W, W)))
    want = np.zeros(nj)
    want[: len(half)] = half
    want[0] = 0.0
    return want


def test_mean_pressure_of_steady_poiseuille_flow():
    """Laminar Poiseuille flow u = (Re/2)(1 - y^2) under dp/dx = -1 is steady; after 10 steps the
    mean-mode pressure is U^2/2 without its T_0 part."""
    cfg = config.channel_32()
    o = pyoracle.Oracle(cfg)
    prof, _ = fields.mean_profile(cfg, "laminar")
    U = prof[:3].copy()
    o.set_state("velocity", _mean_only(cfg, U, np.zeros(1)))
    o.run(10)
    u = o.get_state("velocity")
    assert np.abs(u[cfg["NZS"], :3, 0, 0].real - U).max() < 1e-10 * np.abs(U).max()  # steady
    P = o.get_state("pressure")
    want = _expected_mean_pressure(cfg["NYS"] + 1, U, np.zeros(1))
    err = np.abs(P[cfg["NZS"], :, 0] - want).max()
    print(f"Poiseuille mean pressure: max |P - U^2/2| = {err:.2e} (|U^2/2| {np.abs(want).max():.3e})")
    assert err < 1e-12 * np.abs(want).max()


def test_mean_pressure_of_a_general_mean_flow():
    """Polynomial U(y), W(y) vanishing at the walls, odd and even parts both present: after the
    first step (whose RHS carries N(u^0) alone, b0 = 1, b1 = 0) the mean pressure is
    (U^2 + W^2)/2 of the initial profiles without the T_0 part; also with constant flux and a
    blowing wall (the reference example), whose mean v is the wall-velocity average."""
    for cfg in (config.channel_32(), config.reference_example()):
        U = C.chebmul(C.chebfromroots([-1.0, 1.0]), [-3.0, 1.2, 0.5])  # (y^2 - 1)(-3 + 1.2 y + 0.5 T2)
        W = C.chebmul(C.chebfromroots([-1.0, 1.0]), [0.0, 0.7, 0.0, -0.4])
        o = pyoracle.Oracle(cfg)
        o.set_state("velocity", _mean_only(cfg, U, W))
        o.step(1)
        P = o.get_state("pressure")[cfg["NZS"], :, 0]
        # a uniform mean v = V0 adds nothing to N_y = W W' + U U' (its vorticity is zero), and
        # its |u|^2/2 share is a constant that the gauge removes
        want = _expected_mean_pressure(cfg["NYS"] + 1, U, W)
        err = np.abs(P - want).max()
        print(f"mean pressure ({cfg['NX']}^3, grad_pre {cfg['INDEX_GRAD_PRE']}): max error {err:.2e}")
        assert err < 1e-12 * np.abs(want).max()


def _lap_v_plus_ny_at_walls(cfg, vel, nl):
    kx = np.arange(cfg["NXS"] + 1) * 2 * np.pi / cfg["XL"]
    kz = np.arange(-cfg["NZS"], cfg["NZS"] + 1) * 2 * np.pi / cfg["ZL"]
    v = vel[..., 1]
    d2v = np.zeros_like(v)
    d2 = C.chebder(np.moveaxis(v, 1, 0), 2)  # Chebyshev index first
    d2v[:, : d2.shape[0], :] = np.moveaxis(d2, 0, 1)
    lap = d2v - (kx[None, None, :] ** 2 + kz[:, None, None] ** 2) * v
    f = lap / cfg["Re"] + nl[..., 1]
    sgn = (-1.0) ** np.arange(cfg["NYS"] + 1)
    return f.sum(axis=1), (f * sgn[None, :, None]).sum(axis=1)


def test_pressure_wall_condition_is_wall_value_of_lap_v_plus_ny():
    for cfg, seed in ((config.channel_32(), 3), (config.channel_64(), 4)):
        o = pyoracle.Oracle(cfg)
        vel = fields.seeded_velocity(cfg, seed=seed, mean="log", spectrum="k53", amplitude=0.2)
        # the seeded field is divergence-free by construction
        assert fields.divergence_l2(cfg, vel) < 1e-10 * np.linalg.norm(vel)
        nl, bc = o.nonlinear_term(vel)
        top, bottom = _lap_v_plus_ny_at_walls(cfg, vel, nl)
        scale = np.abs(top).max()
        e0 = np.abs(bc[:, 0, :] - top).max() / scale
        e1 = np.abs(bc[:, 1, :] + bottom).max() / scale
        print(f"{cfg['NX']}^3: pressure wall condition vs lap v / Re + N_y: {e0:.2e} (y=+1), {e1:.2e} (y=-1)")
        assert e0 < 1e-12 and e1 < 1e-12
        # the identity is not vacuous: both parts carry weight at the walls
        lap_only = _lap_v_plus_ny_at_walls(cfg, vel, np.zeros_like(nl))[0]
        assert np.abs(lap_only).max() > 1e-3 * scale and np.abs(top - lap_only).max() > 1e-3 * scale
