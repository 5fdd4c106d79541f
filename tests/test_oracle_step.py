"""Pins of the oracle's time step (src/sim/simulation.cpp, src/solver/timestepper.cpp)."""
import numpy as np
import pytest

from cdns import config, fields
from helpers import stage3_config
import pyoracle


def test_step_coefficients_follow_startup_scheme():
    """tests/test_simulation.cpp:5-20 (exact)."""
    first = pyoracle.Oracle.step_coefficients(1, 0.25, 10.0)
    assert list(first) == [4.0, 0.0, 1.0, 0.0, 40.0]
    second = pyoracle.Oracle.step_coefficients(2, 0.25, 10.0)
    assert list(second[:4]) == [8.0, -2.0, 2.0, -1.0]
    third = pyoracle.Oracle.step_coefficients(3, 0.25, 10.0)
    assert list(third) == list(second)


def test_zero_step_stays_finite_and_projects_mean_wall_normal_velocity():
    """tests/test_timestepper.cpp:21-54: the mean v becomes the wall value, other coefficients 0."""
    cfg = stage3_config()
    o = pyoracle.Oracle(cfg)
    o.step(1)
    v = o.get_state("velocity")
    assert np.isfinite(v).all()
    assert v[16, 0, 0, 1] == cfg["V_Microblow0"] + 0j
    assert (v[16, 1:, 0, 1] == 0).all()


def test_laminar_poiseuille_is_a_steady_state():
    """u = (Re/2)(1 - y^2) with dp/dx = -1, nu = 1/Re is an exact steady solution; the discrete step
    must keep it (to round-off) for many steps: nonlinear term, pressure and solves all consistent."""
    cfg = config.channel_32()
    o = pyoracle.Oracle(cfg)
    u0 = o.vector()
    uc = cfg["Re"] / 2
    u0[16, 0, 0, 0] = uc / 2
    u0[16, 2, 0, 0] = -uc / 2
    o.set_state("velocity", u0)
    o.run(10)
    u = o.get_state("velocity")
    assert np.abs(u - u0).max() < 1e-11 * uc
    assert np.allclose(o.flow_gradients(), [-1.0, -0.0])


def test_constant_flux_mode_holds_the_flow_rate():
    """INDEX_GRAD_PRE=0: dpdx is chosen so the mean streamwise flow rate equals QXmean=cos(THETA)."""
    cfg = stage3_config()
    o = pyoracle.Oracle(cfg)
    u0 = fields.seeded_velocity(cfg, seed=2)
    o.set_state("velocity", u0)
    o.run(3)
    u = o.get_state("velocity")
    assert abs(o.flow_rate(u, 0) - 1.0) < 1e-12
    assert abs(o.flow_rate(u, 2)) < 1e-12


def test_step_keeps_divergence_free_no_slip_fields():
    cfg = config.channel_32()
    o = pyoracle.Oracle(cfg)
    u0 = fields.seeded_velocity(cfg, seed=7)
    o.set_state("velocity", u0)
    o.run(5)
    u = o.get_state("velocity")
    div = o.divergence(u)
    div[16, :, 0] = 0
    assert np.abs(div).max() < 1e-12 * np.abs(u).max()
    assert np.abs(u.sum(axis=1)).max() < 1e-12 * np.abs(u).max()
    assert np.abs((u * ((-1.0) ** np.arange(u.shape[1]))[None, :, None, None]).sum(axis=1)).max() < 1e-12 * np.abs(u).max()
    assert fields.rel_l2(u, u0) > 1e-6  # the field evolves


def test_history_buffers_follow_the_reference_swaps():
    """After a step: velocity_prev = u^n, nonlinear_prev = N(u^n), pressure_bc_prev = B^n
    (simulation.cpp:161-170)."""
    cfg = config.channel_32()
    o = pyoracle.Oracle(cfg)
    u0 = fields.seeded_velocity(cfg, seed=8)
    o.set_state("velocity", u0)
    o.step(1)
    assert np.array_equal(o.get_state("velocity_prev"), u0)
    ref = pyoracle.Oracle(cfg)
    nl, bc = ref.nonlinear_term(u0)
    assert np.array_equal(o.get_state("nonlinear_prev"), nl)
    assert np.array_equal(o.get_state("pressure_bc_prev"), bc)


def test_mt19937_64_known_answer():
    """The 10000th output of std::mt19937_64 with the default seed 5489 is 9981545732273789042."""
    g = pyoracle.MT19937_64(5489)
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042
