"""Mixed precision (FP32 transforms and nonlinear term, FP64 state, RHS and solves) against
the FP64 path: velocity rel-L2 per component within 1e-5, with the mean profile, wall shear
and divergence reported beside it."""
import numpy as np
import pytest

from cdns import config, fields

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-5


def _run(gpu, cfg, u0, steps, precision):
    dev = gpu.Channel(cfg, precision=precision)
    dev.set_state("velocity", u0)
    dev.run(steps, it_begin=1)
    out = dev.get_state("velocity")
    dev.close()
    return out


def _check(gpu, label, cfg, u0, steps):
    ref = _run(gpu, cfg, u0, steps, gpu.FP64)
    got = _run(gpu, cfg, u0, steps, gpu.MIXED)
    assert np.isfinite(ref).all() and np.isfinite(got).all()
    rep = fields.precision_report(cfg, got, ref)
    print(f"{label}: mixed vs fp64 after {steps} steps: {rep}")
    assert max(rep["velocity_rel_l2"]) < MIXED_TOL, rep
    assert rep["mean_profile_rel_l2"] < MIXED_TOL and max(rep["wall_shear_rel"]) < MIXED_TOL, rep
    # both paths enforce the divergence constraint in FP64
    scale = np.linalg.norm(ref)
    assert max(rep["divergence_l2"]) < 1e-10 * scale, rep


@pytest.mark.parametrize("steps,ystage", [(1, "mfma"), (100, "mfma"), (1, "dct"), (100, "dct")])
def test_mixed_retau180_reference_dt(gpu, monkeypatch, steps, ystage):
    """128x129x128 (product 196x193x196), full k^-5/3 perturbation, rms 10% of the centreline,
    at the reference's DT for this grid (5e-5); y stages on the matrix cores or as a DCT-I (both
    precisions use the same y stage)."""
    if ystage == "dct":
        monkeypatch.setenv("CDNS_YSTAGE", "dct")
    else:
        monkeypatch.delenv("CDNS_YSTAGE", raising=False)
    cfg = config.retau180(dt=5e-5)
    u0 = fields.seeded_velocity(cfg, seed=180, mean="log", spectrum="k53", amplitude=0.1)
    _check(gpu, f"retau180 dt=5e-5 y {ystage}", cfg, u0, steps)


def test_mixed_retau180_dt1e3(gpu):
    """dt = 1e-3: 20 steps, before the explicit-advection growth of the highest kx modes
    (a property of the scheme, shared by the FP64 oracle) reaches the resolved field."""
    cfg = config.retau180()
    u0 = fields.seeded_velocity(cfg, seed=180, mean="log", spectrum="k53", amplitude=0.1, kcut=8)
    _check(gpu, "retau180 dt=1e-3", cfg, u0, 20)


def test_mixed_channel32_100_steps(gpu):
    cfg = config.channel_32()
    u0 = fields.seeded_velocity(cfg, seed=5, mean="log", spectrum="k53", amplitude=0.1, kcut=8)
    _check(gpu, "channel_32", cfg, u0, 100)


def test_mixed_nonlinear_term_single_precision_storage(gpu):
    """The mixed step keeps its nonlinear term and that term's history as single-precision
    planes: state written to them comes back rounded to FP32 (and widened exactly); an FP64
    context keeps them exact."""
    cfg = config.channel_32()
    rng = np.random.default_rng(11)
    for precision, exact in ((gpu.MIXED, False), (gpu.FP64, True)):
        dev = gpu.Channel(cfg, precision=precision)
        for name in ("nonlinear", "nonlinear_prev"):
            shape = dev.get_state(name).shape
            x = rng.standard_normal(shape)
            dev.set_state(name, x)
            back = dev.get_state(name)
            want = x if exact else x.astype(np.float32).astype(np.float64)
            assert np.array_equal(back, want), (precision, name)
        dev.close()


def test_mixed_single_precision_nonlinear_term_is_lossless(gpu, monkeypatch):
    """The matrix-core y forward of the mixed step forms N in FP32, so single-precision storage
    of N and its history loses nothing: 100 mixed steps with N kept in FP32 planes and with N kept
    in FP64 planes (CDNS_MIXED_N=f64) give the same velocity, bit for bit (no body force)."""
    cfg = config.retau180(dt=5e-5)
    u0 = fields.seeded_velocity(cfg, seed=180, mean="log", spectrum="k53", amplitude=0.1)
    monkeypatch.delenv("CDNS_YSTAGE", raising=False)
    monkeypatch.delenv("CDNS_MIXED_N", raising=False)
    got32 = _run(gpu, cfg, u0, 100, gpu.MIXED)
    monkeypatch.setenv("CDNS_MIXED_N", "f64")
    got64 = _run(gpu, cfg, u0, 100, gpu.MIXED)
    assert np.isfinite(got32).all()
    assert np.array_equal(got32, got64)
