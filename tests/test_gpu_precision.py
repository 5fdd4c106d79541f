"""Mixed precision (FP32 transforms and nonlinear term, FP64 state, RHS and solves) against
the FP64 path after 100 steps: relative L2 per velocity component within 1e-5."""
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


@pytest.mark.parametrize("name", ["channel_32", "retau180"])
def test_mixed_vs_fp64_100_steps(gpu, name):
    cfg = config.WORKLOADS[name]()
    # log-law mean with a band-limited k^-5/3 perturbation: a flow that the scheme carries
    # through 100 steps at dt = 1e-3 (a laminar-mean field at Re 180 outgrows it by step 70)
    u0 = fields.seeded_velocity(cfg, seed=180 if name == "retau180" else 5, mean="log", spectrum="k53",
                                amplitude=0.1, kcut=8)
    ref = _run(gpu, cfg, u0, 100, gpu.FP64)
    got = _run(gpu, cfg, u0, 100, gpu.MIXED)
    assert np.isfinite(ref).all() and np.isfinite(got).all()
    errs = fields.rel_l2_components(got, ref)
    print(f"{name} mixed vs fp64 after 100 steps, rel-L2 per component: {errs}")
    assert max(errs) < MIXED_TOL, errs
