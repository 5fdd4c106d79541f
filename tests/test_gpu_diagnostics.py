"""Coarse-grid transform, to_physical, plane moments and diagnostics of the HIP backend against the
oracle (whose versions tests/test_oracle_diagnostics.py and the coarse-transform tests pin), and
the reference's analytic diagnostics known answer (tests/test_diagnostics.cpp:11-55) computed
end to end on the GPU."""
import math

import numpy as np
import pytest

from cdns import config, fields
from helpers import band_spectrum, maxrel, ref_transform_config
import pyoracle
from test_oracle_diagnostics import analytic_fields

pytestmark = pytest.mark.gpu

CONFIGS = {
    "ref32_full": ref_transform_config(16, 30, 16),   # NXS = NX/2, NZS = NZ/2: both Nyquist rules
    "ref32_trunc": ref_transform_config(10, 20, 10),
    "channel_64": config.channel_64(),
    "retau180": config.retau180(),
}


def coarse_spectrum(cfg, seed):
    return band_spectrum(dict(cfg, NX2=cfg["NX"], NZ2=cfg["NZ"]), seed)


@pytest.mark.parametrize("name", list(CONFIGS))
def test_coarse_transform_backward_forward(gpu, name):
    """Transforms::channel_xyzfft both ways (transforms.cpp:452-582) at 1e-13, and the round trip."""
    cfg = CONFIGS[name]
    dev, ref = gpu.Channel(cfg), pyoracle.Oracle(cfg)
    spec = coarse_spectrum(cfg, 0xC0A)
    p_dev, p_ref = dev.coarse_phys(), ref.coarse_phys()
    dev.xyzfft("B", p_dev, spec.copy())
    ref.xyzfft("B", p_ref, spec.copy())
    print(f"{name}: backward max rel {maxrel(p_dev, p_ref):.2e}")
    assert maxrel(p_dev, p_ref) < 1e-13
    s_dev, s_ref = dev.scalar(), ref.scalar()
    dev.xyzfft("F", p_ref.copy(), s_dev)
    ref.xyzfft("F", p_ref.copy(), s_ref)
    print(f"{name}: forward max rel {maxrel(s_dev, s_ref):.2e}, round trip {maxrel(s_dev, spec):.2e}")
    assert maxrel(s_dev, s_ref) < 1e-13
    assert maxrel(s_dev, spec) < 1e-12


@pytest.mark.parametrize("name", ["channel_64", "retau180"])
def test_to_physical_moments_diagnostics_match_oracle(gpu, name):
    """A stepped state: to_physical, the 20 plane sums and EKT / TKE / divergence L2 against the
    oracle's on the same spectral state."""
    cfg = dict(CONFIGS[name], DT=5e-4 if name == "channel_64" else 5e-5)
    u0 = fields.seeded_velocity(cfg, seed=11, mean="log", spectrum="k53", amplitude=0.1, kcut=8)
    dev = gpu.Channel(cfg)
    dev.set_state("velocity", u0)
    dev.run(3, it_begin=1)
    vel_s, pr_s = dev.get_state("velocity"), dev.get_state("pressure")
    ref = pyoracle.Oracle(cfg)
    vel_r, pr_r, fin_r = ref.to_physical(vel_s, pr_s)
    vel_d, pr_d, fin_d = dev.to_physical()
    print(f"{name}: to_physical max rel u {maxrel(vel_d, vel_r):.2e} p {maxrel(pr_d, pr_r):.2e}")
    assert fin_d and fin_r
    assert maxrel(vel_d, vel_r) < 1e-12 and maxrel(pr_d, pr_r) < 1e-12
    sums_d, pf_d = dev.plane_moments(higher=True)
    sums_r, pf_r = ref.plane_moments(vel_r[0], vel_r[1], vel_r[2], pr_r, higher=True)
    # per sum, relative to the plane size times the field scale to the moment's degree (sums that
    # vanish, like the plane sum of v, carry round-off only)
    npl = cfg["NX"] * cfg["NZ"]
    amp = max(np.abs(vel_r).max(), np.abs(pr_r).max())
    degree = np.array([1] * 4 + [2] * 10 + [3] * 3 + [4] * 3)
    per_sum = np.max(np.abs(sums_d - sums_r), axis=0) / (npl * amp ** degree)
    err = float(per_sum.max())
    print(f"{name}: plane sums max scaled error {err:.2e}; per sum {np.array2string(per_sum, precision=1)}")
    assert pf_d.all() and pf_r.all() and err < 1e-13
    dg = dev.diagnostics()
    dr = ref.diagnostics(*ref.plane_moments(vel_r[0], vel_r[1], vel_r[2], pr_r, higher=False), vel_s, fin_r)
    print(f"{name}: diagnostics gpu {dg} oracle {dr}")
    assert dg["finite"] and dr["finite"]
    for k in ("ekt", "tke", "divergence_l2"):
        assert math.isclose(dg[k], dr[k], rel_tol=1e-10, abs_tol=1e-13), k


def test_diagnostics_analytic_known_answer_on_gpu(gpu):
    """tests/test_diagnostics.cpp:11-55 through the state: spectral velocity and pressure of the
    analytic fields -> GPU to_physical, plane sums, divergence -> the closed forms."""
    cfg = config.reference_example()
    ref = pyoracle.Oracle(cfg)
    (u, v, w, p), kx, kz = analytic_fields(cfg)
    vel = ref.vector()
    for c, f in enumerate((u, v, w)):
        ref.xyzfft("F", f, vel, stride=3, offset=c)
    pr = ref.scalar()
    ref.xyzfft("F", p, pr)
    dev = gpu.Channel(cfg)
    dev.set_state("velocity", vel)
    dev.set_state("pressure", pr)
    d = dev.diagnostics()
    n = u.size
    print(f"analytic: {d}, want ekt {0.5 * n} tke {4 / 3} div {math.sqrt(kx * kx / 6 + 2 * kz * kz)}")
    assert d["finite"]
    assert abs(d["ekt"] - 0.5 * n) < 1e-8 * n
    assert abs(d["tke"] - 4.0 / 3.0) < 1e-12
    assert abs(d["divergence_l2"] - math.sqrt(kx * kx / 6 + 2 * kz * kz)) < 1e-11
    sums, _ = dev.plane_moments(higher=False)
    assert np.all(np.abs(sums[:, 3] / (cfg["NX"] * cfg["NZ"]) - 5) < 1e-12)


def test_to_physical_flags_non_finite_imaginary_part(gpu):
    """tests/test_diagnostics.cpp:58-70: an infinite imaginary part that the real transform would
    discard is still reported."""
    cfg = config.reference_example()
    dev = gpu.Channel(cfg)
    pr = np.zeros((2 * cfg["NZS"] + 1, cfg["NYS"] + 1, cfg["NXS"] + 1), np.complex128)
    pr[cfg["NZS"], 0, 0] = complex(0.0, np.inf)
    dev.set_state("velocity", np.zeros(pr.shape + (3,), np.complex128))
    dev.set_state("pressure", pr)
    _, _, finite = dev.to_physical()
    assert not finite
    assert not dev.diagnostics()["finite"]
