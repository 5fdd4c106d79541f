"""Restart files, physical snapshots and seeded initial fields with the GPU transforms.

The host-side compositions (cdns.field_io, cdns.initial_bc) run on the coarse-grid transforms of
a Channel and are compared with the same compositions on the oracle's transforms (1e-13); the
restart test follows tests/test_simulation.cpp:169-231 (restarted against continuous run,
relative L2 below 1e-7, the restart repeating the first-order start).
"""
import math

import numpy as np
import pytest

from cdns import config, field_io, initial_bc
from helpers import maxrel
import pyoracle

pytestmark = pytest.mark.gpu


def _cfg():
    # the reference's examples/config.toml (32^3) with the restart test's disturbance and time step
    return dict(config.reference_example(), DT=1e-4, EPS_init=0.01, INDEX_disturb_Z=3)


def test_to_physical_on_gpu_matches_oracle(gpu, tmp_path):
    cfg = config.reference_example()
    dev, ref = gpu.Channel(cfg), pyoracle.Oracle(cfg)
    rng = np.random.default_rng(31)
    io_d, io_r = field_io.FieldIO(cfg, tmp_path, dev), field_io.FieldIO(cfg, tmp_path, ref)
    shape = io_d.velocity_shape()
    vel = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    pr = rng.standard_normal(shape[:3]) + 1j * rng.standard_normal(shape[:3])
    a, b = io_d.to_physical(vel, pr, finite=True), io_r.to_physical(vel, pr, finite=True)
    errs = [maxrel(a["velocity"][c], b["velocity"][c]) for c in range(3)] + [maxrel(a["pressure"], b["pressure"])]
    print(f"to_physical max rel: {errs}", flush=True)
    assert max(errs) < 1e-13 and a["finite"] and b["finite"]


def test_initialize_velocity_on_gpu_matches_oracle(gpu):
    cfg = _cfg()
    dev, ref = gpu.Channel(cfg), pyoracle.Oracle(cfg)
    vd = initial_bc.initialize_velocity(dev, initial_bc.DisturbanceEngine(12345))
    vr = initial_bc.initialize_velocity(ref, initial_bc.DisturbanceEngine(12345))
    err = maxrel(vd, vr)
    print(f"initial field max rel: {err:.2e}", flush=True)
    assert err < 1e-13


@pytest.mark.parametrize("fmt", ["ascii", "binary"])
def test_saved_restart_matches_continuous_run(gpu, tmp_path, fmt):
    cfg = _cfg()
    dev = gpu.Channel(cfg)
    initial = initial_bc.initialize_velocity(dev, initial_bc.DisturbanceEngine(12345))
    # continuous: six steps from the initial field
    dev.set_state("velocity", initial)
    dev.run(6, it_begin=1)
    continuous = dev.get_state("velocity")
    dev.close()
    # three steps, saved, read back into a new context, three more (a restart starts first order)
    first = gpu.Channel(cfg)
    first.set_state("velocity", initial)
    first.run(3, it_begin=1)
    checkpoint = first.get_state("velocity")
    io = field_io.FieldIO(cfg, tmp_path, output_format=fmt)
    io.write_restart(3, 3 * cfg["DT"], checkpoint)
    first.close()
    nt, t, restored = io.read_restart()
    assert nt == 3 and t == 3 * cfg["DT"]
    assert np.array_equal(restored.view(np.float64), checkpoint.view(np.float64))
    resumed = gpu.Channel(cfg)
    resumed.set_state("velocity", restored)
    resumed.run(3, it_begin=1)
    final = resumed.get_state("velocity")
    assert np.isfinite(final.view(np.float64)).all()
    err = math.sqrt(np.sum(np.abs(final - continuous) ** 2) / np.sum(np.abs(continuous) ** 2))
    evolution = np.sum(np.abs(continuous - initial) ** 2)
    print(f"{fmt}: restarted vs continuous rel-L2 {err:.2e}, evolution {evolution:.2e}", flush=True)
    assert evolution > 1e-16
    assert err < 1e-7
