"""CPU-side checks: the C ABI library loads and exports every declared symbol; the host
binding, configuration loader and seeded-input generator behave as documented."""
import os
import re
import subprocess

import numpy as np
import pytest

import cdns
from cdns import config, fields
import pyoracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cdns_[a-z0-9_]+)\s*\(", text)))


def _exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


def test_library_exports_every_declared_entry_point():
    assert os.path.exists(cdns.LIB_PATH), "build the HIP library first"
    decl = _declared(os.path.join(ROOT, "include", "cdns.h"))
    exp = _exported(cdns.LIB_PATH)
    missing = [d for d in decl if d not in exp]
    assert not missing, missing
    assert sorted(cdns.ABI_FUNCTIONS) == decl


def test_library_loads_without_a_gpu():
    L = cdns.lib()
    for name in cdns.ABI_FUNCTIONS:
        assert hasattr(L, name)


def test_oracle_library_exports_its_header():
    decl = sorted(set(re.findall(r"\b(orc_[a-z0-9_]+)\s*\(", open(os.path.join(ROOT, "oracle", "cdns_oracle.h")).read())))
    exp = _exported(os.path.join(ROOT, "oracle", "libcdns_oracle.so"))
    assert [d for d in decl if d not in exp] == []


def test_load_reference_config_toml(tmp_path):
    """The reference's config.toml format (examples/config.toml) loads unchanged."""
    text = """[mesh]
NX = 128
NY = 128
NZ = 128
INDEX_disturb_Z = 3
NXS = 64
NYS = 126
NZS = 64
NX2 = 196
NY2 = 192
NZ2 = 196
INDEX_Serial = 1

[parm]
XL = 12.56637061435917
ZL = 6.283185307179586
THETA = 0.0e0
INDEX_GRAD_PRE = 1
Re = 180.0
INDEX_UNSTEADY_BC = 0
G_pre = 0.0e0
a_pre0 = 0.0e0
b_pre0 = 1.0e0
a_V = [1.0e0, 1.0e0, 1.0e0]
b_V = [0.0e0, 0.0e0, 0.0e0]

[control]
INDEX_CONT = 0
EPS_init = 0.0
NSTEPS = 250000
NDUMFRQ = 10000
NPRNFRQ = 1000
DT = 1.0e-3
NCOUNT_disturb = 3
NDISTURBFRQ = 10000
EPS_disturb = 0.01
V_Microblow0 = 0.0
Nstat = 200000
"""
    p = tmp_path / "config.toml"
    p.write_text(text)
    cfg = config.load_toml(str(p))
    ref = config.retau180()
    for k in ("NX", "NY", "NZ", "NXS", "NYS", "NZS", "NX2", "NY2", "NZ2", "Re", "DT", "INDEX_GRAD_PRE"):
        assert cfg[k] == ref[k], k
    assert abs(cfg["XL"] - ref["XL"]) < 1e-14
    assert config.derived(cfg)["GRAD_PRE_X"] == -1.0


def test_seeded_velocity_is_divergence_free_no_slip_and_deterministic():
    cfg = config.channel_32()
    u = fields.seeded_velocity(cfg, seed=5)
    assert np.array_equal(u, fields.seeded_velocity(cfg, seed=5))
    assert not np.array_equal(u, fields.seeded_velocity(cfg, seed=6))
    o = pyoracle.Oracle(cfg)
    assert np.abs(o.divergence(u)).max() < 1e-12 * np.abs(u).max()
    assert np.abs(u.sum(axis=1)).max() < 1e-12 * np.abs(u).max()
    # kx = 0 plane Hermitian in kz (real field)
    NZS = cfg["NZS"]
    assert np.abs(u[:, :, 0, :] - np.conj(u[::-1, :, 0, :])).max() < 1e-14 * np.abs(u).max()
    # amplitude: max |u'| on the coarse grid = 0.1 of the centreline
    pert = u.copy()
    pert[NZS, :, 0, 0] -= fields.mean_profile(cfg, "laminar")[0]
    peak = max(np.abs(fields.physical_values(cfg, pert[..., c])).max() for c in range(3))
    assert abs(peak - 0.1 * cfg["Re"] / 2) < 1e-9 * peak


def test_log_mean_profile_vanishes_at_walls():
    cfg = config.retau180()
    c, uc = fields.mean_profile(cfg, "log")
    assert abs(c.sum()) < 1e-12 * uc and abs((c * (-1.0) ** np.arange(c.size)).sum()) < 1e-12 * uc
    assert 15 < uc < 25


def test_flow_diagnostics_closed_forms():
    """Laminar u = (Re/2)(1 - y^2): du/dy = -Re at y = +1 and +Re at y = -1; a curl field is
    divergence-free; the precision report of a state against itself is zero."""
    from cdns import config, fields

    cfg = config.channel_32()
    u = fields.seeded_velocity(cfg, seed=3, mean="laminar", kmax=4)
    top, bottom = fields.wall_shear(cfg, u)
    assert abs(top + cfg["Re"]) < 1e-12 * cfg["Re"] and abs(bottom - cfg["Re"]) < 1e-12 * cfg["Re"]
    assert fields.divergence_l2(cfg, u) < 1e-12 * np.linalg.norm(u)
    rep = fields.precision_report(cfg, u, u)
    assert max(rep["velocity_rel_l2"]) == 0 and rep["mean_profile_rel_l2"] == 0
