"""The reference's call chain through a compiled C++ caller: channel-dns_amd/host/run_steps, the
Simulation serial loop of channel-dns_amd/host/cdns_simulation.hpp over the C ABI, run as its own
process on the GPU, against the oracle (1e-10 rel-L2 per component after 5 steps)."""
import os
import subprocess

import numpy as np
import pytest

from cdns import config, fields
from helpers import stage3_config
import pyoracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "channel-dns_amd", "host", "run_steps")


def _write_params(path, cfg, precision=0):
    keys = ["NX", "NY", "NZ", "NXS", "NYS", "NZS", "NX2", "NY2", "NZ2", "XL", "ZL", "THETA", "Re", "G_pre", "a_pre0",
            "b_pre0", "INDEX_GRAD_PRE", "DT", "V_Microblow0"]
    with open(path, "w") as f:
        for k in keys:
            f.write(f"{k}={cfg[k]!r}\n")
        for c in range(3):
            f.write(f"a_V{c}={cfg['a_V'][c]!r}\nb_V{c}={cfg['b_V'][c]!r}\n")
        f.write(f"precision={precision}\n")


@pytest.mark.parametrize("name", ["channel_32", "stage3"])
def test_cpp_simulation_loop_matches_oracle(gpu, tmp_path, name):
    assert os.path.exists(EXE), "build first (__graft_entry__.build)"
    cfg = stage3_config() if name == "stage3" else config.channel_32()
    u0 = fields.seeded_velocity(cfg, seed=12)
    _write_params(tmp_path / "params.txt", cfg)
    np.ascontiguousarray(u0).tofile(tmp_path / "u0.bin")
    print(f"run_steps {name} 5 steps", flush=True)
    r = subprocess.run([EXE, str(tmp_path / "params.txt"), str(tmp_path / "u0.bin"), "5", str(tmp_path / "u5.bin")],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(tmp_path / "u5.bin", dtype=np.complex128).reshape(u0.shape)
    ref = pyoracle.Oracle(cfg)
    ref.set_state("velocity", u0)
    ref.run(5)
    errs = fields.rel_l2_components(got, ref.get_state("velocity"))
    assert max(errs) < 1e-10, errs
