"""cdns_adjoint_run_noalloc (the adjoint horizon in the library's fixed device scratch array, no
allocation) against cdns_adjoint_run on the same device inputs, through the C ABI: two calls in
a row, the second with a shorter horizon, and a horizon beyond the array refused."""
import math

import numpy as np
import pytest

from cdns import config, fields

pytestmark = pytest.mark.gpu


def _dev(torch, a):
    return torch.as_tensor(np.ascontiguousarray(a).view(np.float64), device="cuda")


def _host(t):
    return t.cpu().numpy().view(np.complex128)


def test_noalloc_adjoint_run_matches_adjoint_run(gpu):
    import torch

    cfg = config.channel_32()
    u0 = fields.seeded_velocity(cfg, seed=5, mean="log", spectrum="k53", amplitude=0.1, kcut=8)
    f = 0.01 * fields.seeded_velocity(cfg, seed=7, mean="laminar", kmax=4, amplitude=1.0)
    f[cfg["NZS"], :, 0, :] = 0
    ch = gpu.Channel(cfg)
    u0_d, f_d = _dev(torch, u0), _dev(torch, f)
    print("noalloc adjoint: channel_32, horizons (10, 4) then (6, 3)", flush=True)
    for nsteps, kc, objective in ((10, 4, 0), (6, 3, 1)):
        gu_a, gf_a = torch.zeros_like(u0_d), torch.zeros_like(u0_d)
        gu_b, gf_b = torch.full_like(u0_d, np.nan), torch.full_like(u0_d, np.nan)
        torch.cuda.synchronize()  # torch's fills are done before the library's stream writes
        Ja = ch.adjoint_run_device(u0_d, f_d, nsteps, kc, objective, gu_a, gf_a)
        Jb = ch.adjoint_run_device(u0_d, f_d, nsteps, kc, objective, gu_b, gf_b, noalloc=True)
        torch.cuda.synchronize()
        a_u, a_f, b_u, b_f = _host(gu_a), _host(gf_a), _host(gu_b), _host(gf_b)
        eu = np.abs(b_u - a_u).max() / np.abs(a_u).max()
        ef = np.abs(b_f - a_f).max() / np.abs(a_f).max()
        print(f"nsteps {nsteps} k {kc} objective {objective}: J {Jb!r} vs {Ja!r}; grad rel diff {eu:.1e} {ef:.1e}")
        assert np.isfinite(b_u).all() and np.isfinite(b_f).all()
        assert math.isclose(Jb, Ja, rel_tol=1e-13)
        assert eu <= 1e-13 and ef <= 1e-13
    # a horizon whose storage exceeds the 64 MiB array is refused before any work
    with pytest.raises(RuntimeError, match="more than"):
        ch.adjoint_run_device(u0_d, f_d, 400, 400, 0, gu_b, gf_b, noalloc=True)
