"""Benchmark of the channel-dns per-step hot path on one MI355X.

A step is one full semi-implicit time step (nonlinear term through the 3/2-dealiased
product grid, RHS, coupled tau solves, mean-mode and flow corrections) of the
Re_tau=180 channel: base grid 128x129x128, product grid 196x193x196, FP64.  The state
is resident in HBM before timing starts; steps 1-2 (solver setup) are warm-up.  The time
step is the reference's own for this grid (DT=5e-5, examples/Re_tau180/Re_tau180.toml:52):
the cost of a step does not depend on dt, and at dt=1e-3 the explicitly advected highest
kx modes grow by >10% per step, so a 100-step run of a k^-5/3 field does not stay finite.

Prints one JSON line.  `roofline` is for the kernel with the largest share of step time:
algorithmic bytes per launch (see `ALGO`) over its mean launch time from HIP events on
the backend's stream.  `cpu_baseline` times the CPU oracle (a port of the reference
algorithm, OpenMP threads) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "channel-dns_amd"))

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(cfg, esz):
    """Bytes each kernel must move once per launch (nonzero inputs + outputs, 16-byte complex
    in FP64, esz-byte complex for physical-side buffers)."""
    nkx, nzk, nj = cfg["NXS"] + 1, 2 * cfg["NZS"] + 1, cfg["NYS"] + 1
    NQ, NL = nkx * nzk, cfg["NY2"] + 1
    S = NQ * nj            # one spectral component
    Y = NL * NQ            # one y-physical field
    Z = NL * cfg["NZ2"] * nkx  # one z-physical field
    zc = 16
    return {
        "curl": 6 * S * zc,
        "y_backward": 6 * S * zc + 6 * Y * esz,
        # u, v, w in; u, v, w and the vorticity out (the curl's wall sums: 2 complex per column)
        "y_backward_curl": 3 * S * zc + 6 * Y * esz + NQ * 2 * zc,
        "z_backward": 6 * Y * esz + 6 * Z * esz,
        "x_product": 6 * Z * esz + 3 * Z * esz,
        "z_forward": 3 * Z * esz + 3 * Y * esz,
        "y_forward": 3 * Y * esz + 3 * S * zc,
        # u, u_prev, N, N_prev in; the influence/B responses as 9 real columns (P and v real,
        # u = i kx phi, w = i kz phi) and 2 real ad columns, read once for the two columns
        # kz = -m, +m of a kx that share them (nkx (NZS + 1) distinct columns); u, P out
        "solve_step": 12 * S * zc + 11 * nkx * (cfg["NZS"] + 1) * nj * 8 + 4 * S * zc,
    }


# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE on this workload
# (FP64), each scaled by what the counter reads for a 256 MiB device copy (FETCH_SIZE x2.0,
# WRITE_SIZE x1.0); profiles/step_retau180_hbm_traffic.md holds the runs.
# y stage of the step when CDNS_YSTAGE is not set: the matrix-core apply ("mfma") or the DCT-I
# through FFTs of length 2 NY2 ("dct"); the bench times the other one beside it.
DEFAULT_YSTAGE = "mfma"

# SQ_INSTS_VALU of one x-product launch on this workload (FP64; rocprofv3 --pmc) and the issue
# peak it is set against: a SIMD is 16 lanes wide, so it issues one wave64 vector instruction
# per 4 cycles; 1024 SIMDs at 2.4 GHz.
MEASURED_VALU = {"x_product": 42.6e6}
VALU_PEAK_WAVE_INSTS = 1024 * 2.4e9 / 4

MEASURED_TRAFFIC = {
    "x_product": 356.3e6,
    "solve_step": 329.2e6,
}


def adjoint_timing(steps=200, k_ckpt=20, dt=5e-4):
    """Forward and checkpointed adjoint (gradient of the terminal kinetic energy w.r.t. the
    initial velocity and a body force) on the 64x65x64 grid (product 96x97x96)."""
    import numpy as np
    import torch

    import cdns
    from cdns import config, fields

    cfg = dict(config.channel_64(), DT=dt)
    u0 = fields.seeded_velocity(cfg, seed=64, mean="log", spectrum="k53", amplitude=0.1)
    f = 0.01 * fields.seeded_velocity(cfg, seed=65, mean="laminar", kmax=4, amplitude=1.0)
    f[cfg["NZS"], :, 0, :] = 0
    dev = cdns.Channel(cfg)
    dev.set_body_force(f)
    dev.set_state("velocity", u0)
    dev.run(2, it_begin=1)
    dev.run(steps, it_begin=3)  # warm-up of the same shape
    dev.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev.run(steps, it_begin=3)
    dev.sync()
    t_fwd = time.perf_counter() - t0
    dev.adjoint_run(u0, f, steps, k_ckpt, 0)  # warm-up
    t0 = time.perf_counter()
    J, gu, gf = dev.adjoint_run(u0, f, steps, k_ckpt, 0)
    t_adj = time.perf_counter() - t0
    dev.close()
    return {
        "workload": f"64x65x64 (product 96x97x96), dt={dt}, {steps} steps, checkpoint every {k_ckpt}, "
                    "d(KE(T))/d(u0, body force)",
        "forward_ms_per_step": 1e3 * t_fwd / steps,
        "adjoint_run_ms": 1e3 * t_adj,
        "adjoint_run_over_forward": t_adj / t_fwd,
        "backward_pass_over_forward": (t_adj - t_fwd) / t_fwd,
        "finite": bool(np.isfinite(J) and np.isfinite(gu).all() and np.isfinite(gf).all()),
    }


def cpu_baseline(cfg, u0, seconds_target=15.0, threads=None):
    import numpy as np  # noqa: F401

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    n = threads or min(16, os.cpu_count() or 1)
    pyoracle.set_threads(n)
    o = pyoracle.Oracle(cfg)
    o.set_state("velocity", u0)
    o.step(1)
    o.step(2)
    t0 = time.perf_counter()
    k = 0
    while True:
        o.step(3 + k)
        k += 1
        el = time.perf_counter() - t0
        if el > seconds_target or k >= 50:
            break
    return {"value": k / el, "unit": "steps/s", "cores": n, "kind": "port",
            "sample": f"{k} steps (it=3..{2 + k}) of the same workload, oracle on {n} OpenMP threads, {el:.1f} s"}


def counter_run(cdns, cfg, u0, precision):
    """Steps 1-5 of the step in one precision form, then a plain device copy of known size
    (256 MiB each way) whose counter readings scale FETCH_SIZE and WRITE_SIZE to bytes."""
    import torch

    dev = cdns.Channel(cfg, precision=precision)
    dev.set_state("velocity", u0)
    dev.run(5, it_begin=1)
    dev.sync()
    x = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    dev.close()
    print(json.dumps({"counter_run": "done", "steps": 5, "precision": precision, "copy_bytes": x.numel() * 8}),
          flush=True)


def run_form(cdns, cfg, u0, prec, steps, warm, ystage=None):
    """One precision form of the step: a profiled pass (HIP events around every launch, which
    also brings the GPU to its clocks), then W warm-up steps and K timed steps from the same
    initial state.  Returns the timed seconds, the per-kernel (launches, ms) and the final state."""
    import numpy as np
    import torch

    old = os.environ.get("CDNS_YSTAGE")
    if ystage is not None:
        os.environ["CDNS_YSTAGE"] = ystage
    try:
        dev = cdns.Channel(cfg, precision=prec)
    finally:
        if ystage is not None:
            if old is None:
                os.environ.pop("CDNS_YSTAGE", None)
            else:
                os.environ["CDNS_YSTAGE"] = old
    dev.set_state("velocity", u0)
    dev.run(warm, it_begin=1)
    dev.profile(True)
    dev.profile_reset()
    dev.run(steps, it_begin=warm + 1)
    dev.sync()
    prof = dev.profile_read()
    dev.profile(False)
    dev.set_state("velocity", u0)
    for n in ("velocity_prev", "nonlinear_prev"):
        dev.set_state(n, dev.vector())
    dev.set_state("pressure_bc_prev", dev.boundary())
    dev.run(warm, it_begin=1)
    dev.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev.run(steps, it_begin=warm + 1)
    dev.sync()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    final = dev.get_state("velocity")
    dev.close()
    return dict(elapsed=el, prof=prof, finite=bool(np.isfinite(final).all()), final=final)


def precision_100(cdns, fields, cfg, u0):
    """Mixed against FP64 after 1 and 100 steps from the same initial state (untimed)."""
    out = {}
    states = {}
    for prec in (cdns.FP64, cdns.MIXED):
        dev = cdns.Channel(cfg, precision=prec)
        dev.set_state("velocity", u0)
        dev.run(1, it_begin=1)
        s1 = dev.get_state("velocity")
        dev.run(99, it_begin=2)
        states[prec] = (s1, dev.get_state("velocity"))
        dev.close()
    for i, n in ((0, "after_1_step"), (1, "after_100_steps")):
        out[n] = fields.precision_report(cfg, states[cdns.MIXED][i], states[cdns.FP64][i])
    return out


def c5_adjoint(cdns, config, fields):
    """The larger case (256x193x192, product 384x289x288): steps/s of the mixed and FP64 forms,
    and the FP64 gradient of the terminal mean wall shear over 500 steps with a checkpoint every
    25 (forward horizon + checkpointed backward) against the 500-step FP64 forward; device memory
    of the context and of the horizon storage."""
    import numpy as np
    import torch

    cfg = dict(config.retau395(), DT=5e-5)
    nsteps, kc = 500, 25
    u0 = fields.seeded_velocity(cfg, seed=395, mean="log", spectrum="k53", amplitude=0.1, kcut=16)
    res = {"workload": "256x193x192 (product 384x289x288), Re=395, dt=5e-5, 500 steps, checkpoint every 25, "
                       "d(mean wall shear at T)/d(u0)"}
    torch.cuda.synchronize()
    free0, total = torch.cuda.mem_get_info()
    for prec, name in ((cdns.MIXED, "mixed"), (cdns.FP64, "f64")):
        dev = cdns.Channel(cfg, precision=prec)
        dev.set_state("velocity", u0)
        dev.run(2, it_begin=1)
        dev.sync()
        t0 = time.perf_counter()
        dev.run(nsteps - 2, it_begin=3)
        dev.sync()
        res[f"{name}_steps_per_s"] = (nsteps - 2) / (time.perf_counter() - t0)
        if prec == cdns.FP64:
            free1, _ = torch.cuda.mem_get_info()
            res["context_bytes_f64"] = free0 - free1
            dev.set_state("velocity", u0)
            t0 = time.perf_counter()
            dev.run(nsteps, it_begin=1)
            dev.sync()
            t_fwd = time.perf_counter() - t0
            dev.adjoint_run(u0, None, nsteps, kc, 1)  # allocates the horizon storage once
            free2, _ = torch.cuda.mem_get_info()
            t0 = time.perf_counter()
            J, gu, _ = dev.adjoint_run(u0, None, nsteps, kc, 1)
            t_adj = time.perf_counter() - t0
            g = dev.cfg
            S = (2 * g["NZS"] + 1) * (g["NXS"] + 1) * ((g["NYS"] + 1 + 7) // 8 * 8)
            Z = (g["NY2"] + 1) * g["NZ2"] * (g["NXS"] + 1)
            nq = (2 * g["NZS"] + 1) * (g["NXS"] + 1)
            res["horizon_storage_bytes"] = (-(-nsteps // kc) * (9 * S + 2 * nq) + kc * 6 * Z) * 16
            res["adjoint_workspace_bytes"] = free1 - free2
            res["forward_s"] = t_fwd
            res["gradient_s"] = t_adj
            res["adjoint_forward_ratio"] = (t_adj - t_fwd) / t_fwd
            res["finite"] = bool(np.isfinite(J) and np.isfinite(gu).all())
        dev.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="retau180")
    ap.add_argument("--dt", type=float, default=5e-5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip the other y stage and the c5 adjoint case")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--counters", action="store_true",
                    help="no timing: run 5 FP64 steps and three 256 MiB device copies (the calibration "
                         "of FETCH_SIZE / WRITE_SIZE) for a rocprofv3 --pmc pass")
    ap.add_argument("--counters-mixed", action="store_true", help="the counter run in the mixed form")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this backend runs one GPU per process; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    import cdns
    from cdns import config, fields

    torch.cuda.set_device(0)
    cfg = dict(config.WORKLOADS[args.workload](), DT=args.dt)
    u0 = fields.seeded_velocity(cfg, seed=180, mean="log", spectrum="k53", amplitude=0.1)
    warm = max(args.warmup, 2)
    if args.counters or args.counters_mixed:
        counter_run(cdns, cfg, u0, cdns.MIXED if args.counters_mixed else cdns.FP64)
        return

    ystage = os.environ.get("CDNS_YSTAGE") or DEFAULT_YSTAGE
    res = {}
    for prec, name in ((cdns.FP64, "f64"), (cdns.MIXED, "mixed")):
        res[name] = run_form(cdns, cfg, u0, prec, args.steps, warm, ystage)
    other = {}
    if not args.no_compare:
        alt = "mfma" if ystage == "dct" else "dct"
        for prec, name in ((cdns.FP64, "f64"), (cdns.MIXED, "mixed")):
            r = run_form(cdns, cfg, u0, prec, args.steps, warm, alt)
            other[name] = {"ms_per_step": 1e3 * r["elapsed"] / args.steps,
                           "kernels_ms_per_launch": {k: v[1] / max(v[0], 1) for k, v in r["prof"].items()}}
        other["y_stage"] = alt

    r = res["f64"]
    steps_per_s = args.steps / r["elapsed"]
    algo = algorithmic_bytes(cfg, 16)
    kern = {}
    for k, (cnt, tot) in r["prof"].items():
        kern[k] = {"launches": cnt, "ms_per_launch": tot / max(cnt, 1), "ms_per_step": tot / args.steps}
        if k in algo and cnt:
            gbs = algo[k] / (tot / cnt * 1e-3) / 1e9
            kern[k]["achieved_GBs"] = gbs
            kern[k]["frac_hbm"] = gbs / HBM_PEAK_GBS
    dom = max((k for k in kern if k in algo), key=lambda k: kern[k]["ms_per_step"])
    dk = kern[dom]
    roof = {"kernel": dom, "bound": "hbm", "achieved": dk["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": dk["achieved_GBs"] / HBM_PEAK_GBS, "traffic": MEASURED_TRAFFIC.get(dom),
            "algorithmic_bytes_per_launch": algo[dom], "ms_per_launch": dk["ms_per_launch"]}
    if "x_product" in kern and "x_product" in MEASURED_VALU:
        # the resource that binds the x product: vector-ALU issue (wave-instructions per launch from
        # SQ_INSTS_VALU against the issue peak)
        xi = MEASURED_VALU["x_product"]
        rate = xi / (kern["x_product"]["ms_per_launch"] * 1e-3)
        kern["x_product"]["valu"] = {"wave_insts_per_launch": xi, "achieved_G_per_s": rate / 1e9,
                                     "peak_G_per_s": VALU_PEAK_WAVE_INSTS / 1e9, "frac": rate / VALU_PEAK_WAVE_INSTS}
    adj = None if args.no_adjoint else adjoint_timing()
    c5 = None if (args.no_adjoint or args.no_compare) else c5_adjoint(cdns, config, fields)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(cfg, u0, args.cpu_seconds)
    mixed_sps = args.steps / res["mixed"]["elapsed"]
    # the mixed run against the FP64 run from the same initial state after warm + steps steps
    mixed_err = fields.precision_report(cfg, res["mixed"]["final"], r["final"])
    prec100 = precision_100(cdns, fields, cfg, u0)
    line = {
        "metric": "steps/s at 128x129x128 (product grid 196x193x196), FP64 step on one MI355X",
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warm,
        "ms_per_step": 1e3 * r["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded log-law mean + divergence-free k^-5/3 perturbation (rms 10% of centreline)",
        "config": {"workload": f"{args.workload}: base {cfg['NX']}x{cfg['NY'] + 1}x{cfg['NZ']}, product "
                               f"{cfg['NX2']}x{cfg['NY2'] + 1}x{cfg['NZ2']}, Re={cfg['Re']}, dt={cfg['DT']}",
                   "steps_timed_from_iteration": warm + 1, "y_stage": ystage},
        "roofline": roof,
        "cpu_baseline": cpu,
        "mixed": {"steps_per_s": mixed_sps, "ms_per_step": 1e3 * res["mixed"]["elapsed"] / args.steps,
                  "speedup_vs_f64": mixed_sps / steps_per_s, "finite": res["mixed"]["finite"],
                  "error_vs_f64_after_steps": warm + args.steps, "error_vs_f64": mixed_err,
                  "error_vs_f64_1_and_100_steps": prec100,
                  "kernels_ms_per_launch": {k: v[1] / max(v[0], 1) for k, v in res["mixed"]["prof"].items()}},
        "y_stage_alternative": other or None,
        "adjoint_forward_ratio": None if adj is None else adj["backward_pass_over_forward"],
        "adjoint": adj,
        "adjoint_c5": c5,
        "kernels": kern,
        "finite": r["finite"],
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
