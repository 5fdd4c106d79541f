"""Run configurations for the channel-flow step.

A configuration mirrors the reference's `config.toml` keys ([mesh] [parm]
[control]; src/core/params.cpp:130-228) plus the derived flow-control values
(params.cpp:296-307).  `load_toml` reads a reference run directory's
config.toml unchanged.
"""
import math

try:
    import tomli as _toml
except ImportError:  # pragma: no cover
    _toml = None

PI = math.pi


def make(NX, NY, NZ, NXS, NYS, NZS, NX2, NY2, NZ2, XL, ZL, Re, DT, THETA=0.0, INDEX_GRAD_PRE=1,
         G_pre=0.0, a_pre0=0.0, b_pre0=1.0, a_V=(1.0, 1.0, 1.0), b_V=(0.0, 0.0, 0.0), V_Microblow0=0.0, **extra):
    cfg = dict(NX=NX, NY=NY, NZ=NZ, NXS=NXS, NYS=NYS, NZS=NZS, NX2=NX2, NY2=NY2, NZ2=NZ2, XL=float(XL),
               ZL=float(ZL), Re=float(Re), DT=float(DT), THETA=float(THETA), INDEX_GRAD_PRE=int(INDEX_GRAD_PRE),
               G_pre=float(G_pre), a_pre0=float(a_pre0), b_pre0=float(b_pre0), a_V=[float(v) for v in a_V],
               b_V=[float(v) for v in b_V], V_Microblow0=float(V_Microblow0))
    cfg.update(extra)
    return cfg


# Workloads: (name -> config).  Sizes follow the measurement plan of the step.
def channel_32():
    """Plain FP64, 32x33x32 base grid, 48x49x48 product grid, 4pi x 2 x 2pi, Re=180."""
    return make(32, 32, 32, 16, 30, 16, 48, 48, 48, 4 * PI, 2 * PI, 180.0, 1e-3)


def retau180(dt=1e-3):
    """Re_tau=180 channel: 128x129x128 base grid, 196x193x196 product grid (examples/Re_tau180).

    The measurement plan quotes dt = 1e-3; the reference's own run of this grid uses
    DT = 5e-5 (examples/Re_tau180/Re_tau180.toml:52).  With explicit advection (SBDF2) the
    highest retained kx modes at the centreline (k = 32, U ~ 18-23) see omega*dt ~ 0.6-0.75 at
    dt = 1e-3, an amplification of 1.13-1.26 per step, so a k^-5/3 field stays bounded only for
    a few tens of steps there; long runs use the reference's DT."""
    return make(128, 128, 128, 64, 126, 64, 196, 192, 196, 4 * PI, 2 * PI, 180.0, dt)


def channel_64():
    """64x65x64 base, 96x97x96 product grid, 2pi x 2 x pi, Re=180 (adjoint workload)."""
    return make(64, 64, 64, 32, 62, 32, 96, 96, 96, 2 * PI, PI, 180.0, 1e-3)


def retau395():
    """Re_tau~395: 256x193x192 base, 384x289x288 product grid, 2pi x 2 x pi, dt=5e-4."""
    return make(256, 192, 192, 128, 190, 96, 384, 288, 288, 2 * PI, PI, 395.0, 5e-4)


def reference_example():
    """examples/config.toml of the reference: 32^3, product grid equal to the base grid."""
    return make(32, 32, 32, 16, 30, 16, 32, 32, 32, 10.0, 10.0, 2800.0, 3e-3, INDEX_GRAD_PRE=0,
                V_Microblow0=1e-3)


WORKLOADS = {"channel_32": channel_32, "retau180": retau180, "channel_64": channel_64, "retau395": retau395,
             "reference_example": reference_example}


def derived(cfg):
    """GRAD_PRE_X/Z or QXmean/QZmean as params.cpp:296-307 computes them."""
    th = cfg["THETA"] * math.acos(-1.0) / 180.0
    if cfg["INDEX_GRAD_PRE"] == 1:
        return dict(GRAD_PRE_X=-math.cos(th), GRAD_PRE_Z=-math.sin(th), QXmean=0.0, QZmean=0.0)
    return dict(GRAD_PRE_X=0.0, GRAD_PRE_Z=0.0, QXmean=math.cos(th), QZmean=math.sin(th))


def load_toml(path):
    """Read a reference config.toml (case-insensitive keys, params.cpp:115-228)."""
    if _toml is None:
        raise RuntimeError("tomli is not available")
    with open(path, "rb") as f:
        raw = _toml.load(f)
    t = {k.lower(): ({kk.lower(): vv for kk, vv in v.items()} if isinstance(v, dict) else v) for k, v in raw.items()}
    mesh, parm, ctrl = t["mesh"], t["parm"], t["control"]
    cfg = make(*(int(mesh[k.lower()]) for k in ("NX", "NY", "NZ", "NXS", "NYS", "NZS", "NX2", "NY2", "NZ2")),
               XL=parm["xl"], ZL=parm["zl"], Re=parm["re"], DT=ctrl["dt"], THETA=parm["theta"],
               INDEX_GRAD_PRE=parm["index_grad_pre"], G_pre=parm["g_pre"], a_pre0=parm["a_pre0"],
               b_pre0=parm["b_pre0"], a_V=parm["a_v"], b_V=parm["b_v"], V_Microblow0=ctrl["v_microblow0"])
    cfg["NSTEPS"] = int(ctrl["nsteps"])
    cfg["INDEX_Serial"] = int(mesh["index_serial"])
    # initial field (initial_bc.cpp:66-95): fresh start or restart, and the disturbance
    cfg["INDEX_CONT"] = int(ctrl["index_cont"])
    cfg["EPS_init"] = float(ctrl["eps_init"])
    cfg["INDEX_disturb_Z"] = int(mesh["index_disturb_z"])
    return cfg
