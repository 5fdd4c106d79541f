"""Host-side binding of the MI355X channel-dns step backend (libcdns_hip.so, include/cdns.h).

The classes mirror the reference's interfaces for this path:
  Transforms.channel_xyzfft2      src/fft/transforms.hpp:40-43
  Transforms.channel_xyzfft       src/fft/transforms.hpp:39 (coarse grid)
  FieldIO.to_physical             src/io/field_io.cpp:199-220
  Statistics.reduce               src/sim/statistics.cpp:54-109
  Diagnostics.evaluate            src/sim/diagnostics.cpp:39-75
  vector_ops.curl / cross_product src/solver/vector_ops.hpp:22,33
  TimeStepper.nonlinear_term      src/solver/timestepper.hpp:10-11
  EqSolvePV.setup / solve_modes   src/solver/eqsolve_pv.hpp:17-51
  Simulation state, step, run     src/sim/simulation.hpp:21-62
  initial_bc (seeded fields)      src/sim/initial_bc.cpp:15-99        (module initial_bc)
  FieldIO restart / physical / EKT src/io/field_io.cpp:121-156,222-427 (module field_io)
Arrays are numpy (host) or torch CUDA tensors (device) in the reference layouts.
The library is required: nothing here computes the step on the CPU.
"""
import ctypes
import os

import numpy as np

from . import config, fields  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libcdns_hip.so")
_LIB = None

HOST, DEVICE = 0, 1
FP64, MIXED = 0, 1

ABI_FUNCTIONS = [
    "cdns_create", "cdns_destroy", "cdns_last_error", "cdns_sync", "cdns_stream", "cdns_xyzfft2", "cdns_curl",
    "cdns_cross_product", "cdns_nonlinear_term", "cdns_setup", "cdns_solve_modes", "cdns_get_table",
    "cdns_set_state", "cdns_get_state", "cdns_step", "cdns_run", "cdns_flow_gradients", "cdns_time",
    "cdns_profile", "cdns_profile_read", "cdns_profile_reset", "cdns_set_body_force", "cdns_objective",
    "cdns_adjoint_step", "cdns_adjoint_run", "cdns_xyzfft", "cdns_to_physical", "cdns_plane_moments",
    "cdns_diagnostics", "cdns_engine_create", "cdns_engine_draw", "cdns_engine_destroy", "cdns_adjoint_run_noalloc",
]


class CdnsConfig(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("NX", "NY", "NZ", "NXS", "NYS", "NZS", "NX2", "NY2", "NZ2")] + [
        (n, ctypes.c_double) for n in ("XL", "ZL", "THETA", "Re", "G_pre", "a_pre0", "b_pre0")
    ] + [("a_V", ctypes.c_double * 3), ("b_V", ctypes.c_double * 3), ("INDEX_GRAD_PRE", ctypes.c_int),
         ("DT", ctypes.c_double), ("V_Microblow0", ctypes.c_double), ("precision", ctypes.c_int)]


def lib(path=None):
    """Load the backend library (fails loudly when it has not been built)."""
    global _LIB
    if _LIB is None:
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise RuntimeError(f"HIP backend not built: {p} is missing (run __graft_entry__.build())")
        L = ctypes.CDLL(p)
        vp, d, i, cp = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_char_p
        L.cdns_create.argtypes = [ctypes.POINTER(CdnsConfig), ctypes.POINTER(vp)]
        L.cdns_create.restype = i
        L.cdns_destroy.argtypes = [vp]
        L.cdns_last_error.argtypes = [vp]
        L.cdns_last_error.restype = cp
        L.cdns_sync.argtypes = [vp]
        L.cdns_sync.restype = i
        L.cdns_stream.argtypes = [vp]
        L.cdns_stream.restype = vp
        L.cdns_xyzfft2.argtypes = [vp, ctypes.c_char, vp, vp, i, i]
        L.cdns_curl.argtypes = [vp, vp, vp, i]
        L.cdns_cross_product.argtypes = [vp, vp, vp, vp, i]
        L.cdns_nonlinear_term.argtypes = [vp, vp, vp, vp, i]
        L.cdns_setup.argtypes = [vp, d]
        L.cdns_solve_modes.argtypes = [vp, vp, vp, vp, vp, vp, vp, i, i]
        L.cdns_get_table.argtypes = [vp, i, vp]
        L.cdns_set_state.argtypes = [vp, i, vp, i]
        L.cdns_get_state.argtypes = [vp, i, vp, i]
        L.cdns_step.argtypes = [vp, i]
        L.cdns_run.argtypes = [vp, i, i]
        L.cdns_flow_gradients.argtypes = [vp, vp]
        L.cdns_time.argtypes = [vp]
        L.cdns_time.restype = d
        L.cdns_profile.argtypes = [vp, i]
        L.cdns_profile_read.argtypes = [vp, cp, vp, vp, i]
        L.cdns_profile_reset.argtypes = [vp]
        L.cdns_set_body_force.argtypes = [vp, vp, i]
        L.cdns_objective.argtypes = [vp, i, vp, vp, vp, i]
        L.cdns_adjoint_step.argtypes = [vp, i, vp, vp, vp, vp, vp, vp, i]
        L.cdns_adjoint_run.argtypes = [vp, vp, vp, i, i, i, vp, vp, vp, i]
        L.cdns_adjoint_run_noalloc.argtypes = [vp, vp, vp, i, i, i, vp, vp, vp, i]
        L.cdns_xyzfft.argtypes = [vp, ctypes.c_char, vp, vp, i, i]
        L.cdns_to_physical.argtypes = [vp, vp, vp, vp, i]
        L.cdns_plane_moments.argtypes = [vp, i, vp, vp]
        L.cdns_diagnostics.argtypes = [vp, vp, vp]
        L.cdns_engine_create.argtypes = [ctypes.c_ulonglong, ctypes.POINTER(vp)]
        L.cdns_engine_create.restype = i
        L.cdns_engine_draw.argtypes = [vp, i, vp, ctypes.c_ulonglong]
        L.cdns_engine_draw.restype = i
        L.cdns_engine_destroy.argtypes = [vp]
        for n in ("cdns_xyzfft2", "cdns_curl", "cdns_cross_product", "cdns_nonlinear_term", "cdns_setup",
                  "cdns_solve_modes", "cdns_get_table", "cdns_set_state", "cdns_get_state", "cdns_step",
                  "cdns_run", "cdns_flow_gradients", "cdns_profile", "cdns_profile_read", "cdns_profile_reset",
                  "cdns_set_body_force", "cdns_objective", "cdns_adjoint_step", "cdns_adjoint_run",
                  "cdns_adjoint_run_noalloc", "cdns_xyzfft",
                  "cdns_to_physical", "cdns_plane_moments", "cdns_diagnostics"):
            getattr(L, n).restype = i
        _LIB = L
    return _LIB


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags.c_contiguous
        return ctypes.c_void_p(a.ctypes.data)
    return ctypes.c_void_p(a.data_ptr())  # torch tensor


def _where(a):
    return HOST if (a is None or isinstance(a, np.ndarray)) else DEVICE


STATE = {"velocity": 0, "velocity_prev": 1, "nonlinear": 2, "nonlinear_prev": 3, "pressure": 4,
         "pressure_bc": 5, "pressure_bc_prev": 6, "velocity_bc": 7}
TABLES = ["wn_p", "wn_v", "ad_p", "ad_v", "P_infl0", "P_infl1", "V_infl0", "V_infl1", "infl_div0", "infl_div1",
          "P_b", "V_b", "sigma0", "u_dpdx", "w_dpdz", "q_unit"]


class Channel:
    """One device context: Transforms + EqSolvePV + TimeStepper + Simulation state of a serial run."""

    def __init__(self, cfg, precision=FP64):
        self.cfg = dict(cfg)
        c = CdnsConfig()
        for k in ("NX", "NY", "NZ", "NXS", "NYS", "NZS", "NX2", "NY2", "NZ2", "INDEX_GRAD_PRE"):
            setattr(c, k, int(cfg[k]))
        for k in ("XL", "ZL", "THETA", "Re", "G_pre", "a_pre0", "b_pre0", "DT", "V_Microblow0"):
            setattr(c, k, float(cfg[k]))
        for j in range(3):
            c.a_V[j] = float(cfg["a_V"][j])
            c.b_V[j] = float(cfg["b_V"][j])
        c.precision = int(precision)
        h = ctypes.c_void_p()
        if lib().cdns_create(ctypes.byref(c), ctypes.byref(h)) != 0:
            raise RuntimeError("cdns_create failed (see stderr)")
        self._h = h
        m = self.cfg
        self.nzk, self.nj, self.nkx = 2 * m["NZS"] + 1, m["NYS"] + 1, m["NXS"] + 1
        self.it = 0

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(lib().cdns_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().cdns_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self):
        return lib().cdns_stream(self._h)

    def sync(self):
        self._check(lib().cdns_sync(self._h))

    # shapes (reference layouts)
    def scalar(self):
        return np.zeros((self.nzk, self.nj, self.nkx), np.complex128)

    def vector(self):
        return np.zeros((self.nzk, self.nj, self.nkx, 3), np.complex128)

    def boundary(self, comps=1):
        return np.zeros((self.nzk, 2, self.nkx) + ((comps,) if comps > 1 else ()), np.complex128)

    def fine_phys(self):
        m = self.cfg
        return np.zeros((m["NZ2"], m["NY2"] + 1, m["NX2"]), np.float64)

    # Transforms::channel_xyzfft2
    def xyzfft2(self, direction, phys, spec, stride=1, offset=0):
        where = _where(phys)
        if where == HOST:
            sp = ctypes.c_void_p(spec.ctypes.data + 16 * offset)
        else:
            sp = ctypes.c_void_p(spec.data_ptr() + 16 * offset)
        self._check(lib().cdns_xyzfft2(self._h, direction.encode(), _ptr(phys), sp, int(stride), where))

    def coarse_phys(self):
        m = self.cfg
        return np.zeros((m["NZ"], m["NY"] + 1, m["NX"]), np.float64)

    # Transforms::channel_xyzfft (coarse grid)
    def xyzfft(self, direction, phys, spec, stride=1, offset=0):
        where = _where(phys)
        if where == HOST:
            sp = ctypes.c_void_p(spec.ctypes.data + 16 * offset)
        else:
            sp = ctypes.c_void_p(spec.data_ptr() + 16 * offset)
        self._check(lib().cdns_xyzfft(self._h, direction.encode(), _ptr(phys), sp, int(stride), where))

    # FieldIO::to_physical: (velocity [3][NZ][NY+1][NX], pressure [NZ][NY+1][NX], finite)
    def to_physical(self):
        m = self.cfg
        vel = np.zeros((3, m["NZ"], m["NY"] + 1, m["NX"]), np.float64)
        pr = self.coarse_phys()
        fin = ctypes.c_int(0)
        self._check(lib().cdns_to_physical(self._h, _ptr(vel), _ptr(pr), ctypes.byref(fin), HOST))
        return vel, pr, bool(fin.value)

    # Statistics::reduce on the state: (sums [NY+1][20], finite [NY+1])
    def plane_moments(self, higher=True):
        n = self.cfg["NY"] + 1
        sums = np.zeros((n, 20), np.float64)
        fin = np.zeros(n, np.int32)
        self._check(lib().cdns_plane_moments(self._h, int(bool(higher)), _ptr(sums), _ptr(fin)))
        return sums, fin.astype(bool)

    # Diagnostics::evaluate on the state: {"ekt", "tke", "divergence_l2", "finite"}
    def diagnostics(self):
        out = np.zeros(3, np.float64)
        fin = ctypes.c_int(0)
        self._check(lib().cdns_diagnostics(self._h, _ptr(out), ctypes.byref(fin)))
        return {"ekt": float(out[0]), "tke": float(out[1]), "divergence_l2": float(out[2]), "finite": bool(fin.value)}

    # vector_ops
    def curl(self, v):
        out = self.vector()
        self._check(lib().cdns_curl(self._h, _ptr(out), _ptr(np.ascontiguousarray(v)), HOST))
        return out

    def cross_product(self, a, b):
        out = self.vector()
        self._check(lib().cdns_cross_product(self._h, _ptr(out), _ptr(np.ascontiguousarray(a)),
                                             _ptr(np.ascontiguousarray(b)), HOST))
        return out

    # TimeStepper::nonlinear_term
    def nonlinear_term(self, velocity):
        nl, bc = self.vector(), self.boundary()
        self._check(lib().cdns_nonlinear_term(self._h, _ptr(nl), _ptr(np.ascontiguousarray(velocity)), _ptr(bc), HOST))
        return nl, bc

    # EqSolvePV
    def setup(self, g_v):
        self._check(lib().cdns_setup(self._h, float(g_v)))

    def solve_modes(self, rhs_p, bc_p, rhs_v, bc_v, correct_sigma=True):
        P, V = self.scalar(), self.vector()
        rp = None if rhs_p is None else np.ascontiguousarray(rhs_p)
        self._check(lib().cdns_solve_modes(self._h, _ptr(P), _ptr(rp), _ptr(np.ascontiguousarray(bc_p)), _ptr(V),
                                           _ptr(np.ascontiguousarray(rhs_v)), _ptr(np.ascontiguousarray(bc_v)),
                                           int(bool(correct_sigma)), HOST))
        return P, V

    def table(self, name):
        w = TABLES.index(name)
        nzk, nj, nkx = self.nzk, self.nj, self.nkx
        shapes = [((nzk, nkx), np.float64), ((nzk, nkx), np.float64), ((nzk, nj, nkx), np.float64),
                  ((nzk, nj, nkx), np.float64), ((nzk, nj, nkx), np.complex128), ((nzk, nj, nkx), np.complex128),
                  ((nzk, nj, nkx, 3), np.complex128), ((nzk, nj, nkx, 3), np.complex128),
                  ((nzk, 2, nkx), np.complex128), ((nzk, 2, nkx), np.complex128), ((nzk, nj, nkx), np.complex128),
                  ((nzk, nj, nkx, 3), np.complex128), ((nzk, 2, nkx), np.complex128), ((nj,), np.float64),
                  ((nj,), np.float64), ((2,), np.float64)]
        shape, dt = shapes[w]
        out = np.zeros(shape, dt)
        self._check(lib().cdns_get_table(self._h, w, _ptr(out)))
        return out

    # Simulation state
    def _shape(self, name):
        return {"velocity": self.vector, "velocity_prev": self.vector, "nonlinear": self.vector,
                "nonlinear_prev": self.vector, "pressure": self.scalar, "pressure_bc": self.boundary,
                "pressure_bc_prev": self.boundary, "velocity_bc": lambda: self.boundary(3)}[name]()

    def set_state(self, name, data):
        if isinstance(data, np.ndarray):
            arr = self._shape(name)
            arr[...] = data
            self._check(lib().cdns_set_state(self._h, STATE[name], _ptr(arr), HOST))
        else:
            self._check(lib().cdns_set_state(self._h, STATE[name], _ptr(data), DEVICE))

    def get_state(self, name, out=None):
        if out is None:
            arr = self._shape(name)
            self._check(lib().cdns_get_state(self._h, STATE[name], _ptr(arr), HOST))
            return arr
        self._check(lib().cdns_get_state(self._h, STATE[name], _ptr(out), DEVICE))
        return out

    def step(self, it=None):
        it = self.it + 1 if it is None else it
        self._check(lib().cdns_step(self._h, int(it)))
        self.it = it

    def run(self, nsteps, it_begin=None):
        it_begin = self.it + 1 if it_begin is None else it_begin
        self._check(lib().cdns_run(self._h, int(it_begin), int(nsteps)))
        self.it = it_begin + nsteps - 1

    def flow_gradients(self):
        out = np.zeros(2)
        self._check(lib().cdns_flow_gradients(self._h, _ptr(out)))
        return out

    def time(self):
        return lib().cdns_time(self._h)

    # discrete adjoint (no reference counterpart)
    def set_body_force(self, force):
        """Constant body force added to every step's nonlinear term (None removes it)."""
        if force is None:
            self._check(lib().cdns_set_body_force(self._h, None, HOST))
        elif isinstance(force, np.ndarray):
            f = np.ascontiguousarray(force, np.complex128)
            self._check(lib().cdns_set_body_force(self._h, _ptr(f), HOST))
        else:
            self._check(lib().cdns_set_body_force(self._h, _ptr(force), DEVICE))

    def objective(self, which, velocity):
        """(J, dJ/du) of a velocity: which 0 kinetic energy, 1 mean wall shear."""
        u = np.ascontiguousarray(velocity, np.complex128)
        g = self.vector()
        J = ctypes.c_double()
        self._check(lib().cdns_objective(self._h, int(which), _ptr(u), _ptr(g), ctypes.byref(J), HOST))
        return J.value, g

    def adjoint_step(self, it, velocity, L, lf=None):
        """Transpose of step `it` at `velocity`: L = [adj u, adj u_prev, adj N_prev, adj B_prev] of
        the state after the step -> those of the state before it; lf accumulates the force gradient."""
        u = np.ascontiguousarray(velocity, np.complex128)
        out = [np.array(a, np.complex128) for a in L]
        f = None if lf is None else np.array(lf, np.complex128)
        self._check(lib().cdns_adjoint_step(self._h, int(it), _ptr(u), _ptr(out[0]), _ptr(out[1]), _ptr(out[2]),
                                            _ptr(out[3]), _ptr(f), HOST))
        return out, f

    def adjoint_run(self, u0, force, nsteps, k_ckpt, objective=0):
        """J(u(T)) and its gradients w.r.t. u0 and the body force (None: no force, zero gradient)."""
        u = np.ascontiguousarray(u0, np.complex128)
        f = None if force is None else np.ascontiguousarray(force, np.complex128)
        gu, gf = self.vector(), self.vector()
        J = ctypes.c_double()
        self._check(lib().cdns_adjoint_run(self._h, _ptr(u), _ptr(f), int(nsteps), int(k_ckpt), int(objective),
                                           _ptr(gu), _ptr(gf) if f is not None else None, ctypes.byref(J), HOST))
        self.it = nsteps
        return J.value, gu, gf

    def adjoint_run_device(self, u0, force, nsteps, k_ckpt, objective, grad_u0, grad_force, noalloc=False):
        """cdns_adjoint_run (or, with noalloc, cdns_adjoint_run_noalloc) on device arrays (torch
        tensors in the reference layout, complex as (re, im) pairs); returns J."""
        fn = lib().cdns_adjoint_run_noalloc if noalloc else lib().cdns_adjoint_run
        J = ctypes.c_double()
        self._check(fn(self._h, _ptr(u0), _ptr(force), int(nsteps), int(k_ckpt), int(objective), _ptr(grad_u0),
                       _ptr(grad_force), ctypes.byref(J), DEVICE))
        self.it = nsteps
        return J.value

    # per-kernel HIP-event timing
    def profile(self, enable=True):
        self._check(lib().cdns_profile(self._h, int(enable)))

    def profile_reset(self):
        self._check(lib().cdns_profile_reset(self._h))

    def profile_read(self):
        cap = 64
        names = ctypes.create_string_buffer(64 * cap)
        counts = np.zeros(cap, np.int64)
        tot = np.zeros(cap, np.float64)
        n = lib().cdns_profile_read(self._h, names, _ptr(counts), _ptr(tot), cap)
        if n < 0:
            raise RuntimeError(lib().cdns_last_error(self._h).decode())
        raw = names.raw
        out = {}
        for i in range(n):
            nm = raw[64 * i: 64 * (i + 1)].split(b"\0", 1)[0].decode()
            out[nm] = (int(counts[i]), float(tot[i]))
        return out
