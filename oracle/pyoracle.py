"""ctypes binding of the CPU oracle (test infrastructure only).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker.  Arrays are numpy complex128/float64 in the reference's host layouts
(see cdns_oracle.h).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OrcConfig(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("NX", "NY", "NZ", "NXS", "NYS", "NZS", "NX2", "NY2", "NZ2")] + [
        (n, ctypes.c_double) for n in ("XL", "ZL", "THETA", "Re", "G_pre", "a_pre0", "b_pre0")
    ] + [("a_V", ctypes.c_double * 3), ("b_V", ctypes.c_double * 3), ("INDEX_GRAD_PRE", ctypes.c_int),
         ("DT", ctypes.c_double), ("V_Microblow0", ctypes.c_double)]


class Mt64(ctypes.Structure):
    _fields_ = [("mt", ctypes.c_uint64 * 312), ("idx", ctypes.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libcdns_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library not built: run make -C oracle")
        L = ctypes.CDLL(path)
        vp, d, i = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.POINTER(OrcConfig)]
        L.orc_destroy.argtypes = [vp]
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_set_threads.argtypes = [i]
        L.orc_set_threads.restype = i
        L.orc_tables.argtypes = [vp] + [vp] * 5
        L.orc_xyzfft2.argtypes = [vp, ctypes.c_char, vp, vp, i]
        L.orc_xyzfft2.restype = i
        L.orc_xyzfft.argtypes = [vp, ctypes.c_char, vp, vp, i]
        L.orc_xyzfft.restype = i
        L.orc_derivative.argtypes = [vp, i, vp, vp]
        for n in ("orc_curl", "orc_divergence", "orc_gradient", "orc_divergence_boundary"):
            getattr(L, n).argtypes = [vp, vp, vp]
        L.orc_flow_rate.argtypes = [vp, vp, i, i]
        L.orc_flow_rate.restype = d
        L.orc_cross_product.argtypes = [vp, vp, vp, vp]
        L.orc_nonlinear_term.argtypes = [vp, vp, vp, vp]
        L.orc_setup.argtypes = [vp, d]
        L.orc_solve_modes.argtypes = [vp, vp, vp, vp, vp, vp, vp, i]
        L.orc_solve_staged.argtypes = [vp, vp, vp, vp, vp, vp, vp, i]
        L.orc_solve_tau.argtypes = [vp, vp, vp, d, d, d, vp, i]
        L.orc_step_coefficients.argtypes = [i, d, d, vp]
        L.orc_get_table.argtypes = [vp, i, vp]
        L.orc_set_state.argtypes = [vp, i, vp]
        L.orc_get_state.argtypes = [vp, i, vp]
        L.orc_step.argtypes = [vp, i]
        L.orc_run.argtypes = [vp, i]
        L.orc_flow_gradients.argtypes = [vp, vp]
        L.orc_time.argtypes = [vp]
        L.orc_time.restype = d
        L.orc_set_body_force.argtypes = [vp, vp]
        L.orc_setup_g.argtypes = [vp]
        L.orc_setup_g.restype = d
        L.orc_adj_create.restype = vp
        L.orc_adj_create.argtypes = [vp, ctypes.POINTER(OrcConfig)]
        L.orc_adj_destroy.argtypes = [vp]
        L.orc_adj_fine_backward_T.argtypes = [vp, vp, vp]
        L.orc_adj_fine_forward_T.argtypes = [vp, vp, vp]
        L.orc_adj_curl_T.argtypes = [vp, vp, vp]
        L.orc_adj_solve_T.argtypes = [vp, vp, vp, vp]
        L.orc_adj_tl_step.argtypes = [vp, i, vp, vp, vp, vp]
        L.orc_adj_step.argtypes = [vp, i, vp, vp, vp]
        L.orc_adj_objective.argtypes = [vp, i, vp, vp]
        L.orc_adj_objective.restype = d
        L.orc_adj_run.argtypes = [vp, vp, vp, i, i, i, vp, vp]
        L.orc_adj_run.restype = d
        L.orc_plane_moments.argtypes = [vp, vp, vp, vp, vp, i, vp, vp]
        L.orc_cc_weights.argtypes = [vp, vp]
        L.orc_diagnostics.argtypes = [vp, vp, vp, vp, i, vp]
        L.orc_diagnostics.restype = i
        L.orc_mt64_seed.argtypes = [ctypes.POINTER(Mt64), ctypes.c_uint64]
        L.orc_mt64_next.argtypes = [ctypes.POINTER(Mt64)]
        L.orc_mt64_next.restype = ctypes.c_uint64
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


FIELDS = {"velocity": 0, "velocity_prev": 1, "nonlinear": 2, "nonlinear_prev": 3, "pressure": 4,
          "pressure_bc": 5, "pressure_bc_prev": 6, "velocity_bc": 7}


class Oracle:
    """The reference solver restated on the CPU, one serial context."""

    def __init__(self, cfg):
        self.cfg = dict(cfg)
        c = OrcConfig()
        for k in ("NX", "NY", "NZ", "NXS", "NYS", "NZS", "NX2", "NY2", "NZ2", "INDEX_GRAD_PRE"):
            setattr(c, k, int(cfg[k]))
        for k in ("XL", "ZL", "THETA", "Re", "G_pre", "a_pre0", "b_pre0", "DT", "V_Microblow0"):
            setattr(c, k, float(cfg[k]))
        for j in range(3):
            c.a_V[j] = float(cfg["a_V"][j])
            c.b_V[j] = float(cfg["b_V"][j])
        self._c = c
        self._h = lib().orc_create(ctypes.byref(c))
        if not self._h:
            raise ValueError(lib().orc_last_error().decode())
        m = self.cfg
        self.nzk, self.nj, self.nkx = 2 * m["NZS"] + 1, m["NYS"] + 1, m["NXS"] + 1

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # shapes
    def scalar(self):
        return np.zeros((self.nzk, self.nj, self.nkx), np.complex128)

    def vector(self):
        return np.zeros((self.nzk, self.nj, self.nkx, 3), np.complex128)

    def boundary(self, comps=1):
        return np.zeros((self.nzk, 2, self.nkx) + ((comps,) if comps > 1 else ()), np.complex128)

    def fine_phys(self):
        m = self.cfg
        return np.zeros((m["NZ2"], m["NY2"] + 1, m["NX2"]), np.float64)

    def coarse_phys(self):
        m = self.cfg
        return np.zeros((m["NZ"], m["NY"] + 1, m["NX"]), np.float64)

    def tables(self):
        WAVX, WAVZ = np.zeros(self.nkx), np.zeros(self.nzk)
        WL, WD, WR = np.zeros(self.nj), np.zeros(self.nj), np.zeros(self.nj)
        lib().orc_tables(self._h, _p(WAVX), _p(WAVZ), _p(WL), _p(WD), _p(WR))
        return dict(WAVX=WAVX, WAVZ=WAVZ, WL=WL, WD=WD, WR=WR)

    def xyzfft2(self, direction, phys, spec, stride=1, offset=0):
        """channel_xyzfft2 on flat buffers; `spec` is a complex128 array, `offset` in complex units."""
        assert phys.dtype == np.float64 and phys.flags.c_contiguous
        assert spec.dtype == np.complex128 and spec.flags.c_contiguous
        ptr = ctypes.c_void_p(spec.ctypes.data + 16 * offset)
        rc = lib().orc_xyzfft2(self._h, direction.encode(), _p(phys), ptr, int(stride))
        if rc != 0:
            raise ValueError(lib().orc_last_error().decode())

    def xyzfft(self, direction, phys, spec, stride=1, offset=0):
        ptr = ctypes.c_void_p(spec.ctypes.data + 16 * offset)
        rc = lib().orc_xyzfft(self._h, direction.encode(), _p(phys), ptr, int(stride))
        if rc != 0:
            raise ValueError(lib().orc_last_error().decode())

    def to_physical(self, velocity, pressure):
        """FieldIO::to_physical on the coarse grid: (velocity [3][NZ][NY+1][NX], pressure, finite)."""
        m = self.cfg
        vel = np.zeros((3, m["NZ"], m["NY"] + 1, m["NX"]), np.float64)
        pr = self.coarse_phys()
        v = np.ascontiguousarray(velocity)
        for k in range(3):
            self.xyzfft("B", vel[k], v, stride=3, offset=k)
        self.xyzfft("B", pr, np.ascontiguousarray(pressure))
        finite = bool(np.isfinite(v.view(np.float64)).all() and np.isfinite(np.asarray(pressure).view(np.float64)).all())
        return vel, pr, finite

    def plane_moments(self, u, v, w, p, higher=True):
        n = self.cfg["NY"] + 1
        sums = np.zeros((n, 20), np.float64)
        fin = np.zeros(n, np.int32)
        args = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w, p)]
        lib().orc_plane_moments(self._h, *[_p(a) for a in args], int(bool(higher)), _p(sums), _p(fin))
        return sums, fin.astype(bool)

    def cc_weights(self):
        wt = np.zeros(self.cfg["NY"] + 1)
        lib().orc_cc_weights(self._h, _p(wt))
        return wt

    def diagnostics(self, sums, plane_finite, velocity, spectral_finite=True):
        out = np.zeros(3)
        pf = np.ascontiguousarray(plane_finite, dtype=np.int32)
        fin = lib().orc_diagnostics(self._h, _p(np.ascontiguousarray(sums)), _p(pf), _p(np.ascontiguousarray(velocity)),
                                    int(bool(spectral_finite)), _p(out))
        return {"ekt": out[0], "tke": out[1], "divergence_l2": out[2], "finite": bool(fin)}

    def derivative(self, axis, f):
        out = self.scalar()
        lib().orc_derivative(self._h, int(axis), _p(out), _p(np.ascontiguousarray(f)))
        return out

    def curl(self, v):
        out = self.vector()
        lib().orc_curl(self._h, _p(out), _p(np.ascontiguousarray(v)))
        return out

    def divergence(self, v):
        out = self.scalar()
        lib().orc_divergence(self._h, _p(out), _p(np.ascontiguousarray(v)))
        return out

    def gradient(self, f):
        out = self.vector()
        lib().orc_gradient(self._h, _p(out), _p(np.ascontiguousarray(f)))
        return out

    def divergence_boundary(self, v):
        out = self.boundary()
        lib().orc_divergence_boundary(self._h, _p(out), _p(np.ascontiguousarray(v)))
        return out

    def flow_rate(self, f, comp=0):
        f = np.ascontiguousarray(f)
        stride = 3 if f.ndim == 4 else 1
        return lib().orc_flow_rate(self._h, _p(f), stride, comp)

    def cross_product(self, a, b):
        out = self.vector()
        lib().orc_cross_product(self._h, _p(out), _p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(b)))
        return out

    def nonlinear_term(self, velocity):
        nl, bc = self.vector(), self.boundary()
        lib().orc_nonlinear_term(self._h, _p(nl), _p(np.ascontiguousarray(velocity)), _p(bc))
        return nl, bc

    def setup(self, g_v):
        lib().orc_setup(self._h, float(g_v))

    def solve_modes(self, rhs_p, bc_p, rhs_v, bc_v, correct_sigma=True):
        P, V = self.scalar(), self.vector()
        rp = None if rhs_p is None else np.ascontiguousarray(rhs_p)
        lib().orc_solve_modes(self._h, _p(P), _p(rp), _p(np.ascontiguousarray(bc_p)), _p(V),
                              _p(np.ascontiguousarray(rhs_v)), _p(np.ascontiguousarray(bc_v)), int(correct_sigma))
        return P, V

    def solve_staged(self, rhs_p, bc_p, rhs_v, bc_v, correct_sigma=True):
        P, V = self.scalar(), self.vector()
        lib().orc_solve_staged(self._h, _p(P), _p(np.ascontiguousarray(rhs_p)), _p(np.ascontiguousarray(bc_p)), _p(V),
                               _p(np.ascontiguousarray(rhs_v)), _p(np.ascontiguousarray(bc_v)), int(correct_sigma))
        return P, V

    def solve_tau(self, rhs, g, a, b, bc, which=1):
        out = self.scalar()
        lib().orc_solve_tau(self._h, _p(out), _p(np.ascontiguousarray(rhs)), float(g), float(a), float(b),
                            _p(np.ascontiguousarray(bc)), int(which))
        return out

    TABLES = ["wn_p", "wn_v", "ad_p", "ad_v", "P_infl0", "P_infl1", "V_infl0", "V_infl1", "infl_div0",
              "infl_div1", "P_b", "V_b", "sigma0", "u_dpdx", "w_dpdz", "q_unit"]

    def table(self, name):
        w = self.TABLES.index(name)
        shapes = {0: ((self.nzk, self.nkx), np.float64), 1: ((self.nzk, self.nkx), np.float64),
                  2: ((self.nzk, self.nj, self.nkx), np.float64), 3: ((self.nzk, self.nj, self.nkx), np.float64),
                  4: ((self.nzk, self.nj, self.nkx), np.complex128), 5: ((self.nzk, self.nj, self.nkx), np.complex128),
                  6: ((self.nzk, self.nj, self.nkx, 3), np.complex128), 7: ((self.nzk, self.nj, self.nkx, 3), np.complex128),
                  8: ((self.nzk, 2, self.nkx), np.complex128), 9: ((self.nzk, 2, self.nkx), np.complex128),
                  10: ((self.nzk, self.nj, self.nkx), np.complex128), 11: ((self.nzk, self.nj, self.nkx, 3), np.complex128),
                  12: ((self.nzk, 2, self.nkx), np.complex128), 13: ((self.nj,), np.float64), 14: ((self.nj,), np.float64),
                  15: ((2,), np.float64)}
        shape, dt = shapes[w]
        out = np.zeros(shape, dt)
        lib().orc_get_table(self._h, w, _p(out))
        return out

    def _shape(self, name):
        return {"velocity": self.vector, "velocity_prev": self.vector, "nonlinear": self.vector,
                "nonlinear_prev": self.vector, "pressure": self.scalar, "pressure_bc": self.boundary,
                "pressure_bc_prev": self.boundary, "velocity_bc": lambda: self.boundary(3)}[name]()

    def set_state(self, name, data):
        arr = self._shape(name)
        arr[...] = data
        lib().orc_set_state(self._h, FIELDS[name], _p(arr))

    def get_state(self, name):
        arr = self._shape(name)
        lib().orc_get_state(self._h, FIELDS[name], _p(arr))
        return arr

    def step(self, it):
        lib().orc_step(self._h, int(it))

    def run(self, nsteps):
        lib().orc_run(self._h, int(nsteps))

    def set_body_force(self, f):
        """Constant body force added to every step's nonlinear term (None removes it)."""
        if f is None:
            lib().orc_set_body_force(self._h, None)
        else:
            f = np.ascontiguousarray(f, np.complex128)
            lib().orc_set_body_force(self._h, _p(f))

    def setup_g(self):
        return lib().orc_setup_g(self._h)

    def flow_gradients(self):
        out = np.zeros(2)
        lib().orc_flow_gradients(self._h, _p(out))
        return out

    @staticmethod
    def step_coefficients(it, dt, Re):
        out = np.zeros(5)
        lib().orc_step_coefficients(int(it), float(dt), float(Re), _p(out))
        return out


def _ptrs(arrs):
    return (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


class Adjoint:
    """Tangent-linear and discrete adjoint of the oracle's step (cdns_adjoint.c).

    A step state is the list [u, u_prev, N_prev, B_prev] (vector, vector, vector, boundary)."""

    def __init__(self, oracle):
        self.o = oracle
        self._h = lib().orc_adj_create(oracle._h, ctypes.byref(oracle._c))

    def close(self):
        if self._h:
            lib().orc_adj_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def state_like(self):
        o = self.o
        return [o.vector(), o.vector(), o.vector(), o.boundary()]

    def fine_backward_T(self, phys):
        phys = np.ascontiguousarray(phys, np.float64)
        out = self.o.scalar()
        lib().orc_adj_fine_backward_T(self._h, _p(phys), _p(out))
        return out

    def fine_forward_T(self, spec):
        spec = np.ascontiguousarray(spec, np.complex128)
        out = self.o.fine_phys()
        lib().orc_adj_fine_forward_T(self._h, _p(spec), _p(out))
        return out

    def curl_T(self, lw):
        lw = np.ascontiguousarray(lw, np.complex128)
        out = self.o.vector()
        lib().orc_adj_curl_T(self._h, _p(lw), _p(out))
        return out

    def solve_T(self, lv):
        lv = np.ascontiguousarray(lv, np.complex128)
        lr, lbc = self.o.vector(), self.o.boundary()
        lib().orc_adj_solve_T(self._h, _p(lv), _p(lr), _p(lbc))
        return lr, lbc

    def tl_step(self, it, X, dX, df=None):
        X = [np.ascontiguousarray(a, np.complex128) for a in X]
        dX = [np.ascontiguousarray(a, np.complex128) for a in dX]
        out = self.state_like()
        dfa = None if df is None else np.ascontiguousarray(df, np.complex128)
        lib().orc_adj_tl_step(self._h, int(it), _ptrs(X), _ptrs(dX), _p(dfa), _ptrs(out))
        return out

    def adj_step(self, it, X, L, lf=None):
        """Returns the adjoint of the state before step `it` (L is not modified) and lf + df."""
        X = [np.ascontiguousarray(a, np.complex128) for a in X]
        L = [np.array(a, np.complex128) for a in L]
        lf = self.o.vector() if lf is None else np.array(lf, np.complex128)
        lib().orc_adj_step(self._h, int(it), _ptrs(X), _ptrs(L), _p(lf))
        return L, lf

    def objective(self, which, vel):
        vel = np.ascontiguousarray(vel, np.complex128)
        g = self.o.vector()
        J = lib().orc_adj_objective(self._h, int(which), _p(vel), _p(g))
        return J, g

    def run(self, u0, f, nsteps, kc, objective):
        """J(u(T)) and its gradients with respect to u0 and the body force f (None: zero force)."""
        u0 = np.ascontiguousarray(u0, np.complex128)
        fa = None if f is None else np.ascontiguousarray(f, np.complex128)
        gu, gf = self.o.vector(), self.o.vector()
        J = lib().orc_adj_run(self._h, _p(u0), _p(fa), int(nsteps), int(kc), int(objective), _p(gu), _p(gf))
        return J, gu, gf


def set_threads(n):
    return lib().orc_set_threads(int(n))


class MT19937_64:
    def __init__(self, seed):
        self._g = Mt64()
        lib().orc_mt64_seed(ctypes.byref(self._g), ctypes.c_uint64(seed))

    def __call__(self):
        return lib().orc_mt64_next(ctypes.byref(self._g))
