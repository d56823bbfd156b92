"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: used only by tests, the smoke check and bench.py's
CPU-baseline leg, as the checker.
"""
import ctypes as C
import os

import numpy as np

from viral_amd.desc import ViralDesc, NFIELDS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: %s (run __graft_entry__.build())" % LIB_PATH)
        _lib = load(LIB_PATH)
    return _lib


_omp = None


def fast_lib():
    """The OpenMP build of the oracle (oracle/liboracle_omp.so, bitwise equal
    to the serial build: tests/test_oracle_openmp.py) for long reference runs;
    the serial build where it is missing."""
    global _omp
    if _omp is None:
        path = os.path.join(ROOT, "oracle", "liboracle_omp.so")
        _omp = load(path) if os.path.exists(path) else lib()
    return _omp


def load(path):
    """an oracle library at `path` with its prototypes set (the default one,
    or a variant the tests built, such as one with a planted mistake)"""
    L = C.CDLL(path)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)
    L.or_create.restype = vp
    L.or_create.argtypes = [C.POINTER(ViralDesc)]
    L.or_destroy.argtypes = [vp]
    L.or_field.restype = dp
    L.or_field.argtypes = [vp, C.c_int]
    L.or_cells.restype = C.c_size_t
    L.or_cells.argtypes = [vp]
    L.or_step.restype = C.c_int
    L.or_step.argtypes = [vp, C.c_double, C.c_double]
    L.or_flags.restype = C.POINTER(C.c_uint8)
    L.or_flags.argtypes = [vp]
    L.or_iters.restype = ip
    L.or_iters.argtypes = [vp, C.c_int]
    L.or_first_bad.restype = C.c_int
    L.or_first_bad.argtypes = [vp] + [C.POINTER(C.c_int)] * 4 + [C.POINTER(C.c_uint)]
    L.or_calc_source.argtypes = [vp, C.c_double, C.c_double, dp]
    L.or_hydro_explicit.argtypes = [vp, C.c_double, C.c_double, dp]
    L.or_root.restype = C.c_int
    L.or_root.argtypes = [vp, C.c_double, ip, C.POINTER(C.c_uint8)]
    L.or_boundary.argtypes = [vp]
    L.or_max_temp_gev.restype = C.c_double
    L.or_max_temp_gev.argtypes = [vp]
    for name in ("or_eos_p", "or_eos_dpde", "or_eos_T"):
        getattr(L, name).restype = C.c_double
        getattr(L, name).argtypes = [C.POINTER(ViralDesc), C.c_double]
    L.or_eos_s.restype = C.c_double
    L.or_eos_s.argtypes = [C.POINTER(ViralDesc), C.c_double, C.c_double, C.c_double]
    L.or_newton_cell.restype = C.c_double
    L.or_newton_cell.argtypes = [C.POINTER(ViralDesc)] + [C.c_double] * 11 + [ip, C.POINTER(C.c_uint)]
    for name in ("or_weno_L", "or_weno_R"):
        getattr(L, name).restype = C.c_double
        getattr(L, name).argtypes = [C.c_double] * 5
    L.or_weno_der.restype = C.c_double
    L.or_weno_der.argtypes = [C.c_double] * 6
    for name in ("or_vm_exp", "or_vm_log"):
        getattr(L, name).restype = C.c_double
        getattr(L, name).argtypes = [C.c_double]
    L.or_vm_pow.restype = C.c_double
    L.or_vm_pow.argtypes = [C.c_double, C.c_double]
    L.or_cart_pi10.argtypes = [dp, dp, dp]
    L.or_cart_gnewton.restype = C.c_int
    L.or_cart_gnewton.argtypes = [C.POINTER(ViralDesc), dp, dp, C.POINTER(C.c_uint)]
    L.or_cart_f.argtypes = [C.POINTER(ViralDesc), dp, dp, dp]
    L.or_cart_df.argtypes = [C.POINTER(ViralDesc), dp, dp, dp]
    L.or_centre_flux.argtypes = [vp, C.c_int, C.c_double, dp]
    L.or_shear_source.argtypes = [C.c_int, C.c_double, dp, dp, dp, C.c_double, C.c_double, dp]
    L.or_root_safeguard.restype = C.c_double
    L.or_root_safeguard.argtypes = [C.POINTER(ViralDesc), C.c_double, C.c_double, C.c_double, ip]
    lp = C.POINTER(C.c_longlong)
    L.or_check_root.restype = C.c_int
    L.or_check_root.argtypes = [vp, C.c_double, dp, dp, lp]
    L.or_debug_msg.argtypes = [vp, C.c_double, dp, lp]
    L.or_check_physics.argtypes = [vp, C.c_double, lp, lp]
    L.or_output.restype = C.c_size_t
    L.or_output.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
    L.or_ns_init.argtypes = [vp, C.c_double, C.c_double, C.c_int, dp]
    L.or_state_changed.argtypes = [vp]
    L.or_landau.restype = C.c_long
    L.or_landau.argtypes = [vp, C.c_double, dp, C.c_int, lp]
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def cart_pi10(u, p5):
    u = np.ascontiguousarray(u, dtype=np.float64)
    p5 = np.ascontiguousarray(p5, dtype=np.float64)
    out = np.zeros(10)
    lib().or_cart_pi10(_dp(u), _dp(p5), _dp(out))
    return out


def cart_gnewton(desc, prm, guess):
    """(root, iterations, flags) of the Cartesian 4-D solve"""
    prm = np.ascontiguousarray(prm, dtype=np.float64)
    x = np.array(guess, dtype=np.float64)
    fl = C.c_uint()
    it = lib().or_cart_gnewton(C.byref(desc), _dp(prm), _dp(x), C.byref(fl))
    return x, it, fl.value


def cart_f(desc, prm, x):
    prm = np.ascontiguousarray(prm, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(4)
    lib().or_cart_f(C.byref(desc), _dp(prm), _dp(x), _dp(y))
    return y


def cart_df(desc, prm, x):
    prm = np.ascontiguousarray(prm, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    J = np.zeros(16)
    lib().or_cart_df(C.byref(desc), _dp(prm), _dp(x), _dp(J))
    return J.reshape(4, 4)


def shear_source(milne, tau, u, pi10, du, SH, tpi):
    """du[a][b] = d_b u^a"""
    u = np.ascontiguousarray(u, dtype=np.float64)
    pi10 = np.ascontiguousarray(pi10, dtype=np.float64)
    du = np.ascontiguousarray(du, dtype=np.float64).ravel()
    out = np.zeros(10)
    lib().or_shear_source(int(milne), tau, _dp(u), _dp(pi10), _dp(du), SH, tpi, _dp(out))
    return out


class Oracle:
    """One event on the CPU oracle; fields are numpy views of its buffers."""

    def __init__(self, desc, L=None):
        self.desc = desc
        self.L = L if L is not None else lib()
        self.h = self.L.or_create(C.byref(desc))
        self.n = self.L.or_cells(self.h)
        self.dims = desc.ghost_dims()

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_destroy(self.h)
            self.h = None

    def field(self, f):
        p = self.L.or_field(self.h, f)
        return np.ctypeslib.as_array(p, shape=(self.n,))

    def set_state(self, state):
        """state: array (NFIELDS, n)"""
        for f in range(NFIELDS):
            self.field(f)[:] = state[f]
        self.L.or_state_changed(self.h)

    def state(self):
        return np.stack([self.field(f).copy() for f in range(NFIELDS)])

    def step(self, tau, dtau):
        return self.L.or_step(self.h, tau, dtau)

    def flags(self):
        return np.ctypeslib.as_array(self.L.or_flags(self.h), shape=(self.n,)).copy()

    def iters(self, stage):
        return np.ctypeslib.as_array(self.L.or_iters(self.h, stage), shape=(self.n,)).copy()

    def first_bad(self):
        i, j, k, s = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        fl = C.c_uint()
        r = self.L.or_first_bad(self.h, C.byref(i), C.byref(j), C.byref(k), C.byref(s), C.byref(fl))
        if r < 0:
            return None
        return (i.value, j.value, k.value, s.value, fl.value)

    def calc_source(self, tau, ts):
        out = np.zeros(15 * self.n)
        self.L.or_calc_source(self.h, tau, ts, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out.reshape(15, self.n)

    def hydro_explicit(self, tau, taustep):
        out = np.zeros(15 * self.n)
        self.L.or_hydro_explicit(self.h, tau, taustep, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out.reshape(15, self.n)

    def root(self, tau):
        it = np.zeros(self.n, dtype=np.int32)
        fl = np.zeros(self.n, dtype=np.uint8)
        bad = self.L.or_root(self.h, tau, it.ctypes.data_as(C.POINTER(C.c_int32)),
                             fl.ctypes.data_as(C.POINTER(C.c_uint8)))
        return bad, it, fl

    def boundary(self):
        self.L.or_boundary(self.h)

    def centre_flux(self, d, tau):
        out = np.zeros(15 * self.n)
        self.L.or_centre_flux(self.h, d, tau, _dp(out))
        return out.reshape(15, self.n)

    def max_temp_gev(self):
        return self.L.or_max_temp_gev(self.h)

    def check_root(self, tau):
        """(max, avg, at) of CheckRoot's statistics (6 Milne, 9 Cartesian)"""
        mx, avg = np.zeros(9), np.zeros(9)
        at = np.zeros(9, dtype=np.int64)
        ns = self.L.or_check_root(self.h, tau, _dp(mx), _dp(avg), at.ctypes.data_as(C.POINTER(C.c_longlong)))
        return mx[:ns], avg[:ns], at[:ns]

    def debug_msg(self, tau):
        """({vmax, vx, vy, ve, emax, emin}, cells {vmax, emax, emin})"""
        out = np.zeros(6)
        at = np.zeros(3, dtype=np.int64)
        self.L.or_debug_msg(self.h, tau, _dp(out), at.ctypes.data_as(C.POINTER(C.c_longlong)))
        return out, at

    def check_physics(self, tau):
        """(counts, first cells) of |v| >= 1 and e <= 0"""
        n = np.zeros(2, dtype=np.int64)
        first = np.zeros(2, dtype=np.int64)
        self.L.or_check_physics(self.h, tau, n.ctypes.data_as(C.POINTER(C.c_longlong)),
                                first.ctypes.data_as(C.POINTER(C.c_longlong)))
        return n, first

    def output(self, freq=1, freqz=1, xy=0):
        m = self.L.or_output(self.h, freq, freqz, xy, None)
        if m == 0:
            return None
        out = np.zeros(m)
        self.L.or_output(self.h, freq, freqz, xy, _dp(out))
        return out

    def ns_pi(self, tau, ts):
        """CalcNS's Navier-Stokes pi of the current state, (10, cells); with
        bulk (11, cells), row 10 the Navier-Stokes PI"""
        out = np.zeros((11 if self.desc.bulk else 10, self.n))
        assert self.L.or_ns_init(self.h, tau, ts, 0, _dp(out)) == 0
        return out

    def ns_init(self, tau, ts, zero=False):
        """NSInit (or ZeroInit) in place"""
        assert self.L.or_ns_init(self.h, tau, ts, 2 if zero else 1, None) == 0

    def landau(self, tau, tmunu, at_start=True):
        """Landau matching of a (10, cells) T^{mu nu} field in place; returns
        (number of failing cells, first failing cell or -1)"""
        t = np.ascontiguousarray(tmunu, dtype=np.float64)
        first = np.zeros(1, dtype=np.int64)
        nbad = self.L.or_landau(self.h, tau, _dp(t), 1 if at_start else 0, first.ctypes.data_as(C.POINTER(C.c_longlong)))
        return int(nbad), int(first[0])
