"""ctypes host layer over libviral.so (include/viral.h).

This is the Python side of the drop-in boundary: it owns no physics.  The
shared library is built in-tree by ``__graft_entry__.build()``; if it is
missing, loading raises -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from .desc import ViralDesc, NFIELDS, OK, ERR_PHYSICS

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libviral.so")

# every entry point declared in include/viral.h
EXPORTS = [
    "viral_desc_default", "viral_cells", "viral_ghost_dims", "viral_create", "viral_destroy",
    "viral_upload_state", "viral_download_state", "viral_download_field", "viral_step",
    "viral_flags", "viral_iterations", "viral_first_bad_cell", "viral_max_temp_gev",
    "viral_eos_eval", "viral_set_tile", "viral_get_tile", "viral_autotune", "viral_timers", "viral_timer_read",
    "viral_kernel_count", "viral_kernel_name", "viral_last_error", "viral_steps", "viral_get_launch",
    "viral_set_launch", "viral_tile_count", "viral_tile_table", "viral_xcd_tile", "viral_adjoint_begin",
    "viral_adjoint", "viral_adjoint_end", "viral_check_root", "viral_check_root_noalloc", "viral_debug_extrema",
    "viral_check_physics",
    "viral_output_size", "viral_output", "viral_sync", "viral_host_alloc", "viral_host_free", "viral_root",
    "viral_ns_init", "viral_landau",
]

# modes of viral_ns_init
NS_OUTPUT, NS_INIT, ZERO_INIT = 0, 1, 2

# kernel phases of the tile table (VIRAL_PHASE_*)
PHASE_FLUX_X, PHASE_FLUX_Y, PHASE_FLUX_ETA, PHASE_CELLS, PHASE_K0, PHASE_GHOST = range(6)


class ViralLaunch(C.Structure):
    """viral_launch: how the kernel phases of a stage are launched."""
    _fields_ = [("batched", C.c_int), ("graph", C.c_int), ("blocks_per_cu", C.c_int), ("fusion", C.c_int),
                ("xcd", C.c_int), ("flux_nt", C.c_int * 3)]

    def as_dict(self):
        return {"batched": self.batched, "graph": self.graph, "blocks_per_cu": self.blocks_per_cu,
                "fusion": self.fusion, "xcd": self.xcd, "flux_nt": list(self.flux_nt)}


_lib = None


class ViralError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("viral error %d: %s" % (code, msg))
        self.code = code


def load():
    """Load libviral.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libviral.so not built at %s; run __graft_entry__.build()" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp, dp = C.c_void_p, C.POINTER(C.c_double)
    pdesc = C.POINTER(ViralDesc)
    L.viral_desc_default.argtypes = [pdesc, C.c_int, C.c_int, C.c_int]
    L.viral_cells.restype = C.c_size_t
    L.viral_cells.argtypes = [pdesc]
    L.viral_ghost_dims.argtypes = [pdesc] + [C.POINTER(C.c_int)] * 3
    L.viral_create.restype = C.c_int
    L.viral_create.argtypes = [pdesc, C.c_int, C.POINTER(vp)]
    L.viral_destroy.argtypes = [vp]
    L.viral_upload_state.restype = C.c_int
    L.viral_upload_state.argtypes = [vp, C.c_int, C.POINTER(dp)]
    L.viral_download_state.restype = C.c_int
    L.viral_download_state.argtypes = [vp, C.c_int, C.POINTER(dp)]
    L.viral_download_field.restype = C.c_int
    L.viral_download_field.argtypes = [vp, C.c_int, C.c_int, dp]
    L.viral_step.restype = C.c_int
    L.viral_step.argtypes = [vp, C.c_double, C.c_double]
    L.viral_flags.restype = C.c_int
    L.viral_flags.argtypes = [vp, C.c_int, C.POINTER(C.c_uint8)]
    L.viral_iterations.restype = C.c_int
    L.viral_iterations.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int32)]
    L.viral_first_bad_cell.restype = C.c_int
    L.viral_first_bad_cell.argtypes = [vp, C.c_int] + [C.POINTER(C.c_int)] * 4 + [C.POINTER(C.c_uint)]
    L.viral_max_temp_gev.restype = C.c_int
    L.viral_max_temp_gev.argtypes = [vp, C.c_int, dp]
    L.viral_eos_eval.restype = C.c_int
    L.viral_eos_eval.argtypes = [pdesc, C.c_size_t] + [dp] * 5
    L.viral_set_tile.restype = C.c_int
    L.viral_set_tile.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.viral_get_tile.restype = C.c_int
    L.viral_get_tile.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.viral_autotune.restype = C.c_int
    L.viral_autotune.argtypes = [vp, C.c_int, C.c_int]
    L.viral_timers.restype = C.c_int
    L.viral_timers.argtypes = [vp, C.c_int]
    L.viral_timer_read.restype = C.c_int
    L.viral_timer_read.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]
    L.viral_kernel_count.restype = C.c_int
    L.viral_kernel_name.restype = C.c_char_p
    L.viral_kernel_name.argtypes = [C.c_int]
    L.viral_last_error.restype = C.c_char_p
    L.viral_last_error.argtypes = [vp]
    L.viral_steps.restype = C.c_int
    L.viral_steps.argtypes = [vp, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_int)]
    L.viral_get_launch.restype = C.c_int
    L.viral_get_launch.argtypes = [vp, C.POINTER(ViralLaunch)]
    L.viral_set_launch.restype = C.c_int
    L.viral_set_launch.argtypes = [vp, C.POINTER(ViralLaunch)]
    L.viral_tile_count.restype = C.c_longlong
    L.viral_tile_count.argtypes = [pdesc, C.c_int, C.c_int, C.c_int]
    L.viral_tile_table.restype = C.c_int
    L.viral_tile_table.argtypes = [vp, C.c_int, C.POINTER(C.c_longlong)]
    L.viral_xcd_tile.restype = C.c_longlong
    L.viral_xcd_tile.argtypes = [C.c_longlong, C.c_longlong]
    L.viral_adjoint_begin.restype = C.c_int
    L.viral_adjoint_begin.argtypes = [vp, C.c_int]
    L.viral_adjoint.restype = C.c_int
    L.viral_adjoint.argtypes = [vp, C.POINTER(dp), C.POINTER(dp), dp]
    L.viral_adjoint_end.restype = C.c_int
    L.viral_adjoint_end.argtypes = [vp]
    ip = C.POINTER(C.c_int)
    L.viral_check_root.restype = C.c_int
    L.viral_check_root.argtypes = [vp, C.c_int, C.c_double, dp, dp, ip, ip]
    L.viral_check_root_noalloc.restype = C.c_int
    L.viral_check_root_noalloc.argtypes = [vp, C.c_int, C.c_double, dp, dp, ip, ip]
    L.viral_debug_extrema.restype = C.c_int
    L.viral_debug_extrema.argtypes = [vp, C.c_int, C.c_double, dp, ip]
    L.viral_check_physics.restype = C.c_int
    L.viral_check_physics.argtypes = [vp, C.c_int, C.c_double, C.POINTER(C.c_longlong), ip]
    L.viral_output_size.restype = C.c_int
    L.viral_output_size.argtypes = [pdesc, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_size_t)]
    L.viral_output.restype = C.c_int
    L.viral_output.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, dp, C.c_int]
    L.viral_root.restype = C.c_int
    L.viral_root.argtypes = [vp, C.c_int, C.c_double]
    L.viral_ns_init.restype = C.c_int
    L.viral_ns_init.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_int, C.POINTER(C.POINTER(C.c_double))]
    L.viral_landau.restype = C.c_int
    L.viral_landau.argtypes = [vp, C.c_int, C.c_double, C.POINTER(C.POINTER(C.c_double)), C.c_int]
    L.viral_sync.restype = C.c_int
    L.viral_sync.argtypes = [vp]
    L.viral_host_alloc.restype = C.c_int
    L.viral_host_alloc.argtypes = [C.c_size_t, C.POINTER(vp)]
    L.viral_host_free.restype = C.c_int
    L.viral_host_free.argtypes = [vp]
    _lib = L
    return L


def _ptr_array(arrays):
    arr = (C.POINTER(C.c_double) * len(arrays))()
    for q, a in enumerate(arrays):
        arr[q] = a.ctypes.data_as(C.POINTER(C.c_double))
    return arr


class _Pinned:
    """Owner of one viral_host_alloc block (page-locked host memory), freed
    with the last array that views it."""

    def __init__(self, nbytes):
        self.L = load()
        self.p = C.c_void_p()
        r = self.L.viral_host_alloc(nbytes, C.byref(self.p))
        if r != OK:
            raise ViralError(r, "viral_host_alloc")

    def __del__(self):
        if self.p:
            self.L.viral_host_free(self.p)
            self.p = C.c_void_p()


def pinned_empty(shape):
    """A float64 array in page-locked host memory: host<->device copies of
    state-sized buffers (the adjoint's seed and gradient) run at full PCIe
    rate from it instead of through a staging copy."""
    count = int(np.prod(shape))
    owner = _Pinned(max(count, 1) * 8)
    buf = (C.c_double * max(count, 1)).from_address(owner.p.value)
    buf._owner = owner  # the ctypes buffer keeps the block alive
    return np.frombuffer(buf, dtype=np.float64, count=count).reshape(shape)


def tile_count(desc, phase, tp=64, tk=16):
    """Tiles (blocks' worth of work) of one event in one kernel phase."""
    return load().viral_tile_count(C.byref(desc), phase, tp, tk)


def xcd_tile(w, total):
    return load().viral_xcd_tile(w, total)


def eos_eval(desc, e):
    """(p, dp/de, T, s) of the descriptor's EOS at energy densities e (host)."""
    L = load()
    e = np.ascontiguousarray(e, dtype=np.float64).ravel()
    outs = [np.empty_like(e) for _ in range(4)]
    ptrs = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in [e] + outs]
    r = L.viral_eos_eval(C.byref(desc), e.size, *ptrs)
    if r != OK:
        raise ViralError(r, "viral_eos_eval")
    return outs


class Hydro:
    """A device context holding one or more events (the reference's GRID).

    Mirrors the reference's step seam: ``step(tau, ts)`` is
    TVDRK2/TVDRK3/FirstOrder(HydroGrid, tau, ts) of NEWEST10/time.h."""

    def __init__(self, descs):
        if isinstance(descs, ViralDesc):
            descs = [descs]
        self.L = load()
        self.descs = [d.copy() for d in descs]
        arr = (ViralDesc * len(descs))(*descs)
        h = C.c_void_p()
        r = self.L.viral_create(arr, len(descs), C.byref(h))
        if r != OK:
            raise ViralError(r, "viral_create failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.viral_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, r):
        if r != OK:
            raise ViralError(r, self.L.viral_last_error(self.h).decode())
        return r

    def cells(self, ev=0):
        return self.descs[ev].cells()

    def upload(self, state, ev=0):
        """state: (NFIELDS, cells) float64"""
        state = np.ascontiguousarray(state, dtype=np.float64)
        assert state.shape == (NFIELDS, self.cells(ev)), state.shape
        rows = [state[f] for f in range(NFIELDS)]
        self._chk(self.L.viral_upload_state(self.h, ev, _ptr_array(rows)))

    def download(self, ev=0):
        out = np.empty((NFIELDS, self.cells(ev)))
        rows = [out[f] for f in range(NFIELDS)]
        self._chk(self.L.viral_download_state(self.h, ev, _ptr_array(rows)))
        return out

    def field(self, f, ev=0):
        out = np.empty(self.cells(ev))
        self._chk(self.L.viral_download_field(self.h, ev, f, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def step(self, tau, ts, check=True):
        """One RK step; returns the status (raises on errors other than
        VIRAL_ERR_PHYSICS when check is set)."""
        r = self.L.viral_step(self.h, tau, ts)
        if check and r not in (OK, ERR_PHYSICS):
            self._chk(r)
        return r

    def flags(self, ev=0):
        out = np.empty(self.cells(ev), dtype=np.uint8)
        self._chk(self.L.viral_flags(self.h, ev, out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out

    def iterations(self, stage, ev=0):
        out = np.empty(self.cells(ev), dtype=np.int32)
        self._chk(self.L.viral_iterations(self.h, ev, stage, out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out

    def first_bad_cell(self, ev=0):
        i, j, k, s = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        fl = C.c_uint()
        r = self.L.viral_first_bad_cell(self.h, ev, C.byref(i), C.byref(j), C.byref(k), C.byref(s), C.byref(fl))
        if r == -1:
            return None
        self._chk(r)
        return (i.value, j.value, k.value, s.value, fl.value)

    def set_tile(self, d, tp, tk, ev=0):
        self._chk(self.L.viral_set_tile(self.h, ev, d, tp, tk))

    def get_tile(self, d, ev=0):
        tp, tk = C.c_int(), C.c_int()
        self._chk(self.L.viral_get_tile(self.h, ev, d, C.byref(tp), C.byref(tk)))
        return tp.value, tk.value

    def autotune(self, reps=3, ev=-1):
        """Time candidate launch configurations on the current state and keep
        the fastest: ev < 0 tunes the whole context (flux block size and tile
        shapes, fusion depth, persistent blocks, XCD order), ev >= 0 only that
        event's flux tile shapes.  Returns event 0's (or ev's) tile shapes."""
        self._chk(self.L.viral_autotune(self.h, ev, reps))
        return [self.get_tile(d, max(ev, 0)) for d in range(3)]

    def steps(self, tau, ts, n):
        """n steps with one synchronisation; returns (status, steps done
        before the first failure)."""
        done = C.c_int()
        r = self.L.viral_steps(self.h, tau, ts, n, C.byref(done))
        if r not in (OK, ERR_PHYSICS):
            self._chk(r)
        return r, done.value

    def launch(self):
        c = ViralLaunch()
        self._chk(self.L.viral_get_launch(self.h, C.byref(c)))
        return c

    def set_launch(self, **kw):
        """Change fields of the launch configuration (flux_nt as a 3-list)."""
        c = self.launch()
        for k, v in kw.items():
            if k == "flux_nt":
                for d in range(3):
                    c.flux_nt[d] = v[d]
            else:
                setattr(c, k, int(v))
        self._chk(self.L.viral_set_launch(self.h, C.byref(c)))

    def adjoint_begin(self, every):
        """Start recording steps for the adjoint, with a state checkpoint every
        `every` steps (every = 0: keep every stage state, no recompute)."""
        self._chk(self.L.viral_adjoint_begin(self.h, every))

    def adjoint(self, seed, out=None):
        """seed: (NFIELDS, cells) dJ/d(final state).  Returns (dJ/d(initial
        state) as (NFIELDS, cells), [dJ/d eta_coef, dJ/d tpi_coef]); the
        state gradient goes into `out` when given (e.g. a pinned_empty array,
        as the seed may be)."""
        seed = np.ascontiguousarray(seed, dtype=np.float64)
        assert seed.shape == (NFIELDS, self.cells(0)), seed.shape
        if out is None:
            grad = np.empty_like(seed)
        else:
            grad = out
            assert grad.shape == seed.shape and grad.dtype == np.float64 and grad.flags.c_contiguous
        pg = np.zeros(2)
        self._chk(self.L.viral_adjoint(self.h, _ptr_array([seed[f] for f in range(NFIELDS)]),
                                       _ptr_array([grad[f] for f in range(NFIELDS)]),
                                       pg.ctypes.data_as(C.POINTER(C.c_double))))
        return grad, pg

    def adjoint_end(self):
        self._chk(self.L.viral_adjoint_end(self.h))

    def tile_table(self, phase):
        out = (C.c_longlong * (len(self.descs) + 1))()
        self._chk(self.L.viral_tile_table(self.h, phase, out))
        return list(out)

    def timers(self, enable=True):
        self._chk(self.L.viral_timers(self.h, 1 if enable else 0))

    def timer_report(self):
        """{kernel name: (total ms, launches)} since timers were enabled"""
        out = {}
        for k in range(self.L.viral_kernel_count()):
            ms, n = C.c_double(), C.c_longlong()
            self._chk(self.L.viral_timer_read(self.h, k, C.byref(ms), C.byref(n)))
            out[self.L.viral_kernel_name(k).decode()] = (ms.value, n.value)
        return out

    def max_temp_gev(self, ev=0):
        v = C.c_double()
        self._chk(self.L.viral_max_temp_gev(self.h, ev, C.byref(v)))
        return v.value

    def root(self, tau, ev=0):
        """the root and velocity solve alone on the current state (status)"""
        r = self.L.viral_root(self.h, ev, tau)
        if r not in (OK, ERR_PHYSICS):
            self._chk(r)
        return r

    def ns_pi(self, tau, ts, ev=0):
        """CalcNS: the Navier-Stokes pi of every ghosted cell, (10, cells);
        (11, cells) with bulk, row 10 the Navier-Stokes PI"""
        n = self.descs[ev].cells()
        rows = 11 if self.descs[ev].bulk else 10
        out = np.zeros((rows, n))
        ptrs = (C.POINTER(C.c_double) * rows)(*[out[l].ctypes.data_as(C.POINTER(C.c_double)) for l in range(rows)])
        self._chk(self.L.viral_ns_init(self.h, ev, tau, ts, NS_OUTPUT, ptrs))
        return out

    def ns_init(self, tau, ts, zero=False, ev=0):
        """NSInit (or ZeroInit) on the current state in place"""
        self._chk(self.L.viral_ns_init(self.h, ev, tau, ts, ZERO_INIT if zero else NS_INIT, None))

    def landau(self, tau, tmunu, at_start=True, ev=0):
        """Landau matching of a (10, cells) T^{mu nu} field onto the state
        (status: OK, or ERR_PHYSICS with first_bad_cell naming a failed cell)"""
        t = np.ascontiguousarray(tmunu, dtype=np.float64)
        assert t.shape == (10, self.descs[ev].cells())
        ptrs = (C.POINTER(C.c_double) * 10)(*[t[l].ctypes.data_as(C.POINTER(C.c_double)) for l in range(10)])
        r = self.L.viral_landau(self.h, ev, tau, ptrs, 1 if at_start else 0)
        if r not in (OK, ERR_PHYSICS):
            self._chk(r)
        return r

    def _cells(self, ijk, ev):
        """ghosted linear indices of (i, j, k) triples, -1 where none"""
        gx, gy, gz = self.descs[ev].ghost_dims()
        t = np.asarray(ijk, dtype=np.int64).reshape(-1, 3)
        return np.where(t[:, 0] < 0, -1, (t[:, 0] * gy + t[:, 1]) * gz + t[:, 2])

    def check_root(self, tau, ev=0):
        """CheckRoot statistics: (max, avg, cells), 6 (Milne) or 9 (Cartesian)"""
        mx, avg = np.zeros(9), np.zeros(9)
        ijk = np.zeros(27, dtype=np.int32)
        ns = C.c_int()
        self._chk(self.L.viral_check_root(self.h, ev, tau, mx.ctypes.data_as(C.POINTER(C.c_double)),
                                          avg.ctypes.data_as(C.POINTER(C.c_double)),
                                          ijk.ctypes.data_as(C.POINTER(C.c_int)), C.byref(ns)))
        n = ns.value
        return mx[:n], avg[:n], self._cells(ijk[:3 * n], ev)

    def debug_extrema(self, tau, ev=0):
        """DebugMSG extrema: ({vmax, vx, vy, ve, emax, emin}, cells {vmax, emax, emin})"""
        out = np.zeros(6)
        ijk = np.zeros(9, dtype=np.int32)
        self._chk(self.L.viral_debug_extrema(self.h, ev, tau, out.ctypes.data_as(C.POINTER(C.c_double)),
                                             ijk.ctypes.data_as(C.POINTER(C.c_int))))
        return out, self._cells(ijk, ev)

    def check_physics(self, tau, ev=0):
        """CheckPhysics: (counts, first cells) of |v| >= 1 and e <= 0"""
        n = np.zeros(2, dtype=np.int64)
        ijk = np.zeros(6, dtype=np.int32)
        self._chk(self.L.viral_check_physics(self.h, ev, tau, n.ctypes.data_as(C.POINTER(C.c_longlong)),
                                             ijk.ctypes.data_as(C.POINTER(C.c_int))))
        return n, self._cells(ijk, ev)

    def output(self, freq=1, freqz=1, xy=0, ev=0):
        """write.h output records of the current state (17 doubles per cell)"""
        m = C.c_size_t()
        r = self.L.viral_output_size(C.byref(self.descs[ev]), freq, freqz, xy, C.byref(m))
        if r != OK:
            raise ViralError(r, "output strides must divide the grid")
        out = np.empty(m.value)
        self._chk(self.L.viral_output(self.h, ev, freq, freqz, xy, out.ctypes.data_as(C.POINTER(C.c_double)), 0))
        return out
