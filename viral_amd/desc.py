"""ctypes mirror of ``viral_desc`` and the field/flag constants of include/viral.h.

The reference fixes all of these at compile time through profile macros
(NEWEST10/maindef.h:36-339); here they are runtime fields of one descriptor.
"""
import ctypes as C

BORDER = 3
MILNE, CARTESIAN = 0, 1
KT, SHASTA = 0, 1
EOS_CONFORMAL, EOS_S95P = 0, 1
ETA_TIMES_S, ETA_S_OVER_4PI = 0, 1
TPI_OVER_EP, TPI_OVER_P, TPI_CONST = 0, 1, 2
ZETA_CONST, ZETA_S_OVER_4PI = 0, 1
TPIB_CONST, TPIB_OVER_P = 0, 1

Q_GHOST_EIGEN = 1 << 0
Q_K0_BROADCAST = 1 << 1
Q_DPDE_OF_P = 1 << 2
Q_GUBSER_MIRROR = 1 << 3
Q_SHEAR_SLICES = 1 << 4
Q_ROOT_SAFEGUARD = 1 << 5   # opt-in: bracketed re-solve of cells whose Newton solve fails
Q_FIX_VE = 1 << 6           # opt-in: the FIX build's FixTwo (v_eta zeroed after a step with BMax > 0.98)
Q_SHASTA_FULLSTEP = 1 << 7  # opt-in: SHASTA sub-step dtau (the reference's dtau/2, dtau/3 is unstable in 3-D)
Q_DEFAULT = Q_GHOST_EIGEN | Q_K0_BROADCAST | Q_DPDE_OF_P | Q_SHEAR_SLICES

(F_T00, F_T10, F_T20, F_T30) = range(4)
F_PI0 = 4
F_BULK = 14
F_EN, F_P, F_TEMP = 15, 16, 17
F_VX, F_VY, F_VE = 18, 19, 20
F_U0, F_U1, F_U2, F_U3 = 21, 22, 23, 24
F_PU0 = 25
NFIELDS = 29

FIELD_NAMES = (["T00", "T10", "T20", "T30"] + ["pi%d" % l for l in range(10)] +
               ["PI", "En", "P", "Temp", "Vx", "Vy", "Ve", "u0", "u1", "u2", "u3",
                "prevu0", "prevu1", "prevu2", "prevu3"])

FLAG_ACAUSAL = 1 << 0
FLAG_ENEG = 1 << 1
FLAG_SOLVER = 1 << 2
FLAG_MAXITER = 1 << 3
FLAG_FREEZEOUT = 1 << 4
FLAG_RESCUED = 1 << 5

OK, ERR_ARG, ERR_HIP, ERR_PHYSICS, ERR_UNSUPPORTED = 0, -1, -2, -3, -4

GEVFM = 0.1973


class ViralDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("lbi", C.c_int),
        ("coords", C.c_int),
        ("scheme", C.c_int),
        ("rk", C.c_int),
        ("eos", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("deta", C.c_double),
        ("dof", C.c_double),
        ("eta_law", C.c_int), ("eta_coef", C.c_double),
        ("tpi_law", C.c_int), ("tpi_coef", C.c_double),
        ("quirks", C.c_uint),
        ("k0", C.c_int),
        ("max_iter", C.c_int),
        ("t_freeze", C.c_double),
        ("fast_math", C.c_int),
        ("eos_table", C.c_int),
        ("bulk", C.c_int),
        ("zeta_law", C.c_int), ("zeta_coef", C.c_double),
        ("tpib_law", C.c_int), ("tpib_coef", C.c_double),
    ]

    def ghost_dims(self):
        gz = 1 if self.lbi else self.nz + 2 * BORDER
        return (self.nx + 2 * BORDER, self.ny + 2 * BORDER, gz)

    def cells(self):
        gx, gy, gz = self.ghost_dims()
        return gx * gy * gz

    def interior_slices(self):
        gx, gy, gz = self.ghost_dims()
        kz = slice(0, 1) if self.lbi else slice(BORDER, BORDER + self.nz)
        return (slice(BORDER, BORDER + self.nx), slice(BORDER, BORDER + self.ny), kz)

    def k0_index(self):
        if self.lbi:
            return 0
        if self.k0 >= 0:
            return self.k0
        return (self.nz + 2 * BORDER - 1) // 2

    def copy(self):
        d = ViralDesc()
        C.pointer(d)[0] = self
        return d


def make_desc(nx, ny, nz, **kw):
    """Reference defaults: Milne, KT, RK2, conformal EOS with DOF 42.25,
    eta = 0.2 s and tau_pi = 5 eta/(e+p) (the GUBSER profile's laws,
    NEWEST10/maindef.h:328-329), quirks q1-q3 on, Newton cap 5000."""
    d = ViralDesc()
    d.nx, d.ny, d.nz = nx, ny, nz
    d.lbi = 0
    d.coords = MILNE
    d.scheme = KT
    d.rk = 2
    d.eos = EOS_CONFORMAL
    d.dx = d.dy = d.deta = 0.1
    d.dof = 42.25
    d.eta_law, d.eta_coef = ETA_TIMES_S, 0.2
    d.tpi_law, d.tpi_coef = TPI_OVER_EP, 5.0
    d.quirks = Q_DEFAULT
    d.k0 = -1
    d.max_iter = 5000
    d.t_freeze = 0.0
    d.fast_math = 0
    d.eos_table = 0
    # bulk off; when on, the BULKTEST profile's zeta = 1, tau_PI = 0.1 (NEWEST10/maindef.h:190-193)
    d.bulk = 0
    d.zeta_law, d.zeta_coef = ZETA_CONST, 1.0
    d.tpib_law, d.tpib_coef = TPIB_CONST, 0.1
    for k, v in kw.items():
        if not hasattr(d, k):
            raise AttributeError(k)
        setattr(d, k, v)
    if d.lbi and d.nz != 1:
        raise ValueError("lbi grids have nz == 1")
    return d
