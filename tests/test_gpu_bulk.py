"""GPU: the bulk-viscous build (viral_desc.bulk; the reference's -DBULK,
BULKTEST profile NEWEST10/maindef.h:157-196) against the CPU oracle.

The oracle's bulk row is pinned to the reference's own CalcSource
(NEWEST10/Source.h:519-527, tests/test_oracle_vs_reference.py, BULKTEST
harness profile) and to the 0+1D bulk Bjorken solution at RK1/2/3
(tests/test_oracle_pins.py).  Here the HIP step must equal the oracle bit for
bit in exact mode (state, Newton iteration counts, flags) on random states
carrying a bulk pressure, in Milne and Cartesian coordinates and in an
ensemble; fast mode within 1e-10 after one step; the Navier-Stokes PI of
viral_ns_init and the kept PI of viral_landau bit for bit; the adjoint with
the bulk source against finite differences."""
import numpy as np
import pytest

from viral_amd import capi, ic
from viral_amd.desc import (make_desc, NFIELDS, FIELD_NAMES, CARTESIAN, ETA_S_OVER_4PI, TPI_OVER_P, TPI_CONST,
                            ETA_TIMES_S, ZETA_CONST, ZETA_S_OVER_4PI, TPIB_CONST, TPIB_OVER_P, EOS_S95P, F_BULK,
                            GEVFM, OK, ERR_UNSUPPORTED)
import metrics
import oracle_ctypes
from test_gpu_parity import run_both, assert_bitwise, same_bits

pytestmark = pytest.mark.gpu

LAWS = {"bulktest": dict(zeta_law=ZETA_CONST, zeta_coef=1.0, tpib_law=TPIB_CONST, tpib_coef=0.1),
        "s_over_4pi": dict(zeta_law=ZETA_S_OVER_4PI, zeta_coef=0.5, tpib_law=TPIB_OVER_P, tpib_coef=1.5)}


@pytest.mark.parametrize("laws", sorted(LAWS))
@pytest.mark.parametrize("rk", [1, 2, 3])
def test_bulk_step_bitwise(rk, laws):
    d = make_desc(15, 13, 9, rk=rk, eta_law=ETA_S_OVER_4PI, eta_coef=1.0, tpi_law=TPI_OVER_P, tpi_coef=1.5,
                  bulk=1, **LAWS[laws])
    tau, ts = 0.8, 0.004
    st = ic.random_smooth(d, seed=60 + rk, tau=tau, bulk_frac=0.03)
    print("case bulk rk=%d laws=%s" % (rk, laws), flush=True)
    h, o = run_both(d, st, tau, ts, 2)
    assert_bitwise(h, o, rk)
    assert np.abs(h.download()[F_BULK]).max() > 0


def test_bulk_s95p_bitwise():
    d = make_desc(12, 11, 7, eos=EOS_S95P, bulk=1, **LAWS["s_over_4pi"])
    tau, ts = 0.6, 0.002
    eosp = lambda e: oracle_ctypes.lib().or_eos_p(d, e)
    st = ic.random_smooth(d, seed=66, tau=tau, e_scale=40.0, eos_p=eosp, bulk_frac=0.02)
    print("case bulk s95p", flush=True)
    h, o = run_both(d, st, tau, ts, 2)
    assert_bitwise(h, o, 2)


@pytest.mark.parametrize("rk", [2, 3])
def test_bulk_cartesian_bitwise(rk):
    d = make_desc(12, 11, 9, rk=rk, coords=CARTESIAN, bulk=1, **LAWS["bulktest"])
    tau, ts = 1.0, 0.004
    st = ic.random_smooth(d, seed=70 + rk, tau=tau, bulk_frac=0.03)
    print("case bulk cartesian rk=%d" % rk, flush=True)
    h, o = run_both(d, st, tau, ts, 2)
    assert_bitwise(h, o, rk)


def test_bulktest_profile_bitwise():
    """the BULKTEST profile itself: uniform e = 30 GeV/fm^3 at rest at
    tau0 = 0.6 (initBulktest, NEWEST10/init.h:356-413), 10x10x41, dtau = 0.01,
    eta = 0.2 s, tau_pi = 0.005, zeta = 1, tau_PI = 0.1: ten steps bitwise,
    and PI of the centre cell (what main.cpp:137-141 writes to temp.txt)"""
    d = make_desc(10, 10, 41, eta_law=ETA_TIMES_S, eta_coef=0.2, tpi_law=TPI_CONST, tpi_coef=0.005, bulk=1,
                  **LAWS["bulktest"])
    n = d.cells()
    e = np.full(n, 30.0 / GEVFM)
    u = np.zeros((4, n))
    u[0] = 1.0
    st = ic.assemble(d, e, e / 3.0, u, np.zeros((10, n)), u.copy(), 0.6, Temp=ic.conformal_T(d, e))
    print("case bulktest profile", flush=True)
    h, o = run_both(d, st, 0.6, 0.01, 10)
    assert_bitwise(h, o, 2)
    PI = h.download()[F_BULK].reshape(d.ghost_dims())
    c = PI[8, 8, 20]   # HydroGrid[XCM/2][YCM/2][ZCMA/2]
    assert c < 0 and np.all(PI[d.interior_slices()] == c)


def test_bulk_fast_mode_within_1e10():
    d = make_desc(24, 22, 9, fast_math=1, bulk=1, **LAWS["s_over_4pi"])
    dref = d.copy()
    dref.fast_math = 0
    tau, ts = 0.9, 0.005
    st = ic.random_smooth(d, seed=75, tau=tau, bulk_frac=0.03)
    print("case bulk fast", flush=True)
    h = capi.Hydro(d)
    h.upload(st)
    o = oracle_ctypes.Oracle(dref)
    o.set_state(st)
    assert h.step(tau, ts) == o.step(tau, ts) == OK
    errs = metrics.state_errors(d, h.download(), o.state())
    print(errs)
    assert metrics.worst(errs) <= 1e-10, errs


def test_bulk_ensemble_graph_bitwise():
    """two uneven events with different bulk laws, batched and graph-captured"""
    ds = [make_desc(13, 11, 7, bulk=1, **LAWS["bulktest"]), make_desc(9, 12, 9, bulk=1, **LAWS["s_over_4pi"])]
    tau, ts = 0.8, 0.004
    sts = [ic.random_smooth(d, seed=80 + q, tau=tau, bulk_frac=0.03) for q, d in enumerate(ds)]
    print("case bulk ensemble", flush=True)
    h = capi.Hydro(ds)
    h.set_launch(batched=1, graph=1)
    for q, st in enumerate(sts):
        h.upload(st, ev=q)
    r, done = h.steps(tau, ts, 3)
    assert (r, done) == (OK, 3)
    for q, (d, st) in enumerate(zip(ds, sts)):
        o = oracle_ctypes.Oracle(d)
        o.set_state(st)
        for s in range(3):
            assert o.step(tau + s * ts, ts) == OK
        g, c = h.download(ev=q), o.state()
        for f in range(NFIELDS):
            assert same_bits(g[f], c[f]), (q, FIELD_NAMES[f])


def test_bulk_ns_init_and_landau_bitwise():
    import landau_fields as lf
    d = make_desc(11, 10, 8, bulk=1, **LAWS["s_over_4pi"])
    tau, ts = 0.9, 0.01
    st = ic.random_smooth(d, seed=90, tau=tau)
    print("case bulk ns_init / landau", flush=True)
    h = capi.Hydro(d)
    h.upload(st)
    o = oracle_ctypes.Oracle(d)
    o.set_state(st)
    g, c = h.ns_pi(tau, ts), o.ns_pi(tau, ts)
    assert g.shape == (11, d.cells()) and np.array_equal(g.view(np.int64), c.view(np.int64))
    assert np.abs(c[10]).max() > 0
    h.ns_init(tau, ts)
    o.ns_init(tau, ts)
    gs, cs = h.download(), o.state()
    for f in range(NFIELDS):
        assert np.array_equal(gs[f].view(np.int64), cs[f].view(np.int64)), FIELD_NAMES[f]
    assert np.array_equal(gs[F_BULK], c[10])
    # Landau matching keeps PI = P - e/3 of the field (zero for a conformal fluid field plus noise)
    L = oracle_ctypes.lib()
    T, e, v, pi = lf.random_fluid_field(d.cells(), 5, 0.6, lambda x: np.array([L.or_eos_p(d, float(y))
                                                                            for y in np.atleast_1d(x)]), vmax=0.6)
    h.upload(st)
    o.set_state(st)
    assert h.landau(0.6, T, True) == OK
    nbad, _ = o.landau(0.6, T, True)
    assert nbad == 0
    gs, cs = h.download(), o.state()
    for f in range(NFIELDS):
        assert np.array_equal(gs[f].view(np.int64), cs[f].view(np.int64)), FIELD_NAMES[f]


@pytest.mark.parametrize("case", ["milne_rk2", "milne_rk3_s_over_4pi", "cartesian", "milne_s95p_table"])
def test_bulk_adjoint_matches_fd(case):
    """the adjoint with the bulk source (bulk_vjp: PI's relaxation, zeta(s)
    and tau_Pi(zeta, P) in reverse) along a random direction of the state,
    and the shear coefficients, against four-point central differences of
    the GPU forward (and of the CPU oracle in exact mode) to 1e-6"""
    from test_gpu_adjoint import run_gpu, run_cpu, direction, fd4, close, HP, TAU
    kw = dict(bulk=1, eta_law=ETA_S_OVER_4PI, eta_coef=1.0, tpi_law=TPI_OVER_P, tpi_coef=1.5)
    if case == "milne_rk2":
        d = make_desc(10, 9, 7, rk=2, **LAWS["bulktest"], **kw)
    elif case == "milne_rk3_s_over_4pi":
        d = make_desc(10, 9, 7, rk=3, **LAWS["s_over_4pi"], **kw)
    elif case == "cartesian":
        d = make_desc(10, 9, 8, rk=2, coords=CARTESIAN, **LAWS["s_over_4pi"], **kw)
    else:
        d = make_desc(10, 9, 7, rk=2, fast_math=1, eos=EOS_S95P, eos_table=1, **LAWS["s_over_4pi"], **kw)
    if d.eos == EOS_S95P:
        dref = d.copy()
        dref.fast_math = dref.eos_table = 0
        st = ic.random_smooth(d, seed=95, tau=TAU, pi_frac=0.02, bulk_frac=0.03, e_scale=30.0,
                              eos_p=lambda e: capi.eos_eval(dref, e)[0].reshape(np.shape(e)))
    else:
        st = ic.random_smooth(d, seed=95, tau=TAU, pi_frac=0.02, bulk_frac=0.03)
    rng = np.random.default_rng(96)
    seed = rng.standard_normal(st.shape)
    print("case bulk adjoint %s" % case, flush=True)
    _, grad, pg = run_gpu(d, st, 3, 2, seed)
    runs = [run_gpu] + ([run_cpu] if not d.fast_math else [])
    dx = direction(st, rng)
    ad = np.sum(grad * dx)
    for run in runs:
        fd = fd4(lambda h: np.sum(seed * run(d, st + h * dx, 3)))
        print("state", run.__name__, ad, fd, abs(ad - fd) / abs(fd), flush=True)
        assert close(ad, fd), (run.__name__, ad, fd)
    # the bulk pressure's own direction
    dP = np.zeros_like(st)
    dP[F_BULK] = rng.standard_normal(st.shape[1]) * np.abs(st[F_BULK]).max()
    fd = fd4(lambda h: np.sum(seed * run_gpu(d, st + h * dP, 3)))
    ad = np.sum(grad * dP)
    print("PI direction", ad, fd, flush=True)
    assert close(ad, fd), (ad, fd)
    for q, name in enumerate(["eta_coef", "tpi_coef"]):
        c0 = getattr(d, name)

        def J(h):
            dd = d.copy()
            setattr(dd, name, c0 * (1 + h))
            return np.sum(seed * run_gpu(dd, st, 3))

        fd = fd4(J, HP) / c0
        print(name, pg[q], fd, flush=True)
        assert close(pg[q], fd), (name, pg[q], fd)
