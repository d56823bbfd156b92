"""Pins of the oracle pieces that the reference's own code cannot run here:
the RK1/2/3 orchestration (NEWEST10/time.h:300-478; time.h needs root.h,
which needs GSL), the k0 velocity broadcast (NEWEST10/root.h:426-449) and
the GUBSER mirror of u^1, u^2 in the ghosts (NEWEST10/boundary.h:32-37,
77-81, 118-122, 160-165; the GUBSER profile runs on 6x6 ranks), and the bulk
source's time evolution (the BULKTEST profile integrates it, main.cpp:137-141).

- RK orchestration: Bjorken flow (the reference's BJORKEN profile,
  NEWEST10/maindef.h:200-232: conformal EOS, e0 = 30 GeV/fm^3 at tau0 = 0.6,
  eta = 0.2 s, tau_pi = 0.005) integrated to tau = 1 must converge to the
  textbook 0+1D solution at order 1, 2 and 3 for RK1, RK2 and RK3: the ideal
  closed form e = e0 (tau0/tau)^(4/3), and for Israel-Stewart
      de/dtau   = -(4/3) e/tau + Phi/tau,
      dPhi/dtau = -Phi/tau_pi + (4/3) eta/(tau tau_pi) - (4/3) Phi/tau,
  Phi = -tau^2 pi^{eta eta} = 2 pi^{xx} = 2 pi^{yy}, integrated by scipy's
  Radau.  The flow has u = (1,0,0,0), so the time derivative of u (the
  stage time steps of CalcSource and the prevu bookkeeping, time.h:358,418)
  is pinned by the viscous Gubser flow at RK1/2/3: its error must shrink
  under refinement.
- Bulk: the BULKTEST profile (BJORKEN with -DBULK, zeta = 1, tau_PI = 0.1)
  against the same equations with dPI/dtau = -(PI + zeta/tau)/tau_PI
  - (4/3) PI/tau and the energy equation's e + P - PI, at orders 1, 2, 3.
- k0 broadcast: after the root solve every eta slice carries slice k0's
  (vx, vy, ve), with u rebuilt from them at the solve's tau; e, P, T and the
  ghost cells are the unbroadcast solve's.
- GUBSER mirror: the reference's boundary() run on its 6x6 rank grid under
  MPICH's mpiexec (oracle/_ref, this container) and the committed fixture of
  that run (tests/golden/ref_mirror.npz, everywhere).

Each pin is shown to fail on an oracle with a planted mistake (OR_MUTANT in
oracle/oracle.c, built into a temporary directory).
"""
import os
import subprocess

import numpy as np
import pytest
from scipy.integrate import solve_ivp

import oracle_ctypes
import refdata
from gubser_viscous import ViscousGubser, central_errors
from viral_amd import ic
from viral_amd.desc import (make_desc, ETA_TIMES_S, TPI_CONST, Q_DEFAULT, Q_K0_BROADCAST, Q_GUBSER_MIRROR,
                            F_EN, F_P, F_TEMP, F_VX, F_VY, F_VE, F_U0, F_U1, F_U2, F_U3, F_PI0, NFIELDS,
                            GEVFM, OK)

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FACTOR = 3.0 * 42.25 * np.pi ** 2 / 90.0   # conformal e = FACTOR T^4 (NEWEST10/maindef.h:9-10)
E0, TAU0, TAU1 = 30.0 / GEVFM, 0.6, 1.0     # initBjorken (NEWEST10/init.h:294-345), TAUSTART
ETA_C, TPI = 0.2, 0.005                     # Feta = 0.2 s, Ftaupi = 0.005 (maindef.h:227-228)


# ------------------------------------------------------------------ Bjorken

def bjorken_exact(viscous):
    """(e, Phi) at TAU1 of the 0+1D equations in the module docstring"""
    if not viscous:
        return E0 * (TAU0 / TAU1) ** (4.0 / 3.0), 0.0

    def rhs(t, y):
        e, phi = y
        T = (e / FACTOR) ** 0.25
        eta = ETA_C * (4.0 / 3.0) * e / T
        return [-(4.0 / 3.0) * e / t + phi / t, -phi / TPI + (4.0 / 3.0) * eta / (t * TPI) - (4.0 / 3.0) * phi / t]

    sol = solve_ivp(rhs, (TAU0, TAU1), [E0, 0.0], method="Radau", rtol=1e-13, atol=1e-14)
    return sol.y[0, -1], sol.y[1, -1]


def bjorken_run(L, rk, nsteps, viscous):
    """the uniform Bjorken state through nsteps oracle steps from TAU0 to TAU1;
    returns (e, Phi) of a central cell and the largest spread over the grid"""
    d = make_desc(4, 4, 3, rk=rk, eta_law=ETA_TIMES_S, eta_coef=ETA_C if viscous else 0.0, tpi_law=TPI_CONST,
                  tpi_coef=TPI)
    n = d.cells()
    e = np.full(n, E0)
    u = np.zeros((4, n))
    u[0] = 1.0
    st = ic.assemble(d, e, e / 3.0, u, np.zeros((10, n)), u.copy(), TAU0, Temp=ic.conformal_T(d, e))
    o = oracle_ctypes.Oracle(d, L)
    o.set_state(st)
    ts = (TAU1 - TAU0) / nsteps
    for q in range(nsteps):
        assert o.step(TAU0 + q * ts, ts) == OK
    s = o.state().reshape((NFIELDS,) + d.ghost_dims())[(slice(None),) + d.interior_slices()]
    phi = -TAU1 * TAU1 * s[F_PI0 + 9]
    spread = max(np.ptp(s[F_EN]), np.ptp(phi))
    # Bjorken symmetry: pi^{xx} = pi^{yy} = Phi/2 (up to the integrator's
    # error: tau^2 pi^{eta eta} is not linear in the evolved pi^{eta eta}), the
    # other components zero, u at rest
    assert np.array_equal(s[F_PI0 + 4], s[F_PI0 + 7])
    assert np.allclose(s[F_PI0 + 4], phi / 2, rtol=2e-3, atol=1e-12 * E0)
    for l in (0, 1, 2, 3, 5, 6, 8):
        assert np.abs(s[F_PI0 + l]).max() <= 1e-12 * E0, l
    assert np.array_equal(s[F_U0], np.ones_like(s[F_U0]))
    return s[F_EN].flat[0], phi.flat[0], spread


# step counts in the asymptotic range: the ideal flow is smooth; the viscous
# one relaxes on tau_pi = 0.005, so dtau = 0.002 .. 0.0005
NSTEPS = {0: (10, 20, 40), 1: (200, 400, 800)}


def bjorken_orders(L, rk, viscous):
    e_ref, phi_ref = bjorken_exact(viscous)
    errs = []
    for n in NSTEPS[viscous]:
        e, phi, spread = bjorken_run(L, rk, n, viscous)
        assert spread == 0.0, "the uniform state must stay uniform"
        errs.append(abs(e - e_ref) / e_ref + abs(phi - phi_ref) / e_ref)
    return errs, [np.log2(errs[q] / errs[q + 1]) for q in range(len(errs) - 1)]


def check_bjorken(L, rk, viscous):
    errs, orders = bjorken_orders(L, rk, viscous)
    print("bjorken rk=%d viscous=%d errors %s orders %s" % (rk, viscous, errs, orders))
    for p in orders:
        assert rk - 0.2 <= p <= rk + 0.4, (rk, viscous, errs, orders)
    assert errs[-1] < {1: 3e-3, 2: 1e-5, 3: 5e-8}[rk], errs


@pytest.mark.parametrize("viscous", [0, 1], ids=["ideal", "israel_stewart"])
@pytest.mark.parametrize("rk", [1, 2, 3])
def test_bjorken_converges_at_the_rk_order(rk, viscous):
    check_bjorken(oracle_ctypes.lib(), rk, viscous)


# the BULKTEST profile (NEWEST10/maindef.h:157-196): the BJORKEN flow with the
# bulk source on, zeta = 1, tau_PI = 0.1 (initBulktest, init.h:356-413)
ZETA, TPIB = 1.0, 0.1


def bulk_exact():
    """(e, Phi, PI) at TAU1 of
        de/dtau   = -(e + P - PI)/tau + Phi/tau,
        dPhi/dtau = as above,
        dPI/dtau  = -(PI + zeta/tau)/tau_PI - (4/3) PI/tau"""
    def rhs(t, y):
        e, phi, PI = y
        T = (e / FACTOR) ** 0.25
        eta = ETA_C * (4.0 / 3.0) * e / T
        return [-(e + e / 3.0 - PI) / t + phi / t,
                -phi / TPI + (4.0 / 3.0) * eta / (t * TPI) - (4.0 / 3.0) * phi / t,
                -(PI + ZETA / t) / TPIB - (4.0 / 3.0) * PI / t]

    sol = solve_ivp(rhs, (TAU0, TAU1), [E0, 0.0, 0.0], method="Radau", rtol=1e-13, atol=1e-14)
    return sol.y[:, -1]


def bulk_run(L, rk, nsteps):
    from viral_amd.desc import ZETA_CONST, TPIB_CONST, F_BULK
    d = make_desc(4, 4, 3, rk=rk, eta_law=ETA_TIMES_S, eta_coef=ETA_C, tpi_law=TPI_CONST, tpi_coef=TPI, bulk=1,
                  zeta_law=ZETA_CONST, zeta_coef=ZETA, tpib_law=TPIB_CONST, tpib_coef=TPIB)
    n = d.cells()
    e = np.full(n, E0)
    u = np.zeros((4, n))
    u[0] = 1.0
    st = ic.assemble(d, e, e / 3.0, u, np.zeros((10, n)), u.copy(), TAU0, Temp=ic.conformal_T(d, e))
    o = oracle_ctypes.Oracle(d, L)
    o.set_state(st)
    ts = (TAU1 - TAU0) / nsteps
    for q in range(nsteps):
        assert o.step(TAU0 + q * ts, ts) == OK
    s = o.state().reshape((NFIELDS,) + d.ghost_dims())[(slice(None),) + d.interior_slices()]
    for f in (F_EN, F_PI0 + 9, F_BULK):
        assert np.ptp(s[f]) == 0.0, f
    return s[F_EN].flat[0], -TAU1 * TAU1 * s[F_PI0 + 9].flat[0], s[F_BULK].flat[0]


def check_bulk(L, rk):
    ref = bulk_exact()
    errs = []
    for n in NSTEPS[1]:
        got = bulk_run(L, rk, n)
        errs.append(sum(abs(g - r) for g, r in zip(got, ref)) / ref[0])
    orders = [np.log2(errs[q] / errs[q + 1]) for q in range(len(errs) - 1)]
    print("bulk bjorken rk=%d errors %s orders %s (PI = %g)" % (rk, errs, orders, ref[2]))
    for p in orders:
        assert rk - 0.2 <= p <= rk + 0.4, (rk, errs, orders)
    assert errs[-1] < {1: 3e-3, 2: 1e-5, 3: 5e-8}[rk], errs


@pytest.mark.parametrize("rk", [1, 2, 3])
def test_bulk_bjorken_converges_at_the_rk_order(rk):
    """the bulk row (NEWEST10/Source.h:519-527) with the shear rows, the
    energy equation's -(P - PI) and the root solve's A + P - PI"""
    check_bulk(oracle_ctypes.lib(), rk)


# ------------------------------------------------------------------ Gubser, all RK orders

_GUBSER = {}


def gubser_solution():
    if "g" not in _GUBSER:
        _GUBSER["g"] = ViscousGubser()
    return _GUBSER["g"]


def gubser_errors(L, rk, n, dx, ts, t_end=1.2):
    g = gubser_solution()
    d = make_desc(n, n, 1, lbi=1, dx=dx, dy=dx, rk=rk, eta_coef=0.2, tpi_coef=5.0,
                  quirks=Q_DEFAULT | Q_GUBSER_MIRROR)
    o = oracle_ctypes.Oracle(d, L)
    o.set_state(g.state(d, 1.0, ts))
    t = 1.0
    for _ in range(int(round((t_end - 1.0) / ts))):
        assert o.step(t, ts) == OK
        t = t + ts
    return central_errors(d, o.state(), g.fields(d, t), t, rmax=1.0)


def check_gubser(L, rk):
    """dx and dtau halved: the (e, pi) errors relative to e + P must shrink
    (first order in dtau through the one-sided d_tau u of the shear source)"""
    coarse = gubser_errors(L, rk, 50, 0.1, 0.005)
    fine = gubser_errors(L, rk, 100, 0.05, 0.0025)
    print("gubser rk=%d errors coarse %s fine %s" % (rk, coarse, fine))
    bound = {1: (3e-3, 1.5e-3), 2: (2.5e-4, 1.5e-4), 3: (1.5e-4, 6e-5)}[rk]
    shrink = {1: (1.7, 1.7), 2: (4.0, 2.0), 3: (4.0, 3.0)}[rk]
    assert fine[0] < bound[0] and fine[1] < bound[1], (rk, fine)
    assert fine[0] < coarse[0] / shrink[0] and fine[1] < coarse[1] / shrink[1], (rk, coarse, fine)


@pytest.mark.parametrize("rk", [1, 3])
def test_viscous_gubser_converges_at_rk1_and_rk3(rk):
    """(RK2 is tests/test_gubser_viscous.py)"""
    check_gubser(oracle_ctypes.lib(), rk)


# ------------------------------------------------------------------ k0 broadcast

def check_k0_broadcast(L, tau=1.3):
    d_on = make_desc(6, 5, 7)                         # default quirks: broadcast on
    d_off = make_desc(6, 5, 7, quirks=Q_DEFAULT & ~Q_K0_BROADCAST)
    st = ic.random_smooth(d_on, seed=17, tau=tau, pi_frac=0.05, v_scale=0.4)
    res = []
    for d in (d_on, d_off):
        o = oracle_ctypes.Oracle(d, L)
        o.set_state(st)
        bad, _, _ = o.root(tau)
        assert bad < 0
        res.append(o.state().reshape((NFIELDS,) + d.ghost_dims()))
    on, off = res
    gx, gy, gz = d_on.ghost_dims()
    k0 = (gz - 1) // 2                                # NEWEST10/init.h:641-645
    I = d_on.interior_slices()
    ghost = np.ones((gx, gy, gz), bool)
    ghost[I] = False
    # e, P, T everywhere, and every field of the ghost cells, are the unbroadcast solve's
    for f in (F_EN, F_P, F_TEMP):
        assert np.array_equal(on[f], off[f]), f
    for f in range(NFIELDS):
        assert np.array_equal(on[f][ghost], off[f][ghost]), f
    # slice k0 keeps its own velocities; every other interior slice carries them
    ii, jj = I[0], I[1]
    for k in range(I[2].start, I[2].stop):
        for f in (F_VX, F_VY, F_VE):
            assert np.array_equal(on[f][ii, jj, k], off[f][ii, jj, k0]), (f, k)
        vx, vy, ve = (off[f][ii, jj, k0] for f in (F_VX, F_VY, F_VE))
        if k == k0:
            for f in (F_U0, F_U1, F_U2, F_U3):
                assert np.array_equal(on[f][ii, jj, k], off[f][ii, jj, k]), (f, k)
            continue
        u0 = 1.0 / np.sqrt(1.0 - vx * vx - vy * vy - tau * tau * ve * ve)
        assert np.array_equal(on[F_U0][ii, jj, k], u0), k
        for f, v in ((F_U1, vx), (F_U2, vy), (F_U3, ve)):
            assert np.array_equal(on[f][ii, jj, k], u0 * v), (f, k)
        # u^mu u_mu = 1 in Milne coordinates at this tau
        uu = on[F_U0][ii, jj, k] ** 2 - on[F_U1][ii, jj, k] ** 2 - on[F_U2][ii, jj, k] ** 2 - \
            tau * tau * on[F_U3][ii, jj, k] ** 2
        assert np.abs(uu - 1.0).max() < 1e-13, k
    # the slices really differ without the broadcast (the check is not vacuous)
    assert not np.array_equal(off[F_VX][ii, jj, I[2].start], off[F_VX][ii, jj, k0])


def test_k0_broadcast_index_level():
    check_k0_broadcast(oracle_ctypes.lib())


# ------------------------------------------------------------------ GUBSER mirror

GUBSER = dict(refdata.PROFILES["ref_kt_gubser"], quirks=Q_DEFAULT | Q_GUBSER_MIRROR)
SIDES = ("xl", "xr", "yl", "yr")


def bands(a, d):
    gx, gy, _ = d.ghost_dims()
    a = a.reshape(gx, gy)
    B = 3
    return dict(xl=a[0:2 * B, B:gy - B], xr=a[gx - 2 * B:gx, B:gy - B],
                yl=a[B:gx - B, 0:2 * B], yr=a[B:gx - B, gy - 2 * B:gy])


def ghost_rows(side):
    """rows (x sides) or columns (y sides) of a 6-wide band that are ghosts"""
    return slice(0, 3) if side.endswith("l") else slice(3, 6)


def check_mirror_fixture(L):
    """the oracle's ghost fill of u^1, u^2 on the committed bands of the
    reference's 36-rank boundary() run"""
    z = np.load(os.path.join(HERE, "golden", "ref_mirror.npz"))
    d = make_desc(**GUBSER)
    gx, gy, _ = d.ghost_dims()
    n = d.cells()
    st = np.zeros((NFIELDS, n))
    st[F_U0] = 1.0
    B = 3
    for name, f in (("u1", F_U1), ("u2", F_U2)):
        a = st[f].reshape(gx, gy)
        a[0:2 * B, B:gy - B] = z[name + "_xl_before"]
        a[gx - 2 * B:gx, B:gy - B] = z[name + "_xr_before"]
        a[B:gx - B, 0:2 * B] = z[name + "_yl_before"]
        a[B:gx - B, gy - 2 * B:gy] = z[name + "_yr_before"]
    o = oracle_ctypes.Oracle(d, L)
    o.set_state(st)
    o.boundary()
    after = o.state()
    for name, f in (("u1", F_U1), ("u2", F_U2)):
        got = bands(after[f], d)
        for side in SIDES:
            ref = z["%s_%s_after" % (name, side)]
            g = ghost_rows(side)
            if side.startswith("x"):
                assert np.array_equal(got[side][g], ref[g]), (name, side)
            else:
                assert np.array_equal(got[side][:, g], ref[:, g]), (name, side)
            # the mirror is not the zero-gradient copy (the check is not vacuous)
            before = z["%s_%s_before" % (name, side)]
            edge = before[3] if side == "xl" else before[2] if side == "xr" else None
            if edge is not None:
                assert not np.array_equal(ref[g], np.broadcast_to(edge, ref[g].shape))


def test_gubser_mirror_matches_reference_fixture():
    check_mirror_fixture(oracle_ctypes.lib())


def test_gubser_mirror_matches_reference_on_its_rank_grid():
    """the reference's own boundary() on the 6x6 rank grid of the GUBSER
    profile (mpi_init, NEWEST10/init.h:118-141): the (0,0) rank fills the left
    x and y ghosts, the (5,5) rank the right ones; the oracle's one-rank fill
    of the same state must equal them on those faces, every field"""
    if not refdata.available("ref_kt_gubser") or not refdata.mpi_available():
        pytest.skip("oracle/_ref/ref_kt_gubser or mpiexec absent (reference tree absent)")
    ranks = refdata.run_ranks("ref_kt_gubser", 5, 1.1, 0.005, 36)
    lo, hi = ranks[(0, 0)], ranks[(5, 5)]
    assert np.array_equal(lo["state"], hi["state"])
    d = make_desc(**GUBSER)
    o = oracle_ctypes.Oracle(d)
    o.set_state(lo["state"])
    o.boundary()
    mine = o.state().reshape((NFIELDS,) + d.ghost_dims())[..., 0]
    gx, gy, _ = d.ghost_dims()
    B = 3
    faces = {(0, 0): [(slice(0, B), slice(B, gy - B)), (slice(B, gx - B), slice(0, B))],
             (5, 5): [(slice(gx - B, gx), slice(B, gy - B)), (slice(B, gx - B), slice(gy - B, gy))]}
    for key, r in ((0, 0), lo), ((5, 5), hi):
        ref = r["after_boundary"].reshape((NFIELDS,) + d.ghost_dims())[..., 0]
        for fi in range(NFIELDS):
            if fi == F_TEMP:
                continue   # boundary() leaves Temp alone (it is not in the copied fields)
            for sl in faces[key]:
                assert np.array_equal(mine[fi][sl], ref[fi][sl]), (key, fi)


# ------------------------------------------------------------------ planted mistakes

# OR_MUTANT value -> (what it plants, the check that must fail)
MUTANTS = {
    1: ("RK3 second-stage weights 1/2, 1/2", lambda L: check_bjorken(L, 3, 0)),
    2: ("RK3 middle stage at tau + dtau", lambda L: check_bjorken(L, 3, 1)),
    3: ("RK3 third stage's d_tau u over dtau, not dtau/2", lambda L: check_gubser(L, 3)),
    4: ("RK3 final weights 1/4, 3/4", lambda L: check_bjorken(L, 3, 0)),
    5: ("RK2 final weight dtau L1", lambda L: check_bjorken(L, 2, 1)),
    6: ("prevu also saved in the second stage", lambda L: check_gubser(L, 2)),
    7: ("k0 broadcast from slice k0 + 1", check_k0_broadcast),
    8: ("k0 broadcast's u rebuilt without tau^2", check_k0_broadcast),
    9: ("GUBSER mirror without the reflection", check_mirror_fixture),
    10: ("GUBSER mirror of u^1 only", check_mirror_fixture),
    11: ("bulk source without the 4/3 PI theta term", lambda L: check_bulk(L, 2)),
}


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("oracle_mutants")


@pytest.mark.parametrize("m", sorted(MUTANTS))
def test_pin_fails_on_planted_mistake(m, mutant_dir):
    what, check = MUTANTS[m]
    path = os.path.join(str(mutant_dir), "liboracle_m%d.so" % m)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-DOR_MUTANT=%d" % m, "-o", path,
                    os.path.join(ROOT, "oracle", "oracle.c"), "-lm"], check=True)
    L = oracle_ctypes.load(path)
    with pytest.raises(AssertionError):
        check(L)
    print("planted mistake %d (%s) caught" % (m, what))
