"""Pins of the oracle's root solve (NEWEST10/root.h:212-450), which the
reference harness cannot run (root.h needs GSL):

* conformal EOS: f(e) = e - A + M^2/(A + e/3 - Pi) = 0 is the quadratic
  e^2/3 + e(2A/3 - Pi) - A(A - Pi) + M^2 = 0 (A = T^tt - pi^tt), whose
  positive root the Newton result must equal to a few ulps;
* round trip (e, u, pi) -> T^{tau mu} -> root -> (e, u);
* GSL newton semantics: iteration counts, the f' == 0 error, max_iter;
* the exact period-orbit shortcut: identical results to running the loop;
* the opt-in safeguard (VIRAL_Q_ROOT_SAFEGUARD, no counterpart in the
  reference): its bracketed solve equals the quadratic's positive root, and
  cells whose Newton solve lands on the unphysical branch are rescued while
  every other cell is untouched.
"""
import ctypes

import numpy as np

import oracle_ctypes
from viral_amd import ic
from viral_amd.desc import (make_desc, EOS_S95P, FLAG_SOLVER, FLAG_MAXITER, FLAG_ENEG, FLAG_RESCUED, F_EN, F_U0,
                            F_T00, F_PI0, F_BULK, NFIELDS, Q_DEFAULT, Q_ROOT_SAFEGUARD)

L = oracle_ctypes.lib()
L.or_set_newton_shortcut.argtypes = [ctypes.c_int]


def newton(d, tau, T, pi, PI, guess):
    it = ctypes.c_int32()
    fl = ctypes.c_uint()
    e = L.or_newton_cell(d, tau, T[0], T[1], T[2], T[3], pi[0], pi[1], pi[2], pi[3], PI, guess,
                         ctypes.byref(it), ctypes.byref(fl))
    return e, it.value, fl.value


def test_conformal_root_is_the_quadratic_root():
    d = make_desc(4, 4, 4)
    rng = np.random.default_rng(0)
    worst = 0.0
    for _ in range(3000):
        tau = rng.uniform(0.3, 5.0)
        A = rng.uniform(0.1, 50.0)
        PI = rng.uniform(-0.05, 0.05) * A
        pi = np.zeros(4)
        pi[0] = rng.uniform(-0.05, 0.05) * A
        # momentum below the causal bound
        frac = rng.uniform(0, 0.9)
        M = frac * A
        ang = rng.uniform(0, 2 * np.pi, 2)
        Mx, My, Me = M * np.sin(ang[0]) * np.cos(ang[1]), M * np.sin(ang[0]) * np.sin(ang[1]), M * np.cos(ang[0]) / tau
        T = np.array([A + pi[0], Mx, My, Me])
        M2 = Mx * Mx + My * My + tau * tau * Me * Me
        b = 2 * A / 3 - PI
        c = -A * (A - PI) + M2
        exact = (-b + np.sqrt(b * b - 4 * c / 3)) / (2.0 / 3)
        e, it, fl = newton(d, tau, T, pi, PI, guess=exact * rng.uniform(0.5, 1.5))
        assert fl & FLAG_SOLVER == 0
        worst = max(worst, abs(e - exact) / exact)
    assert worst < 1e-14


def test_round_trip_recovers_e_and_u():
    d = make_desc(7, 6, 5)
    tau = 1.2
    st = ic.random_smooth(d, seed=4, tau=tau)
    o = oracle_ctypes.Oracle(d)
    e0 = st[F_EN].copy()
    u0 = st[F_U0:F_U0 + 4].copy()
    st[F_EN] *= 1.3  # a poor guess
    o.set_state(st)
    bad, it, fl = o.root(tau)
    assert bad < 0
    I = np.zeros(d.ghost_dims(), bool)
    I[d.interior_slices()] = True
    I = I.ravel()
    got = o.state()
    assert np.max(np.abs(got[F_EN][I] - e0[I]) / e0[I]) < 1e-13
    # k0 broadcast copies slice k0's velocity into the others: compare there
    gx, gy, gz = d.ghost_dims()
    k0 = d.k0_index()
    U = got[F_U0:F_U0 + 4].reshape(4, gx, gy, gz)[:, 3:-3, 3:-3, k0]
    U0 = u0.reshape(4, gx, gy, gz)[:, 3:-3, 3:-3, k0]
    assert np.max(np.abs(U - U0)) < 1e-13
    assert it[I].min() >= 1


def test_s95p_round_trip():
    d = make_desc(5, 5, 5, eos=EOS_S95P)
    tau = 0.6
    eosp = lambda e: L.or_eos_p(d, e)
    st = ic.random_smooth(d, seed=8, tau=tau, e_scale=40.0, eos_p=eosp)
    e0 = st[F_EN].copy()
    st[F_EN] *= 0.8
    o = oracle_ctypes.Oracle(d)
    o.set_state(st)
    o.root(tau)
    I = np.zeros(d.ghost_dims(), bool)
    I[d.interior_slices()] = True
    I = I.ravel()
    assert np.max(np.abs(o.state()[F_EN][I] - e0[I]) / e0[I]) < 1e-12


def test_zero_derivative_is_a_solver_error():
    """GSL newton returns GSL_EZERODIV when f'(x) == 0 (the reference aborts).
    Conformal f'(e) = 1 - M^2/(3 (A + e/3)^2) vanishes at M^2 = 3 (A + e/3)^2."""
    d = make_desc(4, 4, 4)
    A, e = 2.0, 3.0
    M = np.sqrt(3.0) * (A + e / 3)
    T = np.array([A, M, 0.0, 0.0])
    # pick the guess so the double arithmetic gives f' == 0 exactly
    for guess in np.linspace(2.9, 3.1, 2001):
        den = A + guess / 3.0
        if 1 - ((M * M / (den * den)) * (1.0 / 3.0)) == 0.0:
            ee, it, fl = newton(d, 1.0, T, np.zeros(4), 0.0, guess)
            assert fl & FLAG_SOLVER and it == 1 and ee == guess
            return
    raise AssertionError("no guess with f' == 0 in double arithmetic")


def test_max_iter_flag():
    d = make_desc(4, 4, 4, max_iter=2)
    T = np.array([10.0, 3.0, 1.0, 0.2])
    e, it, fl = newton(d, 1.0, T, np.zeros(4), 0.0, guess=1.0)
    assert it == 2 and fl & FLAG_MAXITER


def test_period_shortcut_is_exact():
    """The period-orbit shortcut gives the same roots, counts and flags as
    running GSL's loop to max_iter (config 1's Gubser grid, where a fifth of
    the cells end in a period-2 orbit, and random cells)."""
    d = make_desc(64, 64, 8)
    st = ic.gubser(d, 1.0, 0.01)
    res = []
    for sc in (1, 0):
        L.or_set_newton_shortcut(sc)
        o = oracle_ctypes.Oracle(d)
        o.set_state(st)
        o.step(1.0, 0.01)
        res.append((o.state(), o.iters(0), o.iters(1), o.flags()))
    L.or_set_newton_shortcut(1)
    (a, b) = res
    assert np.array_equal(a[0].view(np.int64), b[0].view(np.int64))
    for q in (1, 2, 3):
        assert np.array_equal(a[q], b[q])
    assert (a[1] == 5000).sum() > 1000
    rng = np.random.default_rng(5)
    for _ in range(2000):
        T = np.array([rng.uniform(0.01, 30), rng.normal(0, 3), rng.normal(0, 3), rng.normal(0, 1)])
        T[0] += np.sqrt(T[1] ** 2 + T[2] ** 2 + T[3] ** 2) * 1.2
        g = rng.uniform(0.001, 50)
        L.or_set_newton_shortcut(1)
        r1 = newton(d, 1.0, T, np.zeros(4), 0.0, g)
        L.or_set_newton_shortcut(0)
        r0 = newton(d, 1.0, T, np.zeros(4), 0.0, g)
        L.or_set_newton_shortcut(1)
        assert r1 == r0 or (np.isnan(r1[0]) and np.isnan(r0[0]) and r1[1:] == r0[1:])


def test_safeguard_solve_is_the_quadratic_root():
    d = make_desc(4, 4, 4)
    ds = make_desc(4, 4, 4, eos=EOS_S95P)
    rng = np.random.default_rng(7)
    worst = 0.0
    for _ in range(2000):
        A = rng.uniform(0.05, 80.0)
        PI = rng.uniform(-0.05, 0.05) * A
        M2 = (rng.uniform(0, 0.95) * A) ** 2 * (A - PI) / A
        b = 2 * A / 3 - PI
        exact = (-b + np.sqrt(b * b - 4 * (M2 - A * (A - PI)) / 3)) / (2.0 / 3)
        it = ctypes.c_int32()
        e = L.or_root_safeguard(d, M2, A, PI, ctypes.byref(it))
        worst = max(worst, abs(e - exact) / exact)
        assert 0 < it.value <= 200
        # s95p: a root of f inside the bracket
        e = L.or_root_safeguard(ds, M2, A, PI, ctypes.byref(it))
        Pe = L.or_eos_p(ds, e)
        assert 0 < e <= A and abs(e - A + M2 / (A + Pe - PI)) <= 1e-12 * A
    assert worst < 1e-14
    # no bracket for an acausal cell
    assert np.isnan(L.or_root_safeguard(d, 1.01 * 4.0, 2.0, 0.0, None))


def test_safeguard_rescues_cells_on_the_unphysical_branch():
    """A Newton guess just above the pole of f at e = -3 (A - PI) runs to the
    quadratic's negative root (ENEG: the reference would abort); the
    safeguard re-solves those cells to the positive root and leaves every
    other cell bit for bit as it was."""
    tau = 1.1
    base = make_desc(9, 8, 6)
    st = ic.random_smooth(base, seed=8, tau=tau)
    gx, gy, gz = base.ghost_dims()
    planted = [((4 * gy + 5) * gz + 4), ((8 * gy + 3) * gz + 6)]
    for x in planted:
        st[F_EN, x] = -2.95 * (st[F_T00, x] - st[F_PI0, x] - st[F_BULK, x])
    runs = {}
    for q in (Q_DEFAULT, Q_DEFAULT | Q_ROOT_SAFEGUARD):
        d = make_desc(9, 8, 6, quirks=q)
        o = oracle_ctypes.Oracle(d)
        o.set_state(st)
        bad, it, fl = o.root(tau)
        runs[q] = (bad, fl, o.state())
    bad, fl, s0 = runs[Q_DEFAULT]
    assert bad == planted[0] and fl[planted[0]] & (FLAG_ENEG | FLAG_SOLVER)
    bad, fl, s1 = runs[Q_DEFAULT | Q_ROOT_SAFEGUARD]
    assert bad == -1
    for x in planted:
        assert fl[x] & FLAG_RESCUED and not fl[x] & (FLAG_ENEG | FLAG_SOLVER)
        A, PI = st[F_T00, x] - st[F_PI0, x], st[F_BULK, x]
        M2 = sum((st[F_T00 + m, x] - st[F_PI0 + m, x]) ** 2 * (tau * tau if m == 3 else 1.0) for m in (1, 2, 3))
        b = 2 * A / 3 - PI
        exact = (-b + np.sqrt(b * b - 4 * (M2 - A * (A - PI)) / 3)) / (2.0 / 3)
        assert abs(s1[F_EN, x] - exact) <= 1e-14 * exact
    keep = np.ones(s0.shape[1], bool)
    keep[planted] = False
    assert np.array_equal(s0[:, keep], s1[:, keep])
