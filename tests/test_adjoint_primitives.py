"""CPU checks of the hand-written adjoint primitives (viral_amd/csrc/adjoint.h)
against central finite differences of the forward pieces they reverse.

The primitives are built for the host from the same header the GPU kernels
use (tests/host/adjoint_host.hip).  Each check is a dot-product test:
<VJP(ob), dx> == <ob, (f(x + h dx) - f(x - h dx)) / 2h> for random dx.
"""
import ctypes as C
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "host", "_build", "libadjoint_host.so")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="host adjoint library not built")


def lib():
    L = C.CDLL(LIB)
    dp = C.POINTER(C.c_double)
    L.ah_weno.restype = C.c_double
    L.ah_weno.argtypes = [dp, C.c_int]
    L.ah_weno_grad.argtypes = [dp, C.c_int, dp]
    L.ah_weno_lr_vjp.argtypes = [dp, C.c_double, C.c_double, dp, dp]
    L.ah_shear.argtypes = [C.c_int, C.c_double, dp, dp, dp, C.c_double, C.c_double, dp]
    L.ah_shear_vjp.argtypes = [C.c_int, C.c_double, dp, dp, dp, C.c_double, C.c_double, dp, dp, dp, dp, dp, dp]
    L.ah_flux_cell.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp, dp, dp]
    L.ah_flux_cell_vjp.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp, dp, dp]
    L.ah_source.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                            C.c_double, dp, dp, dp, dp]
    L.ah_source_vjp.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_double, dp, dp, dp, dp, dp]
    L.ah_update.restype = C.c_int
    L.ah_update.argtypes = [C.c_int, C.c_double, dp, C.c_double, dp]
    L.ah_update_vjp.argtypes = [C.c_int, C.c_double, dp, C.c_double, dp, dp]
    L.ah_eos.argtypes = [C.c_int, C.c_double, dp]
    L.ah_eos_d.argtypes = [C.c_int, C.c_double, dp]
    L.ah_tab.restype = C.c_int
    L.ah_tab.argtypes = [C.c_double, dp]
    L.ah_tab_d.restype = C.c_int
    L.ah_tab_d.argtypes = [C.c_double, dp, dp]
    return L


def P(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))


def arr(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def dot_check(f, x, vjp_x, ob, rng, h=1e-6, tol=1e-6, trials=4):
    """f: x -> y (numpy); vjp_x = J^T ob at x"""
    for _ in range(trials):
        dx = rng.standard_normal(x.size)
        fd = (f(x + h * dx) - f(x - h * dx)) / (2 * h)
        lhs = float(vjp_x @ dx)
        rhs = float(ob @ fd)
        assert abs(lhs - rhs) <= tol * (abs(lhs) + abs(rhs) + 1e-8), (lhs, rhs)


@pytest.mark.parametrize("right", [0, 1])
def test_weno_gradient(right):
    L = lib()
    rng = np.random.default_rng(1 + right)
    for scale in (1.0, 1e-3, 50.0):
        for _ in range(20):
            x = arr(scale * (1.0 + rng.standard_normal(5)))
            g = np.zeros(5)
            L.ah_weno_grad(P(x), right, P(g))
            for i in range(5):
                h = 1e-7 * scale
                xp, xm = x.copy(), x.copy()
                xp[i] += h
                xm[i] -= h
                num = (L.ah_weno(P(xp), right) - L.ah_weno(P(xm), right)) / (2 * h)
                assert abs(g[i] - num) <= 1e-5 * (1 + abs(num)), (i, g[i], num)
    # linear data: weights constant, gradient = linear candidate combination
    x = arr([1.0, 2.0, 3.0, 4.0, 5.0])
    g = np.zeros(5)
    L.ah_weno_grad(P(x), right, P(g))
    assert abs(g.sum() - 1.0) < 1e-12


def test_weno_pair_vjp_matches_single_gradients():
    """weno_lr_vjp (shared indicators) = sL dL/dx + sR dR/dx from two
    weno_grad calls, and its values are the reconstructions"""
    L = lib()
    rng = np.random.default_rng(4)
    for scale in (1e-3, 1.0, 50.0):
        for _ in range(30):
            x = arr(scale * (1.0 + rng.standard_normal(5)))
            sL, sR = rng.standard_normal(2)
            gl, gr, g, LR = np.zeros(5), np.zeros(5), np.zeros(5), np.zeros(2)
            L.ah_weno_grad(P(x), 0, P(gl))
            L.ah_weno_grad(P(x), 1, P(gr))
            L.ah_weno_lr_vjp(P(x), sL, sR, P(LR), P(g))
            want = sL * gl + sR * gr
            assert np.abs(g - want).max() <= 1e-12 * (1 + np.abs(want).max()), (g, want)
            assert abs(LR[0] - L.ah_weno(P(x), 0)) <= 1e-13 * (1 + abs(LR[0]))
            assert abs(LR[1] - L.ah_weno(P(x), 1)) <= 1e-13 * (1 + abs(LR[1]))


def rand_u(rng, tau, milne=1):
    s = rng.uniform(-0.5, 0.5, 3)
    g3 = tau * tau if milne else 1.0
    return np.array([np.sqrt(1 + s[0] ** 2 + s[1] ** 2 + g3 * s[2] ** 2), *s])


@pytest.mark.parametrize("milne", [0, 1])
def test_shear_source_vjp(milne):
    L = lib()
    rng = np.random.default_rng(10 + milne)
    for _ in range(6):
        tau = rng.uniform(0.5, 2.0) if milne else 1.0
        u = rand_u(rng, tau, milne)
        pi = 0.1 * rng.standard_normal(10)
        du = rng.standard_normal(16)
        SH, tpi = rng.uniform(0.1, 1.0), rng.uniform(0.3, 2.0)
        x = np.concatenate([u, pi, du, [SH, tpi]])

        def f(z):
            out = np.zeros(10)
            L.ah_shear(milne, tau, P(z[0:4]), P(z[4:14]), P(z[14:30]), z[30], z[31], P(out))
            return out
        ob = rng.standard_normal(10)
        ub, pib, dub = np.zeros(4), np.zeros(10), np.zeros(16)
        SHb, tpib = C.c_double(0), C.c_double(0)
        L.ah_shear_vjp(milne, tau, P(u), P(pi), P(du), SH, tpi, P(ob), P(ub), P(pib), P(dub), C.byref(SHb),
                       C.byref(tpib))
        vjp = np.concatenate([ub, pib, dub, [SHb.value, tpib.value]])
        dot_check(f, x, vjp, ob, rng)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_flux_cell_vjp(d):
    L = lib()
    rng = np.random.default_rng(20 + d)
    tau = 1.3
    for _ in range(5):
        u = rand_u(rng, tau)
        x = np.concatenate([[rng.uniform(1, 5), rng.uniform(0.3, 1.5), 0.01 * rng.standard_normal()], u,
                            0.1 * rng.standard_normal(10), rng.standard_normal(4)])

        def f(z):
            out = np.zeros(31)
            L.ah_flux_cell(d, tau, z[0], z[1], z[2], P(z[3:7]), P(z[7:17]), P(z[17:21]), P(out))
            return out
        ob = rng.standard_normal(31)
        inb = np.zeros(21)
        L.ah_flux_cell_vjp(d, tau, x[0], x[1], x[2], P(x[3:7]), P(x[7:17]), P(ob), P(inb))
        dot_check(f, x, inb, ob, rng)


# s95p branch points of p, s and T, in 1/fm^4 (GeV/fm^3 / GEVFM)
S95P_BREAKS = np.array([1e-5, 0.5028563305441270, 1.62, 1.86, 9.9878355786273545, 0.1270769021427449,
                        0.446707952467404, 1.9402832534193788, 3.7292474570977285, 0.5143939846236409]) / 0.1973
S95P_BREAKS[0] = 1e-5


def test_s95p_closed_form_derivatives():
    """dP/de, dT/de, ds/de of the s95p fits on every branch against central
    differences of the fits themselves"""
    L = lib()
    rng = np.random.default_rng(50)
    es = np.exp(rng.uniform(np.log(2e-6), np.log(2000.0), 400))
    es = np.concatenate([es, S95P_BREAKS * 1.01, S95P_BREAKS * 0.99])
    for e in es:
        # the fits vary on the scale of e; a relative step of 1e-4 keeps the
        # truncation error near 1e-8 and the cancellation of exp(x) - 1 at
        # small e out of the difference
        h = 1e-4 * e
        if np.any(np.abs(S95P_BREAKS - e) < 2 * h):
            continue
        d = np.zeros(6)
        L.ah_eos_d(1, e, P(d))
        fp, fm = np.zeros(3), np.zeros(3)
        L.ah_eos(1, e + h, P(fp))
        L.ah_eos(1, e - h, P(fm))
        fd = (fp - fm) / (2 * h)
        for got, want in zip(d[[1, 3, 5]], fd):
            assert abs(got - want) <= 1e-6 * abs(want) + 1e-12, (e, d, fd)


def test_table_interpolant_derivative():
    """the derivative of the table's cubic interpolant (what the fast-mode
    adjoint differentiates) against central differences of the interpolant,
    at interval midpoints over the table's range"""
    L = lib()
    rng = np.random.default_rng(51)
    K = 1024
    done = 0
    for _ in range(600):
        E = rng.integers(-13, 13)
        q = rng.integers(0, K)
        e = 2.0 ** E * (1.0 + (q + 0.5) / K)
        out, dout = np.zeros(4), np.zeros(4)
        if not L.ah_tab_d(e, P(out), P(dout)):
            continue
        h = 2.0 ** E / K * 1e-2  # stays inside the interval, where the interpolant is one cubic
        fp, fm = np.zeros(4), np.zeros(4)
        assert L.ah_tab(e + h, P(fp)) and L.ah_tab(e - h, P(fm))
        fd = (fp - fm) / (2 * h)
        assert np.all(np.abs(dout - fd) <= 1e-6 * np.abs(fd) + 1e-9 * np.abs(out) / e), (e, dout, fd)
        # and it follows the closed-form derivatives it interpolates
        d = np.zeros(6)
        L.ah_eos_d(1, e, P(d))
        assert abs(dout[0] - d[1]) <= 1e-6 * abs(d[1]) and abs(dout[2] - d[3]) <= 1e-6 * abs(d[3])
        done += 1
    assert done > 400


@pytest.mark.parametrize("eos,laws", [(0, (0, 0.2, 0, 5.0)), (0, (1, 1.0, 1, 1.5)), (0, (0, 0.08, 2, 0.4)),
                                      (1, (1, 1.0, 1, 1.5)), (1, (0, 0.08, 0, 5.0))])
def test_source_vjp(eos, laws):
    L = lib()
    el, ec, tl, tc = laws
    rng = np.random.default_rng(30 + el + tl)
    tau = 0.9
    for _ in range(5):
        u = rand_u(rng, tau)
        e = rng.uniform(1, 8)
        x = np.concatenate([[e, e / 3.0 * (1 + 0.05 * rng.standard_normal()), 0.0], u, 0.05 * e * rng.standard_normal(10),
                            rng.standard_normal(16), [ec, tc]])

        def f(z):
            S = np.zeros(15)
            L.ah_source(eos, el, z[33], tl, z[34], tau, z[0], z[1], z[2], P(z[3:7]), P(z[7:17]), P(z[17:33]), P(S))
            return S
        ob = rng.standard_normal(15)
        ob[[1, 2, 14]] = 0.0
        inb = np.zeros(35)
        L.ah_source_vjp(eos, el, ec, tl, tc, tau, x[0], x[1], x[2], P(x[3:7]), P(x[7:17]), P(x[17:33]), P(ob),
                        P(inb))
        dot_check(f, x, inb, ob, rng)


@pytest.mark.parametrize("eos", [0, 1])
def test_update_vjp_root_by_implicit_function_theorem(eos):
    L = lib()
    rng = np.random.default_rng(40 + eos)
    tn = 1.2
    for _ in range(8):
        e = rng.uniform(1, 10) if eos == 0 else np.exp(rng.uniform(np.log(0.05), np.log(300.0)))
        u = rand_u(rng, tn)
        pi = 0.02 * e * rng.standard_normal(10)
        ev = np.zeros(3)
        L.ah_eos(eos, e, P(ev))
        Pr = ev[0]
        w = e + Pr
        T = w * u[0] * u - Pr * np.array([1.0, 0, 0, 0]) + pi[[0, 1, 2, 3]]
        V = np.concatenate([T * tn, pi, [0.0]])

        def f(z):
            out = np.zeros(25)
            L.ah_update(eos, tn, P(z), e, P(out))
            return out
        out = f(V)
        assert abs(out[0] - e) < 1e-9 * e
        ob = rng.standard_normal(25)
        Vb = np.zeros(15)
        L.ah_update_vjp(eos, tn, P(V), out[0], P(ob), P(Vb))
        dot_check(f, V, Vb, ob, rng, h=1e-7, tol=1e-5)
