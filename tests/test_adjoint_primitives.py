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
    L.ah_cart_pi10.argtypes = [dp, dp, dp]
    L.ah_cart_pi10_vjp.argtypes = [dp, dp, dp, dp]
    L.ah_cart_update.restype = C.c_int
    L.ah_cart_update.argtypes = [C.c_int, dp, dp, dp]
    L.ah_cart_update_vjp.argtypes = [C.c_int, dp, dp, dp, dp]
    L.ah_source_cart.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, C.c_double, dp, dp,
                                 dp, dp]
    L.ah_source_cart_vjp.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, C.c_double, dp,
                                     dp, dp, dp, dp]
    L.ah_bulk.restype = C.c_double
    L.ah_bulk.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double,
                          C.c_double, dp, C.c_double, dp]
    L.ah_bulk_vjp.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double,
                              C.c_double, dp, C.c_double, dp, C.c_double, dp]
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


# ---------------------------------------------------------------- Cartesian (TXYZ5)

def test_cart_pi10_vjp():
    """the full Cartesian pi^{mu nu} from u and the five evolved components"""
    L = lib()
    rng = np.random.default_rng(50)
    for _ in range(6):
        u = rand_u(rng, 1.0, milne=0)
        pe = 0.1 * rng.standard_normal(5)
        x = np.concatenate([u, pe])

        def f(z):
            out = np.zeros(10)
            L.ah_cart_pi10(P(z[:4]), P(z[4:]), P(out))
            return out
        # the tensor is traceless and transverse to u (g = diag(1, -1, -1, -1))
        pi = f(x)
        full = np.array([[pi[0], pi[1], pi[2], pi[3]], [pi[1], pi[4], pi[5], pi[6]], [pi[2], pi[5], pi[7], pi[8]],
                         [pi[3], pi[6], pi[8], pi[9]]])
        g = np.diag([1.0, -1.0, -1.0, -1.0])
        assert abs(np.trace(g @ full)) < 1e-12
        assert np.abs(full @ (g @ u)).max() < 1e-12
        ob = rng.standard_normal(10)
        inb = np.zeros(9)
        L.ah_cart_pi10_vjp(P(u), P(pe), P(ob), P(inb))
        dot_check(f, x, inb, ob, rng)


@pytest.mark.parametrize("eos", [0, 1])
def test_cart_update_vjp_root_by_implicit_function_theorem(eos):
    """the Cartesian 4-D gnewton root (e, u^x, u^y, u^z) from T^{t mu}, the
    five pi and PI, reversed by the implicit-function theorem: dot-product
    test against central differences of the converged root"""
    L = lib()
    rng = np.random.default_rng(60 + eos)
    for _ in range(8):
        e = rng.uniform(1, 10) if eos == 0 else np.exp(rng.uniform(np.log(0.05), np.log(300.0)))
        u = rand_u(rng, 1.0, milne=0)
        pe = 0.02 * e * rng.standard_normal(5)
        PI = 0.0
        ev = np.zeros(3)
        L.ah_eos(eos, e, P(ev))
        Pr = ev[0]
        pi = np.zeros(10)
        L.ah_cart_pi10(P(u), P(pe), P(pi))
        w = e + Pr - PI
        T = w * u[0] * u - (Pr - PI) * np.array([1.0, 0, 0, 0]) + pi[[0, 1, 2, 3]]
        prm = np.concatenate([T, pe, [PI]])
        guess = np.array([e * 1.01, *(u[1:] * 0.99)])

        def f(z):
            out = np.zeros(20)
            # flag 8: gnewton's test_delta(1e-30) asks for a zero step, so it
            # often runs to its 100 iterations between neighbouring doubles
            assert L.ah_cart_update(eos, P(z), P(guess), P(out)) & 4 == 0
            return out
        out = f(prm)
        assert abs(out[0] - e) < 1e-9 * e and np.abs(out[7:10] - u[1:]).max() < 1e-9
        ob = rng.standard_normal(20)
        Vb = np.zeros(15)
        x = np.array([out[0], *out[7:10]])
        L.ah_cart_update_vjp(eos, P(prm), P(x), P(ob), P(Vb))
        assert np.all(Vb[9:14] == 0.0)
        inb = np.concatenate([Vb[:9], [Vb[14]]])
        dot_check(f, prm, inb, ob, rng, h=1e-7, tol=1e-6)


@pytest.mark.parametrize("eos,laws", [(0, (0, 0.2, 0, 5.0)), (0, (1, 1.0, 1, 1.5)), (1, (1, 1.0, 1, 1.5)),
                                      (1, (0, 0.08, 0, 5.0))])
def test_source_cart_vjp(eos, laws):
    """the Cartesian shear source rows (the five evolved components) with the
    full tensor built from them"""
    L = lib()
    el, ec, tl, tc = laws
    rng = np.random.default_rng(70 + el + tl)
    for _ in range(5):
        u = rand_u(rng, 1.0, milne=0)
        e = rng.uniform(1, 8)
        x = np.concatenate([[e, e / 3.0 * (1 + 0.05 * rng.standard_normal())], u, 0.05 * e * rng.standard_normal(5),
                            rng.standard_normal(16), [ec, tc]])

        def f(z):
            S = np.zeros(5)
            L.ah_source_cart(eos, el, z[27], tl, z[28], z[0], z[1], P(z[2:6]), P(z[6:11]), P(z[11:27]), P(S))
            return S
        ob = rng.standard_normal(5)
        inb = np.zeros(29)
        L.ah_source_cart_vjp(eos, el, ec, tl, tc, x[0], x[1], P(x[2:6]), P(x[6:11]), P(x[11:27]), P(ob), P(inb))
        dot_check(f, x, inb, ob, rng)


# ---------------------------------------------------------------- bulk

@pytest.mark.parametrize("eos,laws,milne", [(0, (0, 1.0, 0, 0.1), 1), (0, (1, 0.5, 1, 1.5), 1), (1, (1, 0.5, 1, 1.5), 1),
                                            (1, (0, 0.3, 0, 0.2), 0), (0, (1, 0.5, 1, 1.5), 0)])
def test_bulk_vjp(eos, laws, milne):
    """the Israel-Stewart bulk source with zeta(s) and tau_Pi(zeta, P)"""
    L = lib()
    zl, zc, tl, tc = laws
    rng = np.random.default_rng(80 + zl + 2 * tl + 4 * milne)
    tau = 1.3 if milne else 1.0
    for _ in range(5):
        u = rand_u(rng, tau, milne)
        e = rng.uniform(1, 8)
        ev = np.zeros(3)
        L.ah_eos(eos, e, P(ev))
        x = np.concatenate([[e, ev[0] * (1 + 0.05 * rng.standard_normal()), -0.05 * e * rng.uniform(0.1, 1.0)], u,
                            rng.standard_normal(16)])

        def f(z):
            return np.array([L.ah_bulk(eos, zl, zc, tl, tc, milne, tau, z[0], z[1], P(z[3:7]), z[2], P(z[7:23]))])
        ob = rng.standard_normal(1)
        inb = np.zeros(23)
        L.ah_bulk_vjp(eos, zl, zc, tl, tc, milne, tau, x[0], x[1], P(x[3:7]), x[2], P(x[7:23]), ob[0], P(inb))
        dot_check(f, x, inb, ob, rng)
