"""Pins of the oracle's scalar building blocks: deterministic exp/log/pow
against libm, the conformal x^(1/4), and WENO5-JS against closed forms and
its convergence order (NEWEST10/utils.h:57-160)."""
import math

import numpy as np

import oracle_ctypes

L = oracle_ctypes.lib()


def ulps(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a.view(np.int64) - b.view(np.int64))


def test_vm_exp_matches_libm():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-50, 50, 20000), rng.uniform(-1, 1, 5000), [0.0, 1.0, -1.0, 700.0, -700.0]])
    got = np.array([L.or_vm_exp(v) for v in x])
    assert np.max(ulps(got, np.exp(x))) <= 2


def test_vm_log_matches_libm():
    rng = np.random.default_rng(1)
    x = np.concatenate([10 ** rng.uniform(-12, 12, 20000), rng.uniform(0.5, 2.0, 5000), [1.0, 2.0, 0.5]])
    got = np.array([L.or_vm_log(v) for v in x])
    ref = np.log(x)
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
    # near x = 1 log is small; compare on an absolute scale there
    near1 = np.abs(ref) < 1e-3
    assert np.max(err[~near1]) < 4e-16
    assert np.max(np.abs(got - ref)[near1]) < 1e-18


def test_vm_pow_accuracy_over_eos_ranges():
    # exponents and bases of NEWEST10/s95p.h (e in GeV/fm^3 from 1e-6 to 200)
    exps = [0.4585, -1.233, 0.4585 - 1, -1.233 - 1, 1.16849, 1.15635, 1.16187, 4.54886, -0.128012, 1.18735,
            -1.18423, 0.75, 0.30679]
    rng = np.random.default_rng(2)
    xs = 10 ** rng.uniform(-6, np.log10(200.0), 3000)
    worst = 0.0
    for y in exps:
        got = np.array([L.or_vm_pow(v, y) for v in xs])
        ref = np.power(xs, y)
        worst = max(worst, float(np.max(np.abs(got - ref) / ref)))
    assert worst < 2e-14


def test_conformal_quarter_power_within_one_ulp():
    rng = np.random.default_rng(3)
    x = 10 ** rng.uniform(-10, 6, 50000)
    a = np.sqrt(np.sqrt(x))
    b = np.power(x, 0.25)
    assert np.max(ulps(a, b)) <= 1


def weno_L(*q):
    return L.or_weno_L(*[float(v) for v in q])


def weno_R(*q):
    return L.or_weno_R(*[float(v) for v in q])


def test_weno_constant_and_linear_data_are_exact():
    # the weights d_c/sum round to 0.3, 0.6, 0.1, whose double sum is 1 - 2^-53
    for c in (0.0, 1.0, -3.25, 1e-8):
        assert abs(weno_L(c, c, c, c, c) - c) <= 2.3e-16 * abs(c)
        assert abs(weno_R(c, c, c, c, c) - c) <= 2.3e-16 * abs(c)
    # linear data q_i = a + b i: every candidate stencil is exact, L is the
    # value at i + 1/2 and R at i - 1/2 of the centre cell
    a, b = 0.7, 0.3
    q = [a + b * i for i in range(-2, 3)]
    assert abs(weno_L(*q) - (a + 0.5 * b)) < 1e-15
    assert abs(weno_R(*q) - (a - 0.5 * b)) < 1e-15
    assert abs(L.or_weno_der(*q, 0.1) - b / 0.1) < 1e-13


def test_weno_mirror_symmetry():
    rng = np.random.default_rng(4)
    for _ in range(2000):
        q = rng.standard_normal(5)
        l = weno_L(*q)
        r = weno_R(*q[::-1])
        assert abs(l - r) <= 1e-14 * max(1.0, abs(l))


def test_weno_fifth_order_flux_difference():
    """Finite-difference WENO: (L_i - L_{i-1})/dx -> f'(x_i) at fifth order
    on smooth data (point values, as the reference feeds it)."""
    errs = []
    for n in (20, 40, 80):
        dx = 1.0 / n
        x = np.arange(-3, n + 3) * dx
        f = np.sin(2 * np.pi * x) + 2.0
        Ls = np.array([weno_L(*f[i - 2:i + 3]) for i in range(2, len(f) - 2)])
        xi = x[2:-2]
        d = (Ls[1:] - Ls[:-1]) / dx
        # (L_i - L_{i-1})/dx approximates f' at x_i
        ref = 2 * np.pi * np.cos(2 * np.pi * xi[1:])
        errs.append(np.max(np.abs(d - ref)[3:-3]))
    order = math.log(errs[1] / errs[2], 2)
    assert order > 4.0, errs


def test_weno_shock_no_new_extrema():
    # a step: the reconstructions stay within the data range (ENO property)
    q = [0.0, 0.0, 0.0, 1.0, 1.0]
    for s in range(5):
        qq = q[s:] + [1.0] * s
        l, r = weno_L(*qq), weno_R(*qq)
        assert -1e-12 <= l <= 1 + 1e-12 and -1e-12 <= r <= 1 + 1e-12
