"""CPU checks of the oracle's Cartesian (TXYZ5) pieces.

The reference's TXYZ5 tree does not compile (its flux and source code use the
derived shear components A1..A5 that only NEWEST5/maindef.h:497-501 defines,
and GSL is absent), so these are pinned by the mathematics:
  - the derived shear components satisfy u_mu pi^{mu nu} = 0, pi^mu_mu = 0
    and agree with an independent linear solve of those constraints;
  - the centred fluxes' T rows are T^{d mu} = w u^d u^mu - (P-PI) g^{d mu} +
    pi^{d mu} built independently;
  - the root's Jacobian is the derivative of its residuals (central
    differences), the residuals vanish at states assembled from known
    (e, u, pi), and the 4-D solve recovers them;
  - GSL's stopping rule (test_delta with relative 1e-30) ends a fluid at rest
    after exactly two iterations at the exact energy; a zero pivot raises the
    solver flag;
  - the Israel-Stewart source equals an independent projector-based
    evaluation (Milne and Cartesian);
  - a uniform flow stays uniform through a Cartesian step.
"""
import numpy as np
import pytest

from viral_amd import ic
from viral_amd.desc import (make_desc, CARTESIAN, MILNE, F_T00, F_EN, F_P, F_U0, F_PI0, F_BULK, FLAG_SOLVER,
                            FLAG_MAXITER, NFIELDS, EOS_S95P)
import oracle_ctypes as O

G = np.diag([1.0, -1.0, -1.0, -1.0])
SLOT = np.array([[0, 1, 2, 3], [1, 4, 5, 6], [2, 5, 7, 8], [3, 6, 8, 9]])


def rand_u(rng, scale=0.6):
    s = rng.uniform(-scale, scale, 3)
    return np.array([np.sqrt(1 + s @ s), *s])


def full_tensor(pi10):
    return np.array([[pi10[SLOT[m, n]] for n in range(4)] for m in range(4)])


@pytest.mark.parametrize("seed", range(6))
def test_derived_components_transverse_traceless_and_unique(seed):
    rng = np.random.default_rng(seed)
    u = rand_u(rng)
    p5 = rng.standard_normal(5)
    pi = full_tensor(O.cart_pi10(u, p5))
    ul = G @ u
    assert np.max(np.abs(ul @ pi)) < 1e-13 * (1 + np.max(np.abs(pi)))
    assert abs(np.trace(G @ pi)) < 1e-13 * (1 + np.max(np.abs(pi)))
    # the five constraints fix tt, tx, ty, tz, zz given xx, yy, xy, xz, yz
    A = np.zeros((5, 5))
    b = np.zeros(5)
    unk = [(0, 0), (0, 1), (0, 2), (0, 3), (3, 3)]
    known = {(1, 1): p5[0], (2, 2): p5[1], (1, 2): p5[2], (1, 3): p5[3], (2, 3): p5[4]}

    def comp(m, n):
        return (min(m, n), max(m, n))
    for nu in range(4):
        for mu in range(4):
            key = comp(mu, nu)
            if key in known:
                b[nu] -= ul[mu] * known[key]
            else:
                A[nu, unk.index(key)] += ul[mu]
    for mu in range(4):
        key = (mu, mu)
        if key in known:
            b[4] -= G[mu, mu] * known[key]
        else:
            A[4, unk.index(key)] += G[mu, mu]
    sol = np.linalg.solve(A, b)
    for q, (m, n) in enumerate(unk):
        assert abs(pi[m, n] - sol[q]) < 1e-12 * (1 + abs(sol[q]))


def test_centre_flux_T_rows_are_the_stress_tensor():
    d = make_desc(6, 5, 4, coords=CARTESIAN)
    st = ic.random_smooth(d, seed=3, tau=1.0)
    o = O.Oracle(d)
    o.set_state(st)
    for dd in range(3):
        F = o.centre_flux(dd, 1.0)
        for x in (0, 37, 100, d.cells() - 1):
            u = st[F_U0:F_U0 + 4, x]
            pi = full_tensor(ic.cart_pi10(u, st[F_PI0:F_PI0 + 5, x]))
            w = st[F_EN, x] + st[F_P, x] - st[F_BULK, x]
            T = w * np.outer(u, u) - (st[F_P, x] - st[F_BULK, x]) * np.linalg.inv(G) + pi
            for mu in range(4):
                assert abs(F[mu, x] - T[dd + 1, mu]) < 1e-13 * (1 + abs(T[dd + 1, mu]))
            for l in range(5):
                assert F[4 + l, x] == pytest.approx(st[F_PI0 + l, x] * u[dd + 1] / u[0], rel=1e-15, abs=1e-300)
            assert np.all(F[9:14, x] == 0.0)


def stress_prm(desc, e, u, p5, PI=0.0):
    P = O.lib().or_eos_p(desc, e)
    pi = full_tensor(O.cart_pi10(u, p5))
    w = e + P - PI
    T0 = w * u[0] * u - (P - PI) * np.array([1.0, 0, 0, 0]) + pi[0]
    return np.concatenate([T0, p5, [PI]])


@pytest.mark.parametrize("eos", [0, EOS_S95P])
def test_jacobian_is_derivative_of_residuals(eos):
    d = make_desc(4, 4, 4, coords=CARTESIAN, eos=eos)
    rng = np.random.default_rng(7)
    for _ in range(5):
        u = rand_u(rng)
        prm = stress_prm(d, 3.0, u, 0.05 * rng.standard_normal(5))
        x = np.array([2.7, *(u[1:] * 1.1)])
        J = O.cart_df(d, prm, x)
        for c in range(4):
            h = 1e-6 * max(1.0, abs(x[c]))
            xp, xm = x.copy(), x.copy()
            xp[c] += h
            xm[c] -= h
            num = (O.cart_f(d, prm, xp) - O.cart_f(d, prm, xm)) / (2 * h)
            assert np.allclose(J[:, c], num, rtol=1e-6, atol=1e-7), (c, J[:, c], num)


@pytest.mark.parametrize("eos", [0, EOS_S95P])
def test_root_recovers_known_state(eos):
    d = make_desc(4, 4, 4, coords=CARTESIAN, eos=eos)
    rng = np.random.default_rng(11)
    for _ in range(10):
        e = rng.uniform(0.5, 20.0)
        u = rand_u(rng, 0.8)
        p5 = 0.03 * e * rng.standard_normal(5)
        prm = stress_prm(d, e, u, p5)
        assert np.max(np.abs(O.cart_f(d, prm, np.array([e, *u[1:]])))) < 1e-12 * e * u[0] ** 3
        guess = np.array([e * 1.2, *(u[1:] * 0.9)])
        r, it, fl = O.cart_gnewton(d, prm, guess)
        assert fl & FLAG_SOLVER == 0 and 1 <= it <= 100
        assert abs(r[0] - e) < 1e-12 * e
        assert np.max(np.abs(r[1:] - u[1:])) < 1e-12 * u[0]


def test_fluid_at_rest_stops_after_two_iterations_exactly():
    d = make_desc(4, 4, 4, coords=CARTESIAN)
    prm = np.array([5.0, 0, 0, 0, 0, 0, 0, 0, 0, 0])
    r, it, fl = O.cart_gnewton(d, prm, [7.5, 0.0, 0.0, 0.0])
    assert (r[0], it, fl) == (5.0, 2, 0)
    assert np.all(r[1:] == 0.0)


def test_zero_pivot_sets_solver_flag():
    # at rest with e + P - PI = 0 the Jacobian has two zero rows
    d = make_desc(4, 4, 4, coords=CARTESIAN)
    prm = np.array([1.0, 0, 0, 0, 0, 0, 0, 0, 0, 4.0 / 3.0])
    r, it, fl = O.cart_gnewton(d, prm, [1.0, 0.0, 0.0, 0.0])
    assert it == 1 and fl & FLAG_SOLVER


def projector_source(milne, tau, u, pi10, du, SH, tpi):
    """Israel-Stewart source by explicit projector contractions."""
    g = np.diag([1.0, -1.0, -1.0, -(tau * tau) if milne else -1.0])
    gi = np.linalg.inv(g)
    Gam = np.zeros((4, 4, 4))  # Gam[a, b, c] = Gamma^a_{bc}
    if milne:
        Gam[0, 3, 3] = tau
        Gam[3, 0, 3] = Gam[3, 3, 0] = 1.0 / tau
    pi = full_tensor(pi10)
    nab = du.T + np.einsum("abc,c->ba", Gam, u)      # nab[b, a] = nabla_b u^a
    theta = np.trace(nab)
    Du = u @ nab                                      # Du^a = u^b nabla_b u^a
    Dul = g @ Du
    Delta = gi - np.outer(u, u)                       # Delta^{mn}
    Dmix = np.eye(4) - np.outer(u, g @ u)             # Delta^m_a
    grad_up = gi @ nab                                # nabla^b u^a  [b, a]
    sym = 0.5 * (np.einsum("ma,nb,ab->mn", Dmix, Dmix, grad_up) + np.einsum("ma,nb,ba->mn", Dmix, Dmix, grad_up))
    trace = np.einsum("ab,ab->", g @ Delta @ g, grad_up)
    sigma = sym - Delta * trace / 3.0
    rhs = -(pi - 2 * SH * sigma) / tpi - 4.0 / 3.0 * pi * theta - (np.outer(u, pi @ Dul) + np.outer(pi @ Dul, u))
    chr = np.einsum("l,mla,an->mn", u, Gam, pi) + np.einsum("l,nla,ma->mn", u, Gam, pi)
    divv = sum(du[i, i] / u[0] - u[i] * du[0, i] / u[0] ** 2 for i in range(1, 4))
    S = (rhs - chr) / u[0] + pi * divv
    return np.array([S[m, n] for (m, n) in [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3),
                                            (3, 3)]])


@pytest.mark.parametrize("milne", [0, 1])
def test_shear_source_matches_projector_form(milne):
    rng = np.random.default_rng(21 + milne)
    for _ in range(8):
        tau = rng.uniform(0.6, 2.0) if milne else 1.0
        s = rng.uniform(-0.5, 0.5, 3)
        g33 = tau * tau if milne else 1.0
        u = np.array([np.sqrt(1 + s[0] ** 2 + s[1] ** 2 + g33 * s[2] ** 2), *s])
        pi10 = 0.1 * rng.standard_normal(10)
        du = rng.standard_normal((4, 4))
        SH, tpi = rng.uniform(0.1, 1.0), rng.uniform(0.2, 2.0)
        a = O.shear_source(milne, tau, u, pi10, du, SH, tpi)
        b = projector_source(milne, tau, u, pi10, du, SH, tpi)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.max(np.abs(b))), (a - b)


def test_uniform_flow_stays_uniform():
    d = make_desc(9, 8, 7, coords=CARTESIAN, rk=2)
    gx, gy, gz = d.ghost_dims()
    shape = (gx, gy, gz)
    e = np.full(shape, 4.0)
    u = [np.full(shape, v) for v in (np.sqrt(1 + 0.3 ** 2 + 0.2 ** 2 + 0.1 ** 2), 0.3, -0.2, 0.1)]
    p5 = [np.zeros(shape) for _ in range(5)]
    st = ic.cartesian_state(d, e, e / 3.0, u, p5, [c.copy() for c in u])
    o = O.Oracle(d)
    o.set_state(st)
    assert o.step(1.0, 0.01) == 0
    out = o.state()
    for f in (F_T00, F_T00 + 1, F_EN, F_U0 + 1):
        assert np.max(np.abs(out[f] - st[f])) <= 1e-12 * np.max(np.abs(st[f])), f


def test_q9_zero_shear_sources_in_top_slices():
    from viral_amd.desc import Q_DEFAULT, Q_SHEAR_SLICES
    d = make_desc(6, 5, 7, coords=CARTESIAN)
    st = ic.random_smooth(d, seed=5, tau=1.0)
    o = O.Oracle(d)
    o.set_state(st)
    S = o.calc_source(1.0, 0.01).reshape(15, *d.ghost_dims())
    inter = S[:, 3:-3, 3:-3, :]
    assert np.all(inter[4:9, :, :, 7:10] == 0.0)        # ghosted k = nz .. nz+2
    assert np.all(np.abs(inter[4:9, :, :, 3:7]) > 0.0)
    d.quirks = Q_DEFAULT & ~Q_SHEAR_SLICES
    o2 = O.Oracle(d)
    o2.set_state(st)
    S2 = o2.calc_source(1.0, 0.01).reshape(15, *d.ghost_dims())
    assert np.all(np.abs(S2[4:9, 3:-3, 3:-3, 7:10]) > 0.0)
    assert np.array_equal(S2[4:9, 3:-3, 3:-3, 3:7], inter[4:9, :, :, 3:7])
