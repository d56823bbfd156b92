"""Host-side initial states on the ghosted grid (reference cell order).

These stand in for the reference's init profiles (NEWEST10/init.h:150-542,
ginit.h): the step under test starts from whatever state the caller uploads,
so the generators here only have to produce physical, seeded states of the
shapes the benchmark configurations name.
"""
import math

import numpy as np

from .desc import (BORDER, NFIELDS, F_T00, F_PI0, F_BULK, F_EN, F_P, F_TEMP, F_VX, F_U0, F_PU0,
                   EOS_CONFORMAL, MILNE, GEVFM)


def coords(desc, x0=None, y0=None, eta0=None):
    """Cell-centre coordinates of the ghosted grid, NEWEST10/init.h:23-51,557-567:
    x runs from -(N/2)*dx to (N/2-1)*dx over the interior (asymmetric for even
    N, q11), ghosts continue the spacing."""
    gx, gy, gz = desc.ghost_dims()
    if x0 is None:
        x0 = -(desc.nx // 2) * desc.dx
    if y0 is None:
        y0 = -(desc.ny // 2) * desc.dy
    if eta0 is None:
        eta0 = 0.0 if desc.lbi else -(desc.nz // 2) * desc.deta
    X = x0 + (np.arange(gx) - BORDER) * desc.dx
    Y = y0 + (np.arange(gy) - BORDER) * desc.dy
    if desc.lbi:
        E = np.array([eta0])
    else:
        E = eta0 + (np.arange(gz) - BORDER) * desc.deta
    return np.meshgrid(X, Y, E, indexing="ij")


def conformal_T(desc, e):
    factor = 3.0 * desc.dof * math.pi * math.pi / 90.0
    return np.sqrt(np.sqrt(e / factor))


def assemble(desc, e, P, u, pi, prevu, tau, Temp=None, PI=None):
    """Pack (e, P, u^mu, pi, prevu) into the 29-field state with
    T^{tau mu} = (e+P-PI) u^tau u^mu - (P-PI) g^{tau mu} + pi^{tau mu}
    (NEWEST10/init.h:537-540)."""
    n = desc.cells()
    st = np.zeros((NFIELDS, n))
    if PI is None:
        PI = np.zeros_like(e)
    w = e + P - PI
    st[F_T00] = (w * u[0] * u[0] - (P - PI) + pi[0]).ravel()
    for m in range(1, 4):
        st[F_T00 + m] = (w * u[0] * u[m] + pi[m]).ravel()
    for l in range(10):
        st[F_PI0 + l] = pi[l].ravel()
    st[F_BULK] = PI.ravel()
    st[F_EN] = e.ravel()
    st[F_P] = P.ravel()
    st[F_TEMP] = (Temp if Temp is not None else np.zeros_like(e)).ravel()
    for m in range(1, 4):
        st[F_VX + m - 1] = (u[m] / u[0]).ravel()
    for m in range(4):
        st[F_U0 + m] = u[m].ravel()
        st[F_PU0 + m] = prevu[m].ravel()
    return st


def gubser(desc, tau, ts, q=1.0, eps0=1.0):
    """Ideal Gubser flow (closed form of NEWideal/init.h:139-195, velocity of
    NEWEST10/init.h:470-521): e = eps0 (2q)^(8/3) / (tau (1+g))^(4/3),
    g = 2q^2(tau^2+r^2) + q^4(tau^2-r^2)^2, u^r = sinh(kappa),
    kappa = atanh(2 q^2 tau r / (1 + q^2 tau^2 + q^2 r^2)); prevu at tau-ts;
    pi = 0.  Conformal EOS."""
    assert desc.eos == EOS_CONFORMAL
    X, Y, _ = coords(desc)
    r = np.sqrt(X * X + Y * Y)
    shape = X.shape

    def flow(t):
        kap = np.arctanh((2 * q * q * t * r) / (1 + q * q * t * t + q * q * r * r))
        safe = np.where(r > 1e-7, r, 1.0)
        ux = np.where(r > 1e-7, X * np.sinh(kap) / safe, 0.0)
        uy = np.where(r > 1e-7, Y * np.sinh(kap) / safe, 0.0)
        return np.cosh(kap), ux, uy

    g = 2 * q * q * (tau * tau + r * r) + q ** 4 * (tau * tau - r * r) ** 2
    e = eps0 * (2 * q) ** (8.0 / 3.0) / (tau * (1 + g)) ** (4.0 / 3.0)
    P = e / 3.0
    u0, ux, uy = flow(tau)
    p0, px, py = flow(tau - ts)
    z = np.zeros(shape)
    u = [u0, ux, uy, z]
    prevu = [p0, px, py, z.copy()]
    pi = [np.zeros(shape) for _ in range(10)]
    return assemble(desc, e, P, u, pi, prevu, tau, Temp=conformal_T(desc, e))


def random_smooth(desc, seed, tau, e_scale=4.0, v_scale=0.3, pi_frac=0.05, prevu_jitter=0.01, eos_p=None):
    """A seeded smooth random state with non-zero shear stress, used by the
    parity tests (values of the kind the step sees: e of a few fm^-4,
    |v| below ~0.5, |pi| a few percent of e)."""
    rng = np.random.default_rng(seed)
    gx, gy, gz = desc.ghost_dims()
    i, j, k = np.meshgrid(np.arange(gx), np.arange(gy), np.arange(gz), indexing="ij")
    ph = rng.uniform(0, 2 * np.pi, 12)
    e = e_scale * (1.0 + 0.4 * np.sin(0.21 * i + ph[0]) * np.cos(0.17 * j + ph[1]) + 0.15 * np.sin(0.3 * k + ph[2]))
    e = e * (1.0 + 0.02 * rng.standard_normal(e.shape))
    if eos_p is None:
        P = e / 3.0
    else:
        P = np.vectorize(eos_p)(e)
    u1 = v_scale * np.sin(0.15 * i + ph[3]) + 0.02 * rng.standard_normal(e.shape)
    u2 = v_scale * np.cos(0.13 * j + ph[4]) + 0.02 * rng.standard_normal(e.shape)
    t3 = tau if desc.coords == MILNE else 1.0
    u3 = (0.1 * np.sin(0.25 * k + ph[5]) + 0.01 * rng.standard_normal(e.shape)) / t3
    u0 = np.sqrt(1 + u1 * u1 + u2 * u2 + t3 * t3 * u3 * u3)
    u = [u0, u1, u2, u3]
    prevu = [c * (1.0 + prevu_jitter * rng.standard_normal(e.shape)) for c in u]
    pi = [pi_frac * e * rng.standard_normal(e.shape) for _ in range(10)]
    for l in (3, 6, 8):
        pi[l] = pi[l] / t3
    pi[9] = pi[9] / (t3 * t3)
    return assemble(desc, e, P, u, pi, prevu, tau)
