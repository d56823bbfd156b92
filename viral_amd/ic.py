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


def cart_pi10(u, p):
    """Full Cartesian pi^{mu nu} (slots tt,tx,ty,tz,xx,xy,xz,yy,yz,zz) from the
    five evolved components p = (xx, yy, xy, xz, yz) by u_mu pi^{mu nu} = 0
    and pi^mu_mu = 0 (the A1..A5 forms of NEWEST5/maindef.h:497-501 at
    tau = 1, which the Cartesian TXYZ5 code uses)."""
    u0, u1, u2, u3 = u
    p1, p2, p3, p4, p5 = p
    zz = ((p1 * u1 * u1 + 2.0 * p3 * u1 * u2 + p2 * u2 * u2 + 2.0 * u3 * (p4 * u1 + p5 * u2) - u0 * u0 * (p1 + p2))
          / (u0 * u0 - u3 * u3))
    tx = (p1 * u1 + p3 * u2 + p4 * u3) / u0
    ty = (p3 * u1 + p2 * u2 + p5 * u3) / u0
    tz = (p4 * u1 + p5 * u2 + zz * u3) / u0
    return [p1 + p2 + zz, tx, ty, tz, p1, p3, p4, p2, p5, zz]


def cartesian_state(desc, e, P, u, p5, prevu, Temp=None, PI=None):
    """29-field Cartesian state: T^{t mu} from the full pi, the five evolved
    components in pi slots 0..4, slots 5..9 zero."""
    full = cart_pi10(u, p5)
    st = assemble(desc, e, P, u, full, prevu, 1.0, Temp=Temp, PI=PI)
    for l in range(5):
        st[F_PI0 + l] = np.asarray(p5[l]).ravel()
    for l in range(5, 10):
        st[F_PI0 + l] = 0.0
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


def random_smooth(desc, seed, tau, e_scale=4.0, v_scale=0.3, pi_frac=0.05, prevu_jitter=0.01, eos_p=None,
                  bulk_frac=0.0):
    """A seeded smooth random state with non-zero shear stress, used by the
    parity tests (values of the kind the step sees: e of a few fm^-4,
    |v| below ~0.5, |pi| a few percent of e; bulk_frac > 0 adds a bulk
    pressure PI of that fraction of e)."""
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
    PI = None
    if bulk_frac:
        PI = bulk_frac * e * (np.sin(0.19 * i + ph[6]) * np.cos(0.23 * j + ph[7]) + 0.1 * rng.standard_normal(e.shape))
    if desc.coords != MILNE:
        return cartesian_state(desc, e, P, u, pi[:5], prevu, PI=PI)
    for l in (3, 6, 8):
        pi[l] = pi[l] / t3
    pi[9] = pi[9] / (t3 * t3)
    return assemble(desc, e, P, u, pi, prevu, tau, PI=PI)


def _thickness(xs, ys, R=6.62, a=0.546, nz=801, zmax=20.0):
    """Woods-Saxon thickness T(x,y) = int rho dz with rho0 = 1 (shape only)."""
    z = np.linspace(-zmax, zmax, nz)
    dz = z[1] - z[0]
    r = np.sqrt(xs[..., None] ** 2 + ys[..., None] ** 2 + z ** 2)
    rho = 1.0 / (1.0 + np.exp((r - R) / a))
    return rho.sum(axis=-1) * dz


def eta_profile(eta, plateau=1.5, falloff=1.0):
    """Flat plateau of half-width `plateau` with Gaussian tails of width
    `falloff` (the form of NEWEST10/ginit.h:876-879)."""
    over = np.maximum(np.abs(eta) - plateau, 0.0)
    return np.exp(-over * over / (2.0 * falloff * falloff))


def glauber(desc, tau, ts, b=7.0, e_peak_gev=30.0, R=6.62, a=0.546, plateau=1.5, falloff=1.0, eos_p=None):
    """Smooth optical-Glauber event: e(x,y,eta) = N T_A(x+b/2,y) T_B(x-b/2,y)
    H(eta), normalised to peak e_peak_gev GeV/fm^3 (converted to fm^-4 with
    GEVFM), fluid at rest, pi = 0, prevu = u.  The reference builds its
    Glauber state in ginit.h (GSL integration); this is the same shape of
    input, generated here."""
    X, Y, E = coords(desc)
    x1, y1 = X[:, :, 0], Y[:, :, 0]
    TA = _thickness(x1 + b / 2.0, y1, R, a)
    TB = _thickness(x1 - b / 2.0, y1, R, a)
    prof = TA * TB
    prof = prof / prof.max()
    e = (e_peak_gev / GEVFM) * prof[:, :, None] * eta_profile(E[0, 0, :], plateau, falloff)[None, None, :]
    e = np.maximum(e, 1e-7 * e.max())
    P = e / 3.0 if eos_p is None else eos_p(e)
    z = np.zeros_like(e)
    o = np.ones_like(e)
    u = [o, z, z, z]
    prevu = [o.copy(), z.copy(), z.copy(), z.copy()]
    pi = [np.zeros_like(e) for _ in range(10)]
    return assemble(desc, e, P, u, pi, prevu, tau)


def hotspots(desc, tau, ts, seed, n_spots=200, width=0.4, b=None, b_max=10.0, R=6.62, a=0.546, gamma_k=1.0,
             e_scale_gev=10.0, plateau=1.5, falloff=1.0, eos_p=None, e_floor=1e-7):
    """Fluctuating event: n_spots Gaussian hot spots (width fm) whose centres
    are drawn from two Woods-Saxon nuclei at +-b/2 (b ~ U(0, b_max) unless
    given) with Gamma(k)-distributed weights, times the plateau eta profile,
    floored at e_floor times the maximum.  Seeded; fluid at rest, pi = 0."""
    rng = np.random.default_rng(seed)
    if b is None:
        b = rng.uniform(0.0, b_max)
    # sample nucleon-like positions from the Woods-Saxon density by rejection
    pts = []
    while len(pts) < n_spots:
        c = rng.uniform(-R - 4 * a, R + 4 * a, size=(4 * n_spots, 3))
        r = np.sqrt((c ** 2).sum(axis=1))
        keep = rng.uniform(0, 1, len(r)) < 1.0 / (1.0 + np.exp((r - R) / a))
        for p in c[keep]:
            if len(pts) < n_spots:
                pts.append(p)
    pts = np.array(pts)
    side = np.where(np.arange(n_spots) % 2 == 0, 0.5 * b, -0.5 * b)
    cx, cy = pts[:, 0] + side, pts[:, 1]
    w = rng.gamma(gamma_k, 1.0 / gamma_k, n_spots)
    X, Y, E = coords(desc)
    x1, y1 = X[:, :, 0], Y[:, :, 0]
    tr = np.zeros_like(x1)
    for q in range(n_spots):
        tr += w[q] * np.exp(-((x1 - cx[q]) ** 2 + (y1 - cy[q]) ** 2) / (2 * width * width))
    tr = tr / max(tr.max(), 1e-300)
    e = (e_scale_gev / GEVFM) * tr[:, :, None] * eta_profile(E[0, 0, :], plateau, falloff)[None, None, :]
    e = np.maximum(e, e_floor * e.max())
    P = e / 3.0 if eos_p is None else eos_p(e)
    z = np.zeros_like(e)
    o = np.ones_like(e)
    u = [o, z, z, z]
    prevu = [o.copy(), z.copy(), z.copy(), z.copy()]
    pi = [np.zeros_like(e) for _ in range(10)]
    return assemble(desc, e, P, u, pi, prevu, tau)
