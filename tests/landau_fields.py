"""T^{mu nu} fields for the Landau-matching tests, and an emulation of the
reference's eigenvector selection (NEWEST10/ginit.h:521-629, FindVar) on top
of LAPACK's nonsymmetric eigensolver (numpy.linalg.eig), the same
Francis-QR family GSL's gsl_eigen_nonsymmv uses.  Test infrastructure."""
import numpy as np

PIDX = [[0, 1, 2, 3], [1, 4, 5, 6], [2, 5, 7, 8], [3, 6, 8, 9]]
PAIRS = [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]


def metric(tau):
    return np.array([1.0, -1.0, -1.0, -tau * tau])


def ref_findvar(T10, tau):
    """(ok, E, v) as FindVar would return them: every eigenpair of
    M^a_b = T^{ab} g_bb with a real eigenvalue and a real eigenvector whose
    v = x^i/x^0 has |v_i| < 1 and |v| < 1 counts; exactly one must."""
    g = metric(tau)
    M = np.array([[T10[PIDX[a][b]] * g[b] for b in range(4)] for a in range(4)])
    w, V = np.linalg.eig(M)
    w = np.asarray(w, dtype=complex)
    V = np.asarray(V, dtype=complex)
    hits = []
    for i in range(4):
        if w[i].imag != 0:
            continue
        x = V[:, i]
        if x[0].real != 0 and not np.any(x.imag != 0):
            v = x[1:].real / x[0].real
            if np.all(np.abs(v) < 1) and np.sqrt(np.sum(v * v)) < 1:
                hits.append((abs(w[i].real), v))
    if len(hits) != 1:
        return False, 0.0, np.zeros(3)
    return True, hits[0][0], hits[0][1]


def fluid(e, P, v, pi10, tau):
    """T^{mu nu} = (e + P) u u - P g^{-1} + pi of one cell (pi10 in the pi
    index map), and u"""
    gi = 1.0 / metric(tau)
    u0 = 1.0 / np.sqrt(1.0 - v[0] ** 2 - v[1] ** 2 - tau * tau * v[2] ** 2)
    u = u0 * np.array([1.0, v[0], v[1], v[2]])
    T = np.array([(e + P) * u[a] * u[b] - (P * gi[a] if a == b else 0.0) + pi10[k] for k, (a, b) in enumerate(PAIRS)])
    return T, u


def transverse_traceless(u, tau, A):
    """pi^{mn} = Delta^m_a Delta^n_b A^{ab} - Delta^{mn} Delta_{ab} A^{ab} / 3
    of a symmetric 4x4 A, in the pi index map"""
    g = metric(tau)
    gi = 1.0 / g
    ul = g * u
    Dmix = np.eye(4) - np.outer(u, ul)           # Delta^m_a
    Dup = np.diag(gi) - np.outer(u, u)           # Delta^{mn}
    Ddn = np.diag(g) - np.outer(ul, ul)          # Delta_{ab}
    P = Dmix @ A @ Dmix.T
    pi = P - Dup * np.sum(Ddn * A) / 3.0
    return np.array([pi[a, b] for a, b in PAIRS])


def random_fluid_field(n, seed, tau, eos_p, vmax=0.6, pi_frac=0.05, e_range=(0.5, 40.0)):
    """n cells of random e, v and transverse traceless pi: (tmunu (10, n),
    e (n), v (3, n), pi (10, n))"""
    rng = np.random.default_rng(seed)
    e = np.exp(rng.uniform(np.log(e_range[0]), np.log(e_range[1]), n))
    P = eos_p(e)
    T = np.zeros((10, n))
    V = np.zeros((3, n))
    PI = np.zeros((10, n))
    for x in range(n):
        # velocities (coordinate components) inside the reference's test,
        # |v| < vmax, and timelike
        while True:
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            v = d * (vmax * rng.uniform() ** 0.5)
            if v[0] ** 2 + v[1] ** 2 + tau * tau * v[2] ** 2 < vmax * vmax:
                break
        _, u = fluid(e[x], P[x], v, np.zeros(10), tau)
        A = rng.normal(size=(4, 4))
        A = (A + A.T) * (pi_frac * e[x])
        pi = transverse_traceless(u, tau, A)
        T[:, x], _ = fluid(e[x], P[x], v, pi, tau)
        V[:, x] = v
        PI[:, x] = pi
    return T, e, V, PI


def ginit_like_field(n, seed, tau, eflat=(0.2, 30.0)):
    """T^{mu nu} of the reference's ginit construction (NEWEST10/ginit.h:
    876-907) with random local values of its inputs: energy density ee,
    gradient terms AX, AY, BX, BY, Le, DB and the rapidity cut-off factor"""
    rng = np.random.default_rng(seed)
    t2 = tau * tau
    T = np.zeros((10, n))
    for x in range(n):
        ee = np.exp(rng.uniform(np.log(eflat[0]), np.log(eflat[1])))
        dt = 100 * ee
        AX, AY, BX, BY = rng.normal(size=4) * ee * 0.3
        Le = rng.normal() * ee * 0.5
        DB = rng.normal() * ee * 0.5
        cut = rng.uniform(0.05, 1.0)
        t = np.zeros((4, 4))
        t[0, 0] = ee - 0.125 * t2 * (-2 * Le + dt)
        t[0, 1] = 0.5 * tau * AX
        t[0, 2] = 0.5 * tau * AY
        t[0, 3] = 0.125 * tau * DB
        t[1, 1] = ee - 0.25 * t2 * (-Le + dt)
        t[1, 2] = 0
        t[1, 3] = 0.5 * BX
        t[2, 2] = t[1, 1]
        t[2, 3] = 0.5 * BY
        t[3, 3] = -ee / t2 + 0.125 * (-2 * Le + 3 * dt)
        T[:, x] = [t[a, b] * cut for a, b in PAIRS]
    return T
