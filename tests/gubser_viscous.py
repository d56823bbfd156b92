"""Semi-analytic viscous Gubser flow for conformal Israel-Stewart hydro.

In the Gubser frame (de Sitter coordinates rho, theta; Weyl-rescaled
T_hat = tau T) the flow is static and the reference's shear equation
(u.d pi = -(pi - 2 eta sigma)/tau_pi - 4/3 pi theta - ..., eta = (eta/s) s,
tau_pi = 5 eta/(e+p)) reduces to two ODEs for T_hat and
pbar = pi_hat^eta_eta / (e_hat + p_hat):

    dT_hat/drho = T_hat tanh(rho) (pbar/3 - 2/3)
    dpbar/drho  = -pbar T_hat / (5 eta/s) + tanh(rho) (4/15 - 4/3 pbar^2)

(Gubser & Yarom 2011; Marrochio et al., Phys. Rev. C 91, 014903).  Back in
Milne coordinates: sinh rho = -(1 - q^2 tau^2 + q^2 r^2)/(2 q tau),
T = T_hat/tau, u = (cosh k, sinh k r_hat, 0), tanh k = 2 q^2 tau r /
(1 + q^2 tau^2 + q^2 r^2), and with P_eta = pbar (e + P) the local-rest-frame
shear eigenvalue along eta,
    pi^{mu nu} = 3/2 P_eta e_eta e_eta + 1/2 P_eta (g^{mu nu} - u^mu u^nu),
e_eta = (0, 0, 0, 1/tau).  Test infrastructure.
"""
import math

import numpy as np
from scipy.integrate import solve_ivp

from viral_amd import ic
from viral_amd.desc import BORDER


class ViscousGubser:
    def __init__(self, eta_s=0.2, q=1.0, That_start=0.1, pbar_start=0.0, rho_start=-4.5, rho_max=1.5,
                 method="DOP853"):
        """The solution that starts at rho_start (far in the past of the grid's
        region) with T_hat = That_start and pbar = pbar_start, integrated
        forward (the relaxation equation is stable only forward in rho)."""
        self.eta_s, self.q = eta_s, q

        def rhs(rho, y):
            T, p = y
            th = math.tanh(rho)
            return [T * th * (p / 3.0 - 2.0 / 3.0), -p * T / (5.0 * eta_s) + th * (4.0 / 15.0 - 4.0 / 3.0 * p * p)]
        self.sol = solve_ivp(rhs, (rho_start, rho_max), [That_start, pbar_start], rtol=1e-12, atol=1e-14,
                             dense_output=True, method=method)
        assert self.sol.status == 0
        self.rho_start = rho_start

    def hat(self, rho):
        rho = np.asarray(rho, dtype=float)
        assert np.all(rho >= self.rho_start)
        return self.sol.sol(rho.ravel()).reshape((2,) + rho.shape)

    def rho(self, tau, r):
        q = self.q
        return -np.arcsinh((1 - q * q * tau * tau + q * q * r * r) / (2 * q * tau))

    def flow(self, tau, X, Y):
        q = self.q
        r = np.sqrt(X * X + Y * Y)
        kap = np.arctanh((2 * q * q * tau * r) / (1 + q * q * tau * tau + q * q * r * r))
        safe = np.where(r > 1e-12, r, 1.0)
        ux = np.where(r > 1e-12, X * np.sinh(kap) / safe, 0.0)
        uy = np.where(r > 1e-12, Y * np.sinh(kap) / safe, 0.0)
        return np.cosh(kap), ux, uy

    def fields(self, desc, tau):
        """(e, P, u[4], pi[10]) on the desc's ghosted grid at tau"""
        X, Y, _ = ic.coords(desc)
        r = np.sqrt(X * X + Y * Y)
        That, pbar = self.hat(self.rho(tau, r))
        T = That / tau
        factor = 3.0 * desc.dof * math.pi * math.pi / 90.0
        e = factor * T ** 4
        P = e / 3.0
        Peta = pbar * (e + P)
        u0, ux, uy = self.flow(tau, X, Y)
        z = np.zeros_like(e)
        u = [u0, ux, uy, z]
        h = 0.5 * Peta
        pi = [h * (1 - u0 * u0), -h * u0 * ux, -h * u0 * uy, z.copy(),
              h * (-1 - ux * ux), -h * ux * uy, z.copy(), h * (-1 - uy * uy), z.copy(), Peta / (tau * tau)]
        return e, P, u, pi

    def state(self, desc, tau, ts):
        e, P, u, pi = self.fields(desc, tau)
        X, Y, _ = ic.coords(desc)
        p0, px, py = self.flow(tau - ts, X, Y)
        prevu = [p0, px, py, np.zeros_like(e)]
        return ic.assemble(desc, e, P, u, pi, prevu, tau, Temp=ic.conformal_T(desc, e))


def central_errors(desc, st, ref, tau, rmax):
    """max relative error of e and of the shear eigenvalue tau^2 pi^{eta eta}
    (relative to e + P) over interior cells with r < rmax"""
    from viral_amd.desc import F_EN, F_PI0
    e_ref, P_ref, u_ref, pi_ref = ref
    X, Y, _ = ic.coords(desc)
    r = np.sqrt(X * X + Y * Y).ravel()
    m = np.zeros(desc.ghost_dims(), bool)
    m[desc.interior_slices()] = True
    m = m.ravel() & (r < rmax)
    e_ref = e_ref.ravel()
    h = (e_ref + P_ref.ravel())
    err_e = np.max(np.abs(st[F_EN][m] - e_ref[m]) / e_ref[m])
    err_pi = np.max(np.abs(st[F_PI0 + 9][m] - pi_ref[9].ravel()[m]) * tau * tau / h[m])
    return err_e, err_pi
