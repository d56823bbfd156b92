"""Per-cell relative error of two states on physical scales.

Fields that vanish identically for a symmetric flow (u^eta, pi^{tau eta}, ...
in boost-invariant Gubser flow) carry only rounding noise, so each quantity is
compared with the cell's own natural scale: e with e, u^mu with u^0,
pi^{mu nu} and PI with the enthalpy e+P, T^{tau mu} with T^{tau tau}, the
three-velocities with 1.
"""
import numpy as np

from viral_amd.desc import BORDER, F_T00, F_PI0, F_BULK, F_EN, F_P, F_TEMP, F_VX, F_U0, F_PU0


def interior_mask(desc):
    m = np.zeros(desc.ghost_dims(), bool)
    m[desc.interior_slices()] = True
    return m.ravel()


def state_errors(desc, a, b):
    """max over interior cells of each group's relative difference between
    states a and b (b is the reference)."""
    I = interior_mask(desc)
    e = np.abs(b[F_EN][I])
    h = np.abs(b[F_EN][I] + b[F_P][I])
    u0 = np.abs(b[F_U0][I])
    t0 = np.abs(b[F_T00][I])
    out = {}
    out["e"] = np.max(np.abs(a[F_EN][I] - b[F_EN][I]) / e)
    out["P"] = np.max(np.abs(a[F_P][I] - b[F_P][I]) / np.abs(b[F_P][I]))
    out["Temp"] = np.max(np.abs(a[F_TEMP][I] - b[F_TEMP][I]) / np.abs(b[F_TEMP][I]))
    out["u"] = max(np.max(np.abs(a[F_U0 + m][I] - b[F_U0 + m][I]) / u0) for m in range(4))
    out["prevu"] = max(np.max(np.abs(a[F_PU0 + m][I] - b[F_PU0 + m][I]) / np.abs(b[F_PU0][I])) for m in range(4))
    out["v"] = max(np.max(np.abs(a[F_VX + m][I] - b[F_VX + m][I])) for m in range(3))
    out["pi"] = max(np.max(np.abs(a[F_PI0 + l][I] - b[F_PI0 + l][I]) / h) for l in range(10))
    out["PI"] = np.max(np.abs(a[F_BULK][I] - b[F_BULK][I]) / h)
    out["T"] = max(np.max(np.abs(a[F_T00 + m][I] - b[F_T00 + m][I]) / t0) for m in range(4))
    return out


def worst(errs):
    return max(errs.values())
