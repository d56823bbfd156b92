"""torch.autograd binding of the hand-written adjoint (include/viral.h,
viral_adjoint_*): a recorded multi-step run as a differentiable function of
the initial state and of the viscosity / relaxation coefficients.

The forward runs on the device through the C ABI; backward calls
viral_adjoint, which recomputes the stage states between checkpoints (every
> 0) or reads the stage states the forward kept in HBM (every = 0), and
applies the adjoint kernels.  torch only carries the host-side arrays.
"""
import math

import numpy as np
import torch

from . import capi
from .desc import (NFIELDS, F_T00, F_PI0, F_BULK, F_EN, F_P, F_TEMP, F_VX, F_U0, F_PU0, BORDER, EOS_CONFORMAL)


class HydroRun(torch.autograd.Function):
    """state (NFIELDS, cells) float64, eta_coef, tpi_coef (0-d float64) ->
    final state after `nsteps` steps of dtau from tau0."""

    @staticmethod
    def forward(ctx, state, eta_coef, tpi_coef, desc, tau0, ts, nsteps, every):
        d = desc.copy()
        d.eta_coef = float(eta_coef)
        d.tpi_coef = float(tpi_coef)
        h = capi.Hydro(d)
        h.upload(state.detach().cpu().numpy())
        h.adjoint_begin(every)
        r, done = h.steps(tau0, ts, nsteps)
        if r != 0:
            raise capi.ViralError(r, "step %d failed" % done)
        ctx.hydro = h
        return torch.from_numpy(h.download())

    @staticmethod
    def backward(ctx, gout):
        grad, pg = ctx.hydro.adjoint(gout.detach().cpu().numpy())
        ctx.hydro.close()
        return (torch.from_numpy(grad), torch.tensor(pg[0], dtype=torch.float64),
                torch.tensor(pg[1], dtype=torch.float64), None, None, None, None, None)


def run(state, eta_coef, tpi_coef, desc, tau0, ts, nsteps, every=20):
    """every: checkpoint interval of the recording; 0 keeps every stage state
    (29 doubles per ghosted cell per RK stage) and skips the recompute."""
    return HydroRun.apply(state, eta_coef, tpi_coef, desc, tau0, ts, nsteps, every)


class EosPressure(torch.autograd.Function):
    """P(e) of a descriptor's EOS evaluated by the library on the host
    (viral_eos_eval: the closed forms; the descriptor's arithmetic mode does
    not matter, the initial state is built with the closed form), with dP/de
    from the same call for the backward pass."""

    @staticmethod
    def forward(ctx, e, desc):
        d = desc.copy()
        d.fast_math = d.eos_table = 0
        p, dpde, _, _ = capi.eos_eval(d, e.detach().cpu().numpy())
        ctx.save_for_backward(torch.from_numpy(dpde).reshape(e.shape))
        return torch.from_numpy(p).reshape(e.shape)

    @staticmethod
    def backward(ctx, gp):
        (dpde,) = ctx.saved_tensors
        return gp * dpde, None


def state_from_energy(desc, e, tau):
    """Differentiable 29-field state of a fluid at rest with zero shear stress
    and energy density e (torch, ghosted cells) under the descriptor's EOS:
    T^{tau tau} = (e + P) - P = e, T^{tau i} = 0 (NEWEST10/init.h:537-540),
    P = P(e) -- e/3 for the conformal EOS, the s95p closed form through
    EosPressure otherwise -- and Temp = T(e) (not read by the step, so it
    carries no gradient)."""
    n = desc.cells()
    e = e.reshape(n)
    if desc.eos == EOS_CONFORMAL:
        P = e / 3.0
        factor = 3.0 * desc.dof * math.pi * math.pi / 90.0
        temp = torch.sqrt(torch.sqrt(e / factor))
    else:
        P = EosPressure.apply(e, desc)
        d = desc.copy()
        d.fast_math = d.eos_table = 0
        temp = torch.from_numpy(capi.eos_eval(d, e.detach().cpu().numpy())[2])
    zero = torch.zeros_like(e)
    one = torch.ones_like(e)
    rows = [zero] * NFIELDS
    rows = list(rows)
    rows[F_T00] = (e + P) - P
    rows[F_EN] = e
    rows[F_P] = P
    rows[F_TEMP] = temp
    rows[F_U0] = one
    rows[F_PU0] = one
    return torch.stack(rows)


def interior_mask(desc):
    gx, gy, gz = desc.ghost_dims()
    m = np.zeros((gx, gy, gz), bool)
    m[desc.interior_slices()] = True
    return torch.from_numpy(m.ravel())


def momentum_anisotropy(desc, state, tau):
    """eps_2 = int (T^xx - T^yy) / int (T^xx + T^yy) over the interior, with
    T^{ij} = (e + P - PI) u^i u^j - (P - PI) g^{ij} + pi^{ij}, g^{xx} = g^{yy} = -1."""
    m = interior_mask(desc)
    w = state[F_EN] + state[F_P] - state[F_BULK]
    pp = state[F_P] - state[F_BULK]
    txx = w * state[F_U0 + 1] ** 2 + pp + state[F_PI0 + 4]
    tyy = w * state[F_U0 + 2] ** 2 + pp + state[F_PI0 + 7]
    return (txx - tyy)[m].sum() / (txx + tyy)[m].sum()
