"""The viscous Gubser solution (conformal Israel-Stewart, eta/s = 0.2,
tau_pi = 5 eta/(e+p): the reference's GUBSER profile, NEWEST10/maindef.h:
295-330) against the step, on the CPU oracle and on the GPU.

The initial state at tau = 1 is the semi-analytic solution (tests/
gubser_viscous.py); after the evolution the central region (away from the
boundary, whose influence travels at most at the speed of light) must match
the solution at the final time to discretisation accuracy."""
import numpy as np
import pytest

from viral_amd.desc import make_desc, Q_DEFAULT, Q_GUBSER_MIRROR, OK
from gubser_viscous import ViscousGubser, central_errors
import oracle_ctypes


def gubser_desc(n, dx, fast=0):
    return make_desc(n, n, 1, lbi=1, dx=dx, dy=dx, rk=2, eta_coef=0.2, tpi_coef=5.0,
                     quirks=Q_DEFAULT | Q_GUBSER_MIRROR, fast_math=fast)


def test_ode_solution_limits():
    """ideal limit T_hat ~ 1/cosh(rho)^(2/3) (Gubser 2010), and the
    Navier-Stokes limit pbar -> (4/3)(eta/s) tanh(rho)/T_hat once the shear
    has relaxed (tau_pi small)"""
    rho = np.linspace(-2, 1, 7)
    g = ViscousGubser(eta_s=1e-7, method="Radau")
    That, _ = g.hat(rho)
    ideal = That[3] * (np.cosh(rho[3]) / np.cosh(rho)) ** (2.0 / 3.0)
    assert np.allclose(That, ideal, rtol=1e-5)
    g = ViscousGubser(eta_s=1e-3, method="Radau")
    That, pbar = g.hat(rho)
    ns = (4.0 / 3.0) * 1e-3 * np.tanh(rho) / That
    assert np.allclose(pbar, ns, rtol=2e-2, atol=1e-5)


def run_oracle(n, dx, ts, t_end, g):
    d = gubser_desc(n, dx)
    o = oracle_ctypes.Oracle(d)
    o.set_state(g.state(d, 1.0, ts))
    t = 1.0
    for _ in range(int(round((t_end - 1.0) / ts))):
        assert o.step(t, ts) == OK
        t = t + ts
    return central_errors(d, o.state(), g.fields(d, t), t, rmax=1.0)


def test_viscous_gubser_oracle_converges():
    """From tau = 1 to 1.2: the error shrinks with the grid and the step (the
    time derivative of u in the shear source is a one-sided difference, so the
    scheme is first order in dtau there)."""
    g = ViscousGubser()
    coarse = run_oracle(50, 0.1, 0.005, 1.2, g)
    fine = run_oracle(100, 0.05, 0.0025, 1.2, g)
    print("oracle errors (e, pi): coarse %s fine %s" % (coarse, fine))
    assert coarse[0] < 1.5e-3 and coarse[1] < 5e-4
    assert fine[0] < 2.5e-4 and fine[1] < 1.5e-4
    assert fine[0] < coarse[0] / 4 and fine[1] < coarse[1] / 2


def run_gpu(n, dx, ts, t_end, g, fast=1):
    from viral_amd import capi
    d = gubser_desc(n, dx, fast=fast)
    nsteps = int(round((t_end - 1.0) / ts))
    print("case viscous gubser gpu n=%d dx=%g ts=%g steps=%d" % (n, dx, ts, nsteps), flush=True)
    h = capi.Hydro(d)
    h.upload(g.state(d, 1.0, ts))
    r, done = h.steps(1.0, ts, nsteps)
    assert (r, done) == (OK, nsteps)
    t = 1.0
    for _ in range(nsteps):
        t = t + ts
    return central_errors(d, h.download(), g.fields(d, t), t, rmax=1.5)


@pytest.mark.gpu
def test_viscous_gubser_gpu_reference_profile():
    """The reference's GUBSER grid (100x100 LBI, dx = 0.05, dtau = 0.005) from
    tau = 1 to 1.5, then the same region on a grid and step halved twice."""
    g = ViscousGubser()
    ref = run_gpu(100, 0.05, 0.005, 1.5, g)
    fine = run_gpu(400, 0.0125, 0.00125, 1.5, g)
    print("gpu errors (e, pi): reference profile %s refined %s" % (ref, fine))
    assert ref[0] < 1e-3 and ref[1] < 1e-3
    assert fine[0] < ref[0] / 3 and fine[1] < ref[1] / 3
