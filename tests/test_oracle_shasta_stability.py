"""Why the reference's SHASTA step breaks down on fluctuating 3-D states.

NEWEST10/hydroshasta.h:101-284 gives each direction the rate
(-U + UTD - A_i + A_{i-1}) / taustep, where the transported-diffused UTD
carries a 1/8 diffusion (UTD = U + (U_{+1} + U_{-1} - 2U)/8 for a fluid at
rest) and the antidiffusion A is limited to zero at extrema; TVDRK2 and
TVDRK3 call it with taustep = dtau/2 and dtau/3 (NEWEST10/time.h:352,411).
For the 3-D checkerboard mode every cell is an extremum, so each direction
contributes -(1/2)(dtau/taustep) of the mode to dtau L: -3 per RK2 stage,
-4.5 per RK3 stage.  Through the RK combinations
(NEWEST10/time.h:54-98) the mode's amplitude per step is
    RK2: 1/2 + (1/2)(-2 + 6) = 2.5,
    RK3: 1/3 + (2/3)(U2 - 4.5 U2), U1 = -3.5, U2 = 3/4 + (1/4)(U1 - 4.5 U1)
         = 3.8125 -> -8.5625,
so grid-scale rounding grows without bound in 3-D (in 2-D, boost-invariant
grids, RK2 is marginal: 1).  With the full step as sub-step
(VIRAL_Q_SHASTA_FULLSTEP) the per-stage factor is -1.5 and the mode decays:
RK2 0.625, RK3 0.0625.

The oracle, bitwise equal to the reference's hydroExplicit
(tests/test_oracle_vs_reference.py), must show exactly these factors on a
fluid at rest with a 1e-6 checkerboard, measured at the centre of the grid
(away from the zero-gradient ghosts), to 1 % (the mode's pressure gradients
and the Bjorken dilution add a few 1e-3), and the planted-mistake oracle
whose RK3 final weights are wrong must not.
"""
import numpy as np
import pytest

import oracle_ctypes
from viral_amd import ic
from viral_amd.desc import make_desc, SHASTA, Q_DEFAULT, Q_SHASTA_FULLSTEP, F_EN, OK

AMP = 1e-6
PREDICTED = {(2, 0): 2.5, (3, 0): -8.5625, (2, 1): 0.625, (3, 1): 0.0625, (1, 0): -0.5, (1, 1): -0.5}


def checkerboard_factor(rk, fullstep, lbi=0, L=None):
    n = 20
    d = make_desc(n, n, 1 if lbi else n, lbi=lbi, rk=rk, scheme=SHASTA, eta_coef=0.12,
                  quirks=Q_DEFAULT | (Q_SHASTA_FULLSTEP if fullstep else 0))
    G = d.ghost_dims()
    i, j, k = np.meshgrid(*[np.arange(g) for g in G], indexing="ij")
    pert = (-1.0) ** (i + j + k)
    e = 5e-6 * (1 + AMP * pert)
    u = np.zeros((4,) + e.shape)
    u[0] = 1.0
    st = ic.assemble(d, e, e / 3.0, u, np.zeros((10,) + e.shape), u.copy(), 0.6, Temp=ic.conformal_T(d, e))
    o = oracle_ctypes.Oracle(d, L)
    o.set_state(st)
    assert o.step(0.6, 0.002) == OK
    E = o.field(F_EN).reshape(G)
    c = (G[0] // 2, G[1] // 2, 0 if lbi else G[2] // 2)
    nb = (c[0] + 1, c[1], c[2])
    a = (E[c] - E[nb]) / (E[c] + E[nb])
    return a / (AMP * pert[c])


@pytest.mark.parametrize("rk,fullstep", sorted(PREDICTED))
def test_checkerboard_factor_per_step(rk, fullstep):
    f = checkerboard_factor(rk, fullstep)
    print("rk %d fullstep %d: factor %.6f (predicted %g)" % (rk, fullstep, f, PREDICTED[(rk, fullstep)]))
    # the checkerboard's pressure gradients and the Bjorken dilution move the
    # factor by a few 1e-3 from the linear diffusion count
    assert abs(f - PREDICTED[(rk, fullstep)]) <= 1e-2 * max(1.0, abs(PREDICTED[(rk, fullstep)]))


def test_boost_invariant_rk2_is_marginal():
    """two directions: -2 per stage, RK2 factor 1/2 + (1/2)(-1 + 2) = 1"""
    f = checkerboard_factor(2, 0, lbi=1)
    assert abs(f - 1.0) <= 1e-2, f


def test_wrong_rk3_weights_change_the_factor(tmp_path):
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    path = str(tmp_path / "liboracle_m4.so")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-DOR_MUTANT=4", "-o", path,
                    os.path.join(root, "oracle", "oracle.c"), "-lm"], check=True)
    f = checkerboard_factor(3, 0, L=oracle_ctypes.load(path))
    assert abs(f - PREDICTED[(3, 0)]) > 0.1, f
