"""GPU: the driver loop's device reductions and output records against the
CPU oracle on the same state.

CheckRoot, DebugMSG extrema and CheckPhysics are functions of the state, so
the oracle is loaded with the state the GPU holds and both are compared:
maxima and reported cells bit for bit, averages within twice the rounding
bound of a sum of n non-negative terms (2 n u relative).  The output records
(write.h) must be equal.
"""
import numpy as np
import pytest

from viral_amd import capi, ic
from viral_amd.desc import (make_desc, CARTESIAN, EOS_S95P, F_EN, F_VX, OK)
import oracle_ctypes

pytestmark = pytest.mark.gpu


def planted(d, seed, tau):
    """a smooth random state with two cells of equal largest |v| >= 1, two of
    equal largest e and two with e <= 0 (as the reference harness plants)"""
    st = ic.random_smooth(d, seed=seed, tau=tau)
    gx, gy, gz = d.ghost_dims()
    lo = 0 if d.lbi else 3
    hi = 0 if d.lbi else 3 + d.nz - 1
    cells = [(4, 5, lo), (3 + d.nx - 2, 3 + d.ny - 1, hi), (6, 3, lo), (3 + d.nx - 1, 4, lo), (5, 7, hi),
             (3 + d.nx - 3, 3 + d.ny - 2, lo)]
    for q, (i, j, k) in enumerate(cells):
        x = (i * gy + j) * gz + k
        if q < 2:
            st[F_VX, x], st[F_VX + 1, x], st[F_VX + 2, x] = 0.8, -0.7, 0.0
        elif q < 4:
            st[F_EN, x] = 90.0
        else:
            st[F_EN, x] = -0.5 if q == 4 else 0.0
    return st


CASES = [
    dict(nx=70, ny=60, nz=40),                        # 168000 cells: the fixed grid strides twice
    dict(nx=24, ny=20, nz=1, lbi=1),
    dict(nx=18, ny=16, nz=10, eos=EOS_S95P, fast_math=1, eos_table=1),
    dict(nx=16, ny=14, nz=9, coords=CARTESIAN),
]


@pytest.mark.parametrize("kw", CASES, ids=["milne", "lbi", "s95p_fast", "cartesian"])
def test_diagnostics_match_oracle(kw):
    d = make_desc(**kw)
    tau = 1.0 if d.coords == CARTESIAN else 1.3
    print("case diagnostics %r" % kw, flush=True)
    h = capi.Hydro(d)
    dref = d.copy()
    dref.fast_math = dref.eos_table = 0
    o = oracle_ctypes.Oracle(dref)
    # a planted state as uploaded, then the state after two steps of a clean one
    for phase in range(2):
        if phase == 0:
            h.upload(planted(d, 21, tau))
            t = tau
        else:
            h.upload(ic.random_smooth(d, seed=22, tau=tau))
            r, done = h.steps(tau, 0.004, 2)
            assert (r, done) == (OK, 2)
            t = tau + 0.004 + 0.004
        st = h.download()
        o.set_state(st)
        mx, avg, at = h.check_root(t)
        omx, oavg, oat = o.check_root(t)
        assert np.array_equal(mx, omx) and np.array_equal(at, oat), (phase, mx, omx, at, oat)
        # sums of n non-negative terms: the oracle's loop-order sum and the
        # device's fixed tree are each within n*u of the exact sum
        tol = 2 * d.nx * d.ny * (1 if d.lbi else d.nz) * 1.12e-16
        assert np.all(np.abs(avg - oavg) <= tol * np.abs(oavg) + 1e-300), (phase, avg, oavg)
        ex, eat = h.debug_extrema(t)
        oex, oeat = o.debug_msg(t)
        assert np.array_equal(ex, oex) and np.array_equal(eat, oeat), (phase, ex, oex, eat, oeat)
        n, first = h.check_physics(t)
        on, ofirst = o.check_physics(t)
        assert np.array_equal(n, on) and np.array_equal(first, ofirst), (phase, n, on, first, ofirst)
        if phase == 0:
            assert n[0] == 2 and n[1] == 2 and ex[5] == -0.5
        assert h.max_temp_gev() == o.max_temp_gev()


@pytest.mark.parametrize("kw,strides", [(dict(nx=20, ny=16, nz=12), [(1, 1, 0), (2, 1, 0), (4, 3, 0), (2, 1, 1)]),
                                        (dict(nx=24, ny=20, nz=1, lbi=1), [(1, 1, 1), (2, 1, 1)])])
def test_output_records_match_oracle(kw, strides):
    d = make_desc(**kw)
    print("case output %r" % kw, flush=True)
    h = capi.Hydro(d)
    h.upload(ic.random_smooth(d, seed=23, tau=1.1))
    r, done = h.steps(1.1, 0.004, 1)
    assert (r, done) == (OK, 1)
    o = oracle_ctypes.Oracle(d)
    o.set_state(h.download())
    for f, fz, xy in strides:
        got = h.output(f, fz, xy)
        want = o.output(f, fz, xy)
        assert np.array_equal(got, want), (f, fz, xy)
    with pytest.raises(capi.ViralError):
        h.output(3, 1, 0)
