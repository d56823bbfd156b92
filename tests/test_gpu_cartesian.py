"""GPU parity of the Cartesian (TXYZ5) path against the CPU oracle.

Exact mode: bitwise equality of every state field, of the per-stage
iteration counts of the 4-D root solve and of the flags.  Fast mode: the
1e-10-per-step tolerance on physical scales.
"""
import numpy as np
import pytest

from viral_amd import capi, ic
from viral_amd.desc import (make_desc, CARTESIAN, NFIELDS, FIELD_NAMES, EOS_S95P, ETA_S_OVER_4PI, TPI_OVER_P,
                            Q_DEFAULT, Q_SHEAR_SLICES, OK)
import metrics
import oracle_ctypes

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(nan | (a.view(np.int64) == b.view(np.int64))))


def cart_state(d, seed, tau=1.0):
    if d.eos == EOS_S95P:
        dref = d.copy()
        dref.fast_math = dref.eos_table = 0
        eosp = lambda e: oracle_ctypes.lib().or_eos_p(dref, e)
        return ic.random_smooth(d, seed=seed, tau=tau, e_scale=30.0, eos_p=eosp)
    return ic.random_smooth(d, seed=seed, tau=tau)


def run_pair(d, st, nsteps, tau=1.0, ts=0.004):
    print("case cartesian nx=%d ny=%d nz=%d rk=%d eos=%d quirks=%d fast=%d" %
          (d.nx, d.ny, d.nz, d.rk, d.eos, d.quirks, d.fast_math), flush=True)
    h = capi.Hydro(d)
    h.upload(st)
    dref = d.copy()
    dref.fast_math = dref.eos_table = 0
    o = oracle_ctypes.Oracle(dref)
    o.set_state(st)
    t = tau
    for s in range(nsteps):
        assert h.step(t, ts) == OK
        assert o.step(t, ts) == OK
        t = t + ts
    return h, o


@pytest.mark.parametrize("rk", [1, 2, 3])
def test_cartesian_exact_bitwise(rk):
    d = make_desc(14, 11, 9, coords=CARTESIAN, rk=rk, eta_law=ETA_S_OVER_4PI, eta_coef=1.0, tpi_law=TPI_OVER_P,
                  tpi_coef=1.5)
    h, o = run_pair(d, cart_state(d, 80 + rk), 2)
    g, c = h.download(), o.state()
    for f in range(NFIELDS):
        assert same_bits(g[f], c[f]), FIELD_NAMES[f]
    for s in range(rk):
        assert np.array_equal(h.iterations(s), o.iters(s)), s
    assert np.array_equal(h.flags(), o.flags())
    assert h.iterations(0).max() > 2


@pytest.mark.parametrize("quirks", [Q_DEFAULT, Q_DEFAULT & ~Q_SHEAR_SLICES])
def test_cartesian_s95p_and_q9_bitwise(quirks):
    d = make_desc(10, 12, 8, coords=CARTESIAN, eos=EOS_S95P, quirks=quirks)
    h, o = run_pair(d, cart_state(d, 91), 2)
    g, c = h.download(), o.state()
    for f in range(NFIELDS):
        assert same_bits(g[f], c[f]), FIELD_NAMES[f]
    assert np.array_equal(h.iterations(1), o.iters(1))


def test_cartesian_fast_mode_within_1e10():
    d = make_desc(24, 20, 10, coords=CARTESIAN, fast_math=1)
    st = cart_state(d, 95)
    h, o = run_pair(d, st, 1)
    errs = metrics.state_errors(d, h.download(), o.state())
    print(errs)
    assert metrics.worst(errs) <= 1e-10, errs


def test_cartesian_ensemble_graph_bitwise():
    ds = [make_desc(9, 7, 6, coords=CARTESIAN), make_desc(6, 10, 11, coords=CARTESIAN, eos=EOS_S95P)]
    sts = [cart_state(d, 97 + q) for q, d in enumerate(ds)]
    print("case cartesian ensemble", flush=True)
    h = capi.Hydro(ds)
    h.set_launch(batched=1, graph=1)
    for ev, st in enumerate(sts):
        h.upload(st, ev=ev)
    r, done = h.steps(1.0, 0.004, 2)
    assert (r, done) == (OK, 2)
    for ev, (d, st) in enumerate(zip(ds, sts)):
        o = oracle_ctypes.Oracle(d)
        o.set_state(st)
        assert o.step(1.0, 0.004) == OK and o.step(1.0 + 0.004, 0.004) == OK
        g, c = h.download(ev), o.state()
        for f in range(NFIELDS):
            assert same_bits(g[f], c[f]), (ev, FIELD_NAMES[f])


def test_cartesian_rejects_shasta_and_lbi():
    from viral_amd.desc import SHASTA
    from viral_amd.capi import ViralError
    for kw in (dict(scheme=SHASTA), dict(lbi=1, nz=1)):
        args = dict(coords=CARTESIAN)
        args.update(kw)
        nz = args.pop("nz", 5)
        with pytest.raises(ViralError):
            capi.Hydro(make_desc(6, 6, nz, **args))
