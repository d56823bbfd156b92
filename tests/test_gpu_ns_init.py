"""GPU: the Navier-Stokes initialisation (viral_ns_init: CalcNS, NSInit,
ZeroInit of NEWEST10/Source.h:542-776) against the CPU oracle, bit for bit,
on every ghosted cell of random smooth states (conformal and s95p, 3+1D and
boost-invariant).  The oracle itself is pinned to the reference's CalcNS in
tests/test_oracle_ns_init.py."""
import numpy as np
import pytest

from viral_amd import capi, ic
from viral_amd.desc import make_desc, NFIELDS, FIELD_NAMES, EOS_S95P, CARTESIAN, ERR_UNSUPPORTED
import oracle_ctypes

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


CASES = [dict(nx=14, ny=12, nz=10), dict(nx=13, ny=11, nz=9, eos=EOS_S95P), dict(nx=20, ny=18, nz=1, lbi=1)]


def state_for(d, seed):
    if d.eos == EOS_S95P:
        return ic.random_smooth(d, seed=seed, tau=0.9, e_scale=30.0,
                                eos_p=lambda e: capi.eos_eval(d, e)[0].reshape(np.shape(e)))
    return ic.random_smooth(d, seed=seed, tau=0.9)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_ns_init_bitwise(case):
    d = make_desc(**CASES[case])
    st = state_for(d, 40 + case)
    tau, ts = 0.9, 0.01
    print("case ns_init", CASES[case], flush=True)
    h = capi.Hydro(d)
    h.upload(st)
    o = oracle_ctypes.Oracle(d)
    o.set_state(st)
    g, c = h.ns_pi(tau, ts), o.ns_pi(tau, ts)
    assert same_bits(g, c)
    assert np.abs(c).max() > 0
    assert same_bits(h.download(), st), "VIRAL_NS_OUTPUT leaves the state alone"
    for zero in (False, True):
        h.upload(st)
        o.set_state(st)
        h.ns_init(tau, ts, zero=zero)
        o.ns_init(tau, ts, zero=zero)
        gs, cs = h.download(), o.state()
        for f in range(NFIELDS):
            assert same_bits(gs[f], cs[f]), (zero, FIELD_NAMES[f])


def test_ns_init_rejects_cartesian():
    d = make_desc(10, 10, 8, coords=CARTESIAN)
    h = capi.Hydro(d)
    assert h.L.viral_ns_init(h.h, 0, 1.0, 0.01, capi.NS_INIT, None) == ERR_UNSUPPORTED
