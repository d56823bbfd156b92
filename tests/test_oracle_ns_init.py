"""The oracle's Navier-Stokes initialisation against the reference's own
CalcNS / NSInit / ZeroInit (NEWEST10/Source.h:542-776), compiled in place
(oracle/_ref harness modes 4 and 5), on every ghosted cell of a random state.

CalcNS is a computer-algebra expansion of 2 eta sigma^{mu nu}; the oracle
evaluates the tensor form of the shear source's sigma, so the two agree to
rounding (a bound on the pi scale).  The velocity gradients follow
CalcDer4Vel's coverage on a freshly zeroed grid: ghost cells get only the
derivatives CalcDer4Vel writes there, which the comparison over all cells
checks.  NSInit/ZeroInit then set PI (the Navier-Stokes -zeta theta in the
-DBULK build, BULKTEST profile; 0 otherwise) and T^{tau mu}; T^{tau tau} uses
u0*u0 for the reference's pow(u0, 2)."""
import numpy as np
import pytest

import oracle_ctypes
import refdata
from viral_amd.desc import F_T00, F_PI0, F_BULK, NFIELDS

PROFILES = ["ref_kt", "ref_kt_ginit", "ref_kt_gubser", "ref_kt_bulktest"]


def need(name):
    if not refdata.available(name):
        pytest.skip("oracle/_ref/%s not built (reference tree absent)" % name)


@pytest.mark.parametrize("name", PROFILES)
@pytest.mark.parametrize("seed,tau", [(7, 0.8), (99, 1.4)])
def test_ns_pi_matches_calcns(name, seed, tau):
    need(name)
    r = refdata.run(name, seed, tau, 0.01, mode=4)
    o = oracle_ctypes.Oracle(r["desc"])
    o.set_state(r["state"])
    ns = o.ns_pi(r["tau"], r["ts"])
    bulk = r["desc"].bulk
    if bulk:
        ns, nsPI = ns[:10], ns[10]
        sb = np.abs(r["nsPI"]).max()
        assert sb > 0
        assert np.abs(nsPI - r["nsPI"]).max() <= 2e-14 * sb, np.abs(nsPI - r["nsPI"]).max() / sb
    else:
        assert not np.any(r["nsPI"])
    ref = r["nspi"]
    scale = np.abs(ref).max()
    assert scale > 0
    assert np.abs(ns - ref).max() <= 2e-14 * scale, np.abs(ns - ref).max() / scale
    # cells where CalcDer4Vel writes nothing still carry the Christoffel part
    assert np.count_nonzero(ns) == np.count_nonzero(ref)
    # NSInit: pi = nspi, PI = 0, T^{tau mu} from e, P, u, pi
    o.ns_init(r["tau"], r["ts"])
    after, got = r["after"], o.state()
    for f in range(NFIELDS):
        a, b = got[f], after[f]
        if F_PI0 <= f < F_PI0 + 10 or F_T00 <= f < F_T00 + 4 or (bulk and f == F_BULK):
            s = max(np.abs(b).max(), 1e-300)
            assert np.abs(a - b).max() <= 2e-14 * s, (f, np.abs(a - b).max() / s)
        else:
            assert np.array_equal(a, b), f


@pytest.mark.parametrize("name", PROFILES)
def test_zero_init_matches_reference(name):
    need(name)
    r = refdata.run(name, 5, 1.1, 0.01, mode=5)
    o = oracle_ctypes.Oracle(r["desc"])
    o.set_state(r["state"])
    o.ns_init(r["tau"], r["ts"], zero=True)
    got = o.state()
    # bitwise, T^{tau tau} included: glibc's pow(u0, 2) rounds as u0*u0 here
    for f in range(NFIELDS):
        assert np.array_equal(got[f], r["after"][f]), f
    assert not np.any(got[F_PI0:F_PI0 + 10]) and not np.any(got[F_BULK])
