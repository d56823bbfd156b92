"""The oracle against the reference's own code, compiled in place from the
reference tree (oracle/build_ref.sh -> oracle/_ref).  Skipped where the
harness binaries were not built (no reference tree); the committed golden
fixtures (tests/test_golden.py) carry the same pins everywhere.

Pinned here, per compile-time profile of NEWEST10/maindef.h:
  CalcSource (geometric + velocity gradients + shear, and the bulk row of
  the BULKTEST profile), NEWEST10/Source.h:195
  CalcCentreFlux + hydroExplicit (KT/WENO Result), NEWEST10/hydroKT.h:16,406
  boundary (ghost fill), NEWEST10/boundary.h:266
  s95p_p/a/s/T, NEWEST10/s95p.h
"""
import numpy as np
import pytest

import oracle_ctypes
import refdata

PROFILES = ["ref_kt", "ref_kt_gubser", "ref_kt_ginit", "ref_shas", "ref_shas_ginit", "ref_kt_bulktest"]


def need(name):
    if not refdata.available(name):
        pytest.skip("oracle/_ref/%s not built (reference tree absent)" % name)


@pytest.fixture(scope="module", params=[(p, s) for p in PROFILES for s in (7, 1234)],
                ids=lambda v: "%s-seed%d" % v)
def case(request):
    name, seed = request.param
    need(name)
    r = refdata.run(name, seed, 1.3 if seed == 7 else 0.7, 0.01)
    o = oracle_ctypes.Oracle(r["desc"])
    o.set_state(r["state"])
    return r, o


def test_source_matches_reference(case):
    r, o = case
    d = r["desc"]
    S = o.calc_source(r["tau"], r["ts"]).reshape((15,) + d.ghost_dims())
    Sr = r["source"].reshape((15,) + d.ghost_dims())
    I = d.interior_slices()
    for l in range(15):
        a, b = S[l][I], Sr[l][I]
        if l < 4 or (l == 14 and not d.bulk):
            # geometric rows are the reference's expressions verbatim (and
            # the bulk row is zero without -DBULK)
            assert np.array_equal(a, b), l
        else:
            # shear rows and the bulk row: covariant form vs the reference's
            # expanded algebra, equal up to rounding
            scale = np.abs(b).max()
            assert np.abs(a - b).max() <= 1e-14 * scale, (l, np.abs(a - b).max() / scale)


def test_kt_result_bitwise(case):
    r, o = case
    d = r["desc"]
    R = o.hydro_explicit(r["tau"], r["ts"] / 2).reshape((15,) + d.ghost_dims())
    Rr = r["result"].reshape((15,) + d.ghost_dims())
    I = d.interior_slices()
    for l in range(15):
        assert np.array_equal(R[l][I], Rr[l][I]), l


def test_boundary_bitwise(case):
    r, o = case
    if not r["one_rank"]:
        pytest.skip("profile has several ranks; boundary() needs MPI neighbours")
    o.boundary()
    st = o.state()
    for f in range(st.shape[0]):
        assert np.array_equal(st[f], r["after_boundary"][f]), f


def test_s95p_matches_reference():
    need("ref_kt")
    tab = refdata.run("ref_kt", 3, 1.0, 0.01, mode=1)
    from viral_amd.desc import make_desc, EOS_S95P
    d = make_desc(4, 4, 4, eos=EOS_S95P)
    L = oracle_ctypes.lib()
    for e, p, a, s, T in tab:
        for got, ref in ((L.or_eos_p(d, e), p), (L.or_eos_dpde(d, e), a), (L.or_eos_T(d, e), T)):
            assert abs(got - ref) <= 1e-13 * abs(ref), (e, got, ref)
        assert abs(L.or_eos_s(d, e, p, T) - s) <= 1e-13 * abs(s), (e,)
