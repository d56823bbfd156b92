"""GPU: Landau matching (viral_landau: FindVar and the per-cell part of
ginit, NEWEST10/ginit.h:521-629, 876-990) against the CPU oracle, bit for
bit, on every ghosted cell: fluid fields with a known Landau frame (conformal
and s95p, tau below and above 1), fields of ginit's own construction, the
prevu call (at_start 0), and fields with failing cells (same first bad
cell).  The oracle is pinned to FindVar's selection rule in
tests/test_oracle_landau.py."""
import numpy as np
import pytest

from viral_amd import capi
from viral_amd.desc import make_desc, NFIELDS, FIELD_NAMES, EOS_S95P, OK, ERR_PHYSICS, FLAG_SOLVER, CARTESIAN
import landau_fields as lf
import oracle_ctypes

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def eos_p_of(d):
    L = oracle_ctypes.lib()
    return lambda e: np.array([L.or_eos_p(d, float(x)) for x in np.atleast_1d(e)])


def compare(d, tau, T, at_start, seed=0):
    rng = np.random.default_rng(seed)
    st = rng.uniform(0.1, 1.0, size=(NFIELDS, d.cells()))
    h = capi.Hydro(d)
    h.upload(st)
    o = oracle_ctypes.Oracle(d)
    o.set_state(st)
    r = h.landau(tau, T, at_start)
    nbad, first = o.landau(tau, T, at_start)
    g, c = h.download(), o.state()
    for f in range(NFIELDS):
        assert same_bits(g[f], c[f]), FIELD_NAMES[f]
    assert (r == OK) == (nbad == 0)
    fl = h.flags()
    assert np.count_nonzero(fl & FLAG_SOLVER) == nbad
    if nbad:
        i, j, k, stage, flags = h.first_bad_cell()
        gx, gy, gz = d.ghost_dims()
        assert (i * gy + j) * gz + k == first and flags == FLAG_SOLVER
    return nbad


@pytest.mark.parametrize("eos,tau", [(0, 0.6), (EOS_S95P, 0.1), (EOS_S95P, 1.7)])
@pytest.mark.parametrize("at_start", [True, False])
def test_known_frame_bitwise(eos, tau, at_start):
    d = make_desc(9, 8, 6, eos=eos)
    print("case landau eos=%d tau=%g at_start=%d" % (eos, tau, at_start), flush=True)
    T, e, v, pi = lf.random_fluid_field(d.cells(), 5, tau, eos_p_of(d), vmax=0.6)
    assert compare(d, tau, T, at_start) == 0


@pytest.mark.parametrize("seed", [1, 2])
def test_ginit_field_bitwise(seed):
    d = make_desc(8, 8, 6, eos=EOS_S95P)
    print("case landau ginit seed=%d" % seed, flush=True)
    T = lf.ginit_like_field(d.cells(), seed, 0.1)
    compare(d, 0.1, T, True, seed)


def test_failing_cells_reported():
    d = make_desc(6, 6, 1, lbi=1)
    n = d.cells()
    print("case landau failures", flush=True)
    T, e, v, pi = lf.random_fluid_field(n, 9, 2.0, lambda e: np.asarray(e) / 3.0, vmax=0.3)
    bad = np.array([1.0, 2.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.3, 0.0, 0.3 / 4.0])   # |T^{tx}| > T^{tt}
    fast_eta, _ = lf.fluid(2.0, 2.0 / 3.0, np.array([0.0, 0.0, 0.35]), np.zeros(10), 2.0)   # two pass at tau = 2
    T[:, 17] = bad
    T[:, 5] = fast_eta
    assert compare(d, 2.0, T, True) == 2


def test_landau_rejects_cartesian():
    d = make_desc(6, 6, 4, coords=CARTESIAN)
    h = capi.Hydro(d)
    with pytest.raises(capi.ViralError):
        h.landau(1.0, np.zeros((10, d.cells())))
