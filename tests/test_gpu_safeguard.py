"""GPU: the root solve alone (viral_root, RootSearchForEnVelUsingDerivatives)
and the opt-in safeguard (VIRAL_Q_ROOT_SAFEGUARD; the reference has none and
aborts).

A Newton guess just above the pole of f at e = -3 (A - PI) runs to the
quadratic's negative root.  Without the safeguard the solve fails as the
reference would (same first bad cell as the oracle); with it those cells are
RESCUED onto the positive root.  Either way the GPU result equals the
oracle's bit for bit in exact mode: every field, the iteration counts and the
flags.  (A full step cannot carry such a guess: the cell's own sources would
read T(e) of a negative e.)
"""
import numpy as np
import pytest

from viral_amd import capi, ic
from viral_amd.desc import (make_desc, NFIELDS, FIELD_NAMES, F_EN, F_T00, F_PI0, F_BULK, EOS_S95P, Q_DEFAULT,
                            Q_ROOT_SAFEGUARD, FLAG_RESCUED, FLAG_ENEG, FLAG_SOLVER, OK, ERR_PHYSICS)
import metrics
import oracle_ctypes

pytestmark = pytest.mark.gpu

TAU = 1.1


def same_bits(a, b):
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(nan | (a.view(np.int64) == b.view(np.int64))))


def planted_state(d, seed=31):
    if d.eos == EOS_S95P:
        dref = d.copy()
        dref.fast_math = dref.eos_table = 0
        st = ic.random_smooth(d, seed=seed, tau=TAU, e_scale=30.0,
                              eos_p=lambda e: capi.eos_eval(dref, e)[0].reshape(np.shape(e)))
    else:
        st = ic.random_smooth(d, seed=seed, tau=TAU)
    gx, gy, gz = d.ghost_dims()
    cells = [((4 * gy + 5) * gz + 4), ((9 * gy + 3) * gz + 6), ((6 * gy + 8) * gz + 5)]
    for x in cells:
        st[F_EN, x] = -2.95 * (st[F_T00, x] - st[F_PI0, x] - st[F_BULK, x])
    return st, cells


def failing_cells(d, st):
    """cells whose plain Newton solve fails (oracle, no safeguard)"""
    dd = d.copy()
    dd.quirks = Q_DEFAULT
    o = oracle_ctypes.Oracle(dd)
    o.set_state(st)
    bad, it, fl = o.root(TAU)
    return [x for x in range(st.shape[1]) if fl[x] & (FLAG_ENEG | FLAG_SOLVER)]


@pytest.mark.parametrize("eos", [0, EOS_S95P])
def test_root_alone_and_safeguard_match_oracle(eos):
    """conformal: two of the planted guesses end on the negative root; s95p:
    its linear p(e) below 1e-5 keeps Newton on the positive branch"""
    for quirks in (Q_DEFAULT, Q_DEFAULT | Q_ROOT_SAFEGUARD):
        d = make_desc(12, 11, 8, eos=eos, quirks=quirks)
        st, cells = planted_state(d)
        print("case root eos=%d quirks=%d" % (eos, quirks), flush=True)
        h = capi.Hydro(d)
        h.upload(st)
        o = oracle_ctypes.Oracle(d)
        o.set_state(st)
        rg = h.root(TAU)
        bad, it, fl = o.root(TAU)
        g, c = h.download(), o.state()
        for f in range(NFIELDS):
            assert same_bits(g[f], c[f]), FIELD_NAMES[f]
        assert np.array_equal(h.iterations(0), it)
        assert np.array_equal(h.flags(), fl)
        failing = failing_cells(d, st)
        assert len(failing) == (2 if eos == 0 else 0)
        if quirks & Q_ROOT_SAFEGUARD:
            assert rg == OK and bad == -1
            assert [x for x in range(st.shape[1]) if fl[x] & FLAG_RESCUED] == failing
        elif failing:
            assert rg == ERR_PHYSICS and bad == failing[0]
            i, j, k, stage, flags = h.first_bad_cell()
            gx, gy, gz = d.ghost_dims()
            assert (i * gy + j) * gz + k == bad and flags == fl[bad]
        else:
            assert rg == OK


def test_root_fast_mode_within_bound():
    d = make_desc(12, 11, 8, quirks=Q_DEFAULT | Q_ROOT_SAFEGUARD, fast_math=1)
    st, cells = planted_state(d)
    print("case root fast", flush=True)
    h = capi.Hydro(d)
    h.upload(st)
    dref = d.copy()
    dref.fast_math = 0
    o = oracle_ctypes.Oracle(dref)
    o.set_state(st)
    assert h.root(TAU) == OK
    bad, it, fl = o.root(TAU)
    assert bad == -1
    errs = metrics.state_errors(d, h.download(), o.state())
    assert metrics.worst(errs) <= 1e-10, errs
    assert all(h.flags()[x] & FLAG_RESCUED for x in failing_cells(d, st))
