"""The product's EOS (viral_eos_eval, host side of the same code the kernels
run): the closed form equals the oracle's bit for bit, and the fast-mode
table stays within 1e-10 of the closed form (NEWEST10/s95p.h), including
around the branch points of the piecewise fit."""
import numpy as np
import pytest

import oracle_ctypes
from viral_amd import capi
from viral_amd.desc import make_desc, EOS_S95P, EOS_CONFORMAL, GEVFM

L = oracle_ctypes.lib()
BREAKS = np.array([0.5028563305441270, 1.62, 1.86, 9.9878355786273545, 0.1270769021427449, 0.446707952467404,
                   1.9402832534193788, 3.7292474570977285, 0.5143939846236409]) / GEVFM


def energies(seed=0, n=40000):
    rng = np.random.default_rng(seed)
    e = 10 ** rng.uniform(-6, np.log10(5000.0), n)
    near = (BREAKS[:, None] * (1 + rng.uniform(-3e-3, 3e-3, (len(BREAKS), 200)))).ravel()
    return np.concatenate([e, near, BREAKS, [1e-5, 1e-5 * (1 + 1e-12), 2e-5]])


@pytest.mark.parametrize("eos", [EOS_CONFORMAL, EOS_S95P])
def test_closed_form_equals_oracle_bitwise(eos):
    d = make_desc(4, 4, 4, eos=eos)
    e = energies(1, 5000)
    p, a, T, s = capi.eos_eval(d, e)
    po = np.array([L.or_eos_p(d, v) for v in e])
    ao = np.array([L.or_eos_dpde(d, v) for v in e])
    To = np.array([L.or_eos_T(d, v) for v in e])
    so = np.array([L.or_eos_s(d, v, pp, tt) for v, pp, tt in zip(e, po, To)])
    for got, ref in ((p, po), (a, ao), (T, To), (s, so)):
        assert np.array_equal(got.view(np.int64), ref.view(np.int64))


def test_table_within_1e10_of_closed_form():
    dc = make_desc(4, 4, 4, eos=EOS_S95P)
    dt = make_desc(4, 4, 4, eos=EOS_S95P, fast_math=1, eos_table=1)
    e = energies(2)
    ref = capi.eos_eval(dc, e)
    got = capi.eos_eval(dt, e)
    worst = [float(np.max(np.abs(g - r) / np.abs(r))) for g, r in zip(got, ref)]
    print("max relative error p, dp/de, T, s:", worst)
    assert max(worst) <= 1e-10
    # the table is actually used away from the branch points
    away = np.min(np.abs(e[:, None] / BREAKS[None, :] - 1), axis=1) > 0.02
    assert np.any(got[0][away & (e > 1e-4)] != ref[0][away & (e > 1e-4)])
