"""The oracle against committed fixtures of the reference's own CalcSource
and KT Result (made by tests/golden/make_golden.py from the reference code
compiled in place).  Runs everywhere, reference tree or not."""
import os

import numpy as np
import pytest

import oracle_ctypes
import refdata
from viral_amd.desc import make_desc, NFIELDS

HERE = os.path.dirname(os.path.abspath(__file__))
DATA = np.load(os.path.join(HERE, "golden", "ref_boxes.npz"))
CASES = sorted({k.split("_")[0] for k in DATA.files})


@pytest.mark.parametrize("c", CASES)
def test_box_matches_reference(c):
    name = str(DATA[c + "_profile"])
    tau, ts = DATA[c + "_tau_ts"]
    box = DATA[c + "_state"]
    prof = dict(refdata.PROFILES[name])
    lbi = prof.get("lbi", 0)
    prof.update(nx=3, ny=3, nz=1 if lbi else 3, quirks=0)   # halo cells are interior cells of the reference grid
    d = make_desc(**prof)
    assert box.shape == (NFIELDS,) + d.ghost_dims()
    o = oracle_ctypes.Oracle(d)
    o.set_state(box.reshape(NFIELDS, -1))
    I = d.interior_slices()
    S = o.calc_source(tau, ts).reshape((15,) + d.ghost_dims())
    R = o.hydro_explicit(tau, ts / 2).reshape((15,) + d.ghost_dims())
    Sref, Rref = DATA[c + "_source"], DATA[c + "_result"]
    for l in range(15):
        assert np.array_equal(R[l][I], Rref[l]), ("result", l)
        if l < 4 or l == 14:
            assert np.array_equal(S[l][I], Sref[l]), ("source", l)
        else:
            scale = max(np.abs(Sref[l]).max(), 1e-300)
            assert np.abs(S[l][I] - Sref[l]).max() <= 1e-14 * scale, ("source", l)
