"""GPU: the FIX build (VIRAL_Q_FIX_VE: FixTwo + DebugMSG's BMax at the end of
every step, NEWEST10/time.h:280-290, 385-392) against the CPU oracle, bit for
bit, in exact mode: a planted fast cell triggers the zeroing of interior
v_eta in the first step, the next steps run with the new BMax.  Steps are
taken one call at a time and in one viral_steps call (hipGraph on and off),
and in a two-event context where only one event carries the quirk.  The
oracle's FIX rule is pinned in tests/test_oracle_fix.py."""
import numpy as np
import pytest

from viral_amd import capi, ic
from viral_amd.desc import make_desc, Q_DEFAULT, Q_FIX_VE, F_VE, F_VX, F_VY, NFIELDS, FIELD_NAMES, \
    ERR_UNSUPPORTED
import oracle_ctypes

pytestmark = pytest.mark.gpu

TAU, TS, NSTEPS = 1.0, 0.01, 3
SHAPES = [dict(nx=14, ny=12, nz=10), dict(nx=16, ny=14, nz=1, lbi=1)]


def same_bits(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def planted(d, seed=5):
    st = ic.random_smooth(d, seed=seed, tau=TAU)
    gx, gy, gz = d.ghost_dims()
    x = (7 * gy + 6) * gz + (0 if d.lbi else 5)
    st[F_VX, x], st[F_VY, x] = 0.99, 0.1
    return st


def oracle_states(d, st):
    o = oracle_ctypes.Oracle(d)
    o.set_state(st)
    out, tau = [], TAU
    for s in range(NSTEPS):
        assert o.step(tau, TS) == 0
        out.append(o.state())
        tau += TS
    return out


def assert_state(g, c, what):
    for f in range(NFIELDS):
        assert same_bits(g[f], c[f]), (what, FIELD_NAMES[f])


@pytest.mark.parametrize("shape", range(len(SHAPES)))
@pytest.mark.parametrize("graph", [0, 1])
def test_fix_matches_oracle(shape, graph):
    d = make_desc(**SHAPES[shape], quirks=Q_DEFAULT | Q_FIX_VE)
    st = planted(d)
    ref = oracle_states(d, st)
    m = np.zeros(d.ghost_dims(), dtype=bool)
    m[d.interior_slices()] = True
    assert np.all(ref[0][F_VE][m.ravel()] == 0.0) and np.abs(ref[1][F_VE]).max() > 0
    print("case fix", SHAPES[shape], "graph", graph, flush=True)
    h = capi.Hydro(d)
    h.set_launch(graph=graph)
    h.upload(st)
    tau = TAU
    for s in range(NSTEPS):
        assert h.step(tau, TS) == capi.OK
        assert_state(h.download(), ref[s], ("step", s))
        tau += TS
    h.upload(st)
    r, done = h.steps(TAU, TS, NSTEPS)
    assert r == capi.OK and done == NSTEPS
    assert_state(h.download(), ref[-1], "steps")


def test_fix_per_event_in_ensemble():
    kw = SHAPES[0]
    dfix = make_desc(**kw, quirks=Q_DEFAULT | Q_FIX_VE)
    dplain = make_desc(**kw)
    st = planted(dfix)
    print("case fix ensemble", flush=True)
    h = capi.Hydro([dfix, dplain])
    h.set_launch(batched=1, graph=1)
    for e in range(2):
        h.upload(st, ev=e)
    r, done = h.steps(TAU, TS, NSTEPS)
    assert r == capi.OK and done == NSTEPS
    assert_state(h.download(0), oracle_states(dfix, st)[-1], "fix event")
    assert_state(h.download(1), oracle_states(dplain, st)[-1], "plain event")


def test_fix_refused_by_adjoint():
    d = make_desc(**SHAPES[0], quirks=Q_DEFAULT | Q_FIX_VE)
    h = capi.Hydro(d)
    assert h.L.viral_adjoint_begin(h.h, 5) == ERR_UNSUPPORTED
