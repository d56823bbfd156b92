"""GPU parity: the HIP step (through the C ABI) against the CPU oracle.

In exact mode the kernels evaluate every expression in the reference's
operation order, so the bar is bitwise equality of all 29 state fields, of
the per-stage Newton iteration counts and of the per-cell flags.
"""
import ctypes as C

import numpy as np
import pytest

from viral_amd import capi, ic
from viral_amd.desc import (make_desc, NFIELDS, FIELD_NAMES, EOS_S95P, Q_DEFAULT, Q_GHOST_EIGEN,
                            Q_K0_BROADCAST, Q_GUBSER_MIRROR, ETA_S_OVER_4PI, TPI_OVER_P, TPI_CONST,
                            F_T10, F_EN, FLAG_ACAUSAL, OK, ERR_PHYSICS)
import oracle_ctypes

pytestmark = pytest.mark.gpu


def run_both(desc, state, tau, ts, nsteps):
    print("case nx=%d ny=%d nz=%d lbi=%d rk=%d eos=%d quirks=%d steps=%d" %
          (desc.nx, desc.ny, desc.nz, desc.lbi, desc.rk, desc.eos, desc.quirks, nsteps), flush=True)
    h = capi.Hydro(desc)
    h.upload(state)
    o = oracle_ctypes.Oracle(desc)
    o.set_state(state)
    for s in range(nsteps):
        t = tau + s * ts
        rg = h.step(t, ts)
        ro = o.step(t, ts)
        assert rg == ro, (s, rg, ro)
    return h, o


def same_bits(a, b):
    """bitwise equality, with any NaN equal to any NaN (payloads differ
    between x86 and gfx950)"""
    nan = np.isnan(a) & np.isnan(b)
    return np.all(nan | (a.view(np.int64) == b.view(np.int64)))


def assert_bitwise(h, o, nstages):
    g = h.download()
    c = o.state()
    for f in range(NFIELDS):
        if not same_bits(g[f], c[f]):
            bad = np.flatnonzero(~((np.isnan(g[f]) & np.isnan(c[f])) | (g[f].view(np.int64) == c[f].view(np.int64))))
            raise AssertionError("field %s differs in %d cells (first %d: gpu %r oracle %r)" %
                                 (FIELD_NAMES[f], bad.size, bad[0], g[f][bad[0]], c[f][bad[0]]))
    for s in range(nstages):
        assert np.array_equal(h.iterations(s), o.iters(s)), "iteration counts of stage %d" % s
    assert np.array_equal(h.flags(), o.flags())


def test_gubser_c1_grid_bitwise():
    """Config 1 grid and parameters (64x64x8, dx=deta=0.1, dtau=0.01, tau0=1,
    eta/s=0.2, tau_pi=5eta/(e+p), RK2), a few steps."""
    d = make_desc(64, 64, 8)
    tau, ts = 1.0, 0.01
    st = ic.gubser(d, tau, ts)
    h, o = run_both(d, st, tau, ts, 3)
    assert_bitwise(h, o, 2)
    it = h.iterations(1)
    assert it.max() > 1


@pytest.mark.parametrize("rk", [1, 2, 3])
def test_random_state_rk_orders_bitwise(rk):
    d = make_desc(21, 18, 9, rk=rk, eta_law=ETA_S_OVER_4PI, eta_coef=1.0, tpi_law=TPI_OVER_P, tpi_coef=1.5)
    tau, ts = 0.8, 0.005
    st = ic.random_smooth(d, seed=11 + rk, tau=tau)
    h, o = run_both(d, st, tau, ts, 2)
    assert_bitwise(h, o, rk)


@pytest.mark.parametrize("quirks", [0, Q_GHOST_EIGEN, Q_K0_BROADCAST, Q_DEFAULT | Q_GUBSER_MIRROR])
def test_quirk_modes_bitwise(quirks):
    d = make_desc(13, 16, 7, quirks=quirks)
    tau, ts = 1.1, 0.004
    st = ic.random_smooth(d, seed=5, tau=tau)
    h, o = run_both(d, st, tau, ts, 2)
    assert_bitwise(h, o, 2)


def test_s95p_eos_bitwise():
    d = make_desc(12, 14, 9, eos=EOS_S95P, eta_law=ETA_S_OVER_4PI, eta_coef=1.0)
    tau, ts = 0.6, 0.004
    eosp = lambda e: oracle_ctypes.lib().or_eos_p(d, e)
    # energies spanning all s95p branches (0.1 .. 100 GeV/fm^3)
    st = ic.random_smooth(d, seed=3, tau=tau, e_scale=60.0, eos_p=eosp)
    h, o = run_both(d, st, tau, ts, 2)
    assert_bitwise(h, o, 2)


def test_lbi_grid_bitwise():
    d = make_desc(24, 20, 1, lbi=1, quirks=Q_DEFAULT | Q_GUBSER_MIRROR, tpi_law=TPI_CONST, tpi_coef=0.3)
    tau, ts = 1.0, 0.01
    st = ic.gubser(d, tau, ts)
    h, o = run_both(d, st, tau, ts, 3)
    assert_bitwise(h, o, 2)


@pytest.mark.parametrize("rk", [2, 3])
def test_shasta_bitwise(rk):
    """SHASTA scheme with the CON gradient sources (-DSHAS)."""
    from viral_amd.desc import SHASTA
    d = make_desc(17, 15, 9, scheme=SHASTA, rk=rk, eta_law=ETA_S_OVER_4PI, eta_coef=1.0,
                  tpi_law=TPI_OVER_P, tpi_coef=1.5)
    tau, ts = 0.7, 0.004
    st = ic.random_smooth(d, seed=31 + rk, tau=tau)
    h, o = run_both(d, st, tau, ts, 2)
    assert_bitwise(h, o, rk)


def test_shasta_launch_options_other_than_its_tested_configuration_are_refused():
    """SHASTA runs its LDS line tiles with 256-thread blocks, one tile per
    block and the sources in their own kernel; persistent blocks, source
    fusion, other block sizes and the autotuner are refused (a host check:
    no kernel runs)"""
    from viral_amd.desc import SHASTA, ERR_UNSUPPORTED
    d = make_desc(12, 10, 6, scheme=SHASTA, fast_math=1)
    h = capi.Hydro(d)
    L = h.launch()
    for field, value in (("blocks_per_cu", 2), ("fusion", 1)):
        c = h.launch()
        setattr(c, field, value)
        assert h.L.viral_set_launch(h.h, C.byref(c)) == ERR_UNSUPPORTED, field
    c = h.launch()
    c.flux_nt[0] = 128
    assert h.L.viral_set_launch(h.h, C.byref(c)) == ERR_UNSUPPORTED
    assert h.L.viral_set_launch(h.h, C.byref(L)) == OK
    assert h.L.viral_autotune(h.h, -1, 1) == ERR_UNSUPPORTED


def test_shasta_s95p_lbi_bitwise():
    from viral_amd.desc import SHASTA
    d = make_desc(20, 18, 1, lbi=1, scheme=SHASTA, eos=EOS_S95P)
    tau, ts = 0.6, 0.004
    eosp = lambda e: oracle_ctypes.lib().or_eos_p(d, e)
    st = ic.random_smooth(d, seed=41, tau=tau, e_scale=30.0, eos_p=eosp)
    h, o = run_both(d, st, tau, ts, 2)
    assert_bitwise(h, o, 2)


def test_acausal_cell_reported_like_reference():
    """A cell pushed to v >= 1 is flagged and reported as the first bad cell
    in (i,j,k) order, where the reference exits (NEWEST10/root.h:396-421)."""
    d = make_desc(10, 9, 7)
    tau, ts = 1.0, 0.01
    st = ic.random_smooth(d, seed=9, tau=tau)
    gx, gy, gz = d.ghost_dims()
    for (i, j, k) in [(6, 5, 4), (8, 3, 5)]:
        x = (i * gy + j) * gz + k
        st[F_T10][x] = 50.0 * st[F_EN][x]
    h, o = run_both(d, st, tau, ts, 1)
    assert h.first_bad_cell() == o.first_bad()
    bad = h.first_bad_cell()
    assert bad is not None and bad[3] == 0 and bad[4] & FLAG_ACAUSAL
    assert np.array_equal(h.flags(), o.flags())


def test_max_temp_matches_oracle():
    d = make_desc(16, 12, 9)
    st = ic.random_smooth(d, seed=1, tau=1.0)
    h = capi.Hydro(d)
    h.upload(st)
    o = oracle_ctypes.Oracle(d)
    o.set_state(st)
    assert h.max_temp_gev() == o.max_temp_gev()


# ---------------------------------------------------------------- fast mode
# fast_math=1 uses one division per WENO reconstruction, shared smoothness
# indicators and FMAs; the bar is the tolerance the target states: relative
# error <= 1e-10 per step and <= 1e-8 after 100 steps against the oracle,
# per cell on physical scales (tests/metrics.py: e vs e, u vs u^0, pi vs e+P,
# T^{tau mu} vs T^{tau tau}).  A one-ulp perturbation of the oracle's own
# input moves these by ~1e-13 over 100 Gubser steps.
import metrics


def test_fast_mode_one_step_within_1e10():
    d = make_desc(40, 36, 9, fast_math=1)
    dref = d.copy()
    dref.fast_math = 0
    tau, ts = 0.9, 0.005
    st = ic.random_smooth(d, seed=21, tau=tau)
    print("case fast one step", flush=True)
    h = capi.Hydro(d)
    h.upload(st)
    o = oracle_ctypes.Oracle(dref)
    o.set_state(st)
    assert h.step(tau, ts) == o.step(tau, ts) == OK
    g, c = h.download(), o.state()
    errs = metrics.state_errors(d, g, c)
    print(errs)
    assert metrics.worst(errs) <= 1e-10, errs
    assert np.array_equal(h.flags() & 0x7, o.flags() & 0x7)


def test_fast_mode_100_steps_gubser_within_1e8():
    """Config 1 (64x64x8 Gubser, RK2, dtau=0.01) for 100 steps."""
    d = make_desc(64, 64, 8, fast_math=1)
    dref = d.copy()
    dref.fast_math = 0
    tau, ts = 1.0, 0.01
    st = ic.gubser(d, tau, ts)
    print("case fast 100 steps", flush=True)
    h = capi.Hydro(d)
    h.upload(st)
    o = oracle_ctypes.Oracle(dref)
    o.set_state(st)
    for s in range(100):
        t = tau + s * ts
        assert h.step(t, ts) == o.step(t, ts) == OK
    g, c = h.download(), o.state()
    errs = metrics.state_errors(d, g, c)
    print(errs)
    assert metrics.worst(errs) <= 1e-8, errs


def test_fast_mode_s95p_table_within_1e10():
    """Fast mode with the tabulated s95p EOS against the closed-form oracle."""
    d = make_desc(30, 28, 9, eos=EOS_S95P, fast_math=1, eos_table=1, eta_law=ETA_S_OVER_4PI, eta_coef=1.0)
    dref = d.copy()
    dref.fast_math = 0
    dref.eos_table = 0
    tau, ts = 0.6, 0.002
    eosp = lambda e: oracle_ctypes.lib().or_eos_p(dref, e)
    st = ic.random_smooth(d, seed=51, tau=tau, e_scale=60.0, eos_p=eosp)
    print("case fast s95p table", flush=True)
    h = capi.Hydro(d)
    h.upload(st)
    o = oracle_ctypes.Oracle(dref)
    o.set_state(st)
    assert h.step(tau, ts) == o.step(tau, ts) == OK
    errs = metrics.state_errors(d, h.download(), o.state())
    print(errs)
    assert metrics.worst(errs) <= 1e-10, errs


@pytest.mark.parametrize("tiles", [((32, 16), (64, 8), (16, 16)), ((64, 16), (128, 8), (23, 4)), ((256, 4), (16, 64), (8, 128))])
def test_fast_flux_tile_shapes_agree(tiles):
    """The LDS tile shape is a launch parameter only: every shape gives the
    same fluxes (bitwise) on a grid that is not a multiple of any tile."""
    d = make_desc(37, 29, 17, fast_math=1)
    tau, ts = 0.9, 0.005
    st = ic.random_smooth(d, seed=61, tau=tau)
    print("case tiles", tiles, flush=True)
    outs = []
    for tl in [None, tiles]:
        h = capi.Hydro(d)
        h.upload(st)
        if tl:
            if max(tp * tk for tp, tk in tl) > 512:
                h.set_launch(flux_nt=[512, 512, 512])  # 1024-slot tiles need 512-thread blocks
            for dd, (tp, tk) in enumerate(tl):
                h.set_tile(dd, tp, tk)
        assert h.step(tau, ts) == OK
        outs.append(h.download())
    for f in range(NFIELDS):
        assert same_bits(outs[0][f], outs[1][f]), FIELD_NAMES[f]


def test_autotune_keeps_results():
    d = make_desc(40, 30, 12, fast_math=1)
    st = ic.random_smooth(d, seed=62, tau=1.0)
    h = capi.Hydro(d)
    h.upload(st)
    tiles = h.autotune(reps=1)
    assert all(tp * tk <= 1024 for tp, tk in tiles)
    h.step(1.0, 0.005)
    h2 = capi.Hydro(d)
    h2.upload(st)
    h2.step(1.0, 0.005)
    a, b = h.download(), h2.download()
    for f in range(NFIELDS):
        assert same_bits(a[f], b[f]), FIELD_NAMES[f]
