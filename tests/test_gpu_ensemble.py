"""GPU: ensembles of uneven events in one context, through the launch layer.

Every launch configuration (one launch per event or one batched launch over
all events, hipGraph replay, persistent blocks, XCD tile order, fused
sources, flux block sizes) must give each event exactly what the CPU oracle
gives that event alone: the batching only changes which block computes a
cell, never the arithmetic.
"""
import numpy as np
import pytest

from viral_amd import capi, ic
from viral_amd.desc import (make_desc, NFIELDS, FIELD_NAMES, EOS_S95P, ETA_S_OVER_4PI, TPI_CONST, F_T10, F_EN,
                            FLAG_ACAUSAL, FLAG_ENEG, FLAG_SOLVER, OK, ERR_PHYSICS)
import metrics
import oracle_ctypes

pytestmark = pytest.mark.gpu

TAU, TS = 0.8, 0.004


def same_bits(a, b):
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(nan | (a.view(np.int64) == b.view(np.int64))))


def ensemble(fast, rk=2):
    """three uneven events: conformal 3-D, s95p 3-D, boost-invariant"""
    descs = [make_desc(12, 10, 6, rk=rk, fast_math=fast),
             make_desc(9, 14, 5, rk=rk, fast_math=fast, eos=EOS_S95P, eos_table=fast, eta_law=ETA_S_OVER_4PI,
                       eta_coef=1.0),
             make_desc(8, 8, 1, rk=rk, fast_math=fast, lbi=1, tpi_law=TPI_CONST, tpi_coef=0.3)]
    states = []
    for q, d in enumerate(descs):
        if d.eos == EOS_S95P:
            dref = d.copy()
            dref.fast_math = dref.eos_table = 0
            eosp = (lambda dd: (lambda e: oracle_ctypes.lib().or_eos_p(dd, e)))(dref)
            states.append(ic.random_smooth(d, seed=70 + q, tau=TAU, e_scale=30.0, eos_p=eosp))
        else:
            states.append(ic.random_smooth(d, seed=70 + q, tau=TAU))
    return descs, states


def run_ensemble(descs, states, cfg, nsteps, per_step=False, tiles=None):
    print("case ensemble cfg=%r steps=%d per_step=%d" % (cfg, nsteps, per_step), flush=True)
    h = capi.Hydro(descs)
    if tiles:
        for ev in range(len(descs)):
            for d, (tp, tk) in enumerate(tiles):
                h.set_tile(d, tp, tk, ev=ev)
    if cfg:
        h.set_launch(**cfg)
    for ev, st in enumerate(states):
        h.upload(st, ev=ev)
    if per_step:
        t = TAU
        for s in range(nsteps):
            assert h.step(t, TS) == OK
            t = t + TS
    else:
        r, done = h.steps(TAU, TS, nsteps)
        assert (r, done) == (OK, nsteps)
    return h


def oracle_run(d, st, nsteps):
    dref = d.copy()
    dref.fast_math = dref.eos_table = 0
    o = oracle_ctypes.Oracle(dref)
    o.set_state(st)
    t = TAU
    for s in range(nsteps):
        assert o.step(t, TS) == OK
        t = t + TS
    return o


CONFIGS = [
    dict(batched=0, graph=0),
    dict(batched=1, graph=0),
    dict(batched=1, graph=1),
    dict(batched=1, graph=1, blocks_per_cu=2, xcd=1),
    dict(batched=1, graph=0, fusion=1),
    dict(batched=0, graph=1, fusion=1, blocks_per_cu=1, xcd=1),
]


@pytest.mark.parametrize("cfg", CONFIGS)
def test_ensemble_exact_matches_oracle_per_event(cfg):
    descs, states = ensemble(fast=0)
    h = run_ensemble(descs, states, cfg, 3)
    for ev, (d, st) in enumerate(zip(descs, states)):
        o = oracle_run(d, st, 3)
        g, c = h.download(ev), o.state()
        for f in range(NFIELDS):
            assert same_bits(g[f], c[f]), "event %d field %s" % (ev, FIELD_NAMES[f])
        for s in range(2):
            assert np.array_equal(h.iterations(s, ev), o.iters(s)), (ev, s)
        assert np.array_equal(h.flags(ev), o.flags())


def test_ensemble_rk3_graph_odd_parity_exact():
    """RK3 flips the state copy every step: the graph of each parity."""
    descs, states = ensemble(fast=0, rk=3)
    h = run_ensemble(descs, states, dict(batched=1, graph=1), 3)
    for ev, (d, st) in enumerate(zip(descs, states)):
        o = oracle_run(d, st, 3)
        g, c = h.download(ev), o.state()
        for f in range(NFIELDS):
            assert same_bits(g[f], c[f]), "event %d field %s" % (ev, FIELD_NAMES[f])


FAST_CONFIGS = CONFIGS + [
    dict(batched=1, graph=1, flux_nt=[64, 64, 64]),
    dict(batched=1, graph=1, flux_nt=[128, 128, 128], blocks_per_cu=4),
    dict(batched=1, graph=1, flux_nt=[512, 512, 512], xcd=1),
]


def test_ensemble_fast_configs_agree_and_meet_tolerance():
    descs, states = ensemble(fast=1)
    ref = run_ensemble(descs, states, dict(batched=0, graph=0), 2, per_step=True, tiles=[(32, 8)] * 3)
    base = [ref.download(ev) for ev in range(len(descs))]
    for ev, (d, st) in enumerate(zip(descs, states)):
        errs = metrics.state_errors(d, base[ev], oracle_run(d, st, 2).state())
        assert metrics.worst(errs) <= 1e-10, (ev, errs)
    for cfg in FAST_CONFIGS:
        h = run_ensemble(descs, states, cfg, 2, tiles=[(32, 8)] * 3)
        for ev in range(len(descs)):
            g = h.download(ev)
            for f in range(NFIELDS):
                assert same_bits(g[f], base[ev][f]), "cfg %r event %d field %s" % (cfg, ev, FIELD_NAMES[f])


def test_ensemble_autotune_keeps_results():
    descs, states = ensemble(fast=1)
    a = run_ensemble(descs, states, dict(batched=1, graph=1), 1)
    tiles = a.autotune(reps=1)
    cfg = a.launch().as_dict()
    print("tuned", tiles, cfg, flush=True)
    r, done = a.steps(TAU + TS, TS, 1)
    assert r == OK
    b = run_ensemble(descs, states, dict(batched=0, graph=0), 2, per_step=True)
    for ev in range(len(descs)):
        x, y = a.download(ev), b.download(ev)
        for f in range(NFIELDS):
            assert same_bits(x[f], y[f]), (ev, FIELD_NAMES[f])


def test_ensemble_failure_names_event_step_and_cell():
    """One event fails in the first step (a cell pushed towards v >= 1 ends
    with e <= 0 or v >= 1): the batched graph run
    reports zero completed steps and that event's first bad cell, exactly as
    the oracle run of that event alone; the healthy events report none."""
    descs, states = ensemble(fast=0)
    d = descs[1]
    gx, gy, gz = d.ghost_dims()
    for (i, j, k) in [(5, 7, 4), (9, 4, 6)]:
        x = (i * gy + j) * gz + k
        states[1][F_T10][x] = 50.0 * states[1][F_EN][x]
    print("case ensemble failure", flush=True)
    h = capi.Hydro(descs)
    h.set_launch(batched=1, graph=1)
    for ev, st in enumerate(states):
        h.upload(st, ev=ev)
    r, done = h.steps(TAU, TS, 3)
    assert (r, done) == (ERR_PHYSICS, 0)
    dref = d.copy()
    o = oracle_ctypes.Oracle(dref)
    o.set_state(states[1])
    assert o.step(TAU, TS) == ERR_PHYSICS
    bad = h.first_bad_cell(1)
    assert bad == o.first_bad() and bad[4] & (FLAG_ACAUSAL | FLAG_ENEG | FLAG_SOLVER)
    assert h.first_bad_cell(0) is None and h.first_bad_cell(2) is None


def test_tile_table_lists_every_event():
    descs, states = ensemble(fast=1)
    h = capi.Hydro(descs)
    for ph in range(6):
        t = h.tile_table(ph)
        tp = [h.get_tile(min(ph, 2), ev) for ev in range(len(descs))]
        counts = [capi.tile_count(d, ph, *tp[ev]) for ev, d in enumerate(descs)]
        assert t == [0] + list(np.cumsum(counts)), (ph, t, counts)
