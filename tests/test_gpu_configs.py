"""GPU: the benchmark configurations against the CPU oracle, at sizes the
oracle finishes in seconds (the OpenMP build of the oracle, bitwise equal to
the serial one, runs the long references).

- Glauber grid (s95p EOS, b = 7 fm, eta/s = 0.08, tau_pi = 5 eta/(e+p), RK3,
  dtau = 0.002): 40x40x16 cells of 0.3 fm / 0.3 in eta, so the grid spans the
  event; exact mode bitwise for 3 steps (state, Newton counts, flags); fast
  mode with the s95p table within 1e-8 of the exact oracle after 100 steps,
  with the Newton iteration counts' agreement rate reported (fast arithmetic
  rounds differently, so counts may differ by an iteration; the flags that
  stop the reference must not).
- Hot-spot ensemble (200 spots of width 0.4 fm from two Woods-Saxon nuclei,
  gamma weights, b ~ U(0,12) fm, s95p, eta/s = 0.08, RK3): five uneven events
  of side 24..40 and 8..16 eta cells in one batched, graph-captured launch,
  every event bitwise equal to the oracle stepping it alone (exact mode), and
  within 1e-9 after 20 steps in the autotuned fast mode.
- Adjoint (smooth Glauber b = 8 fm, s95p table, RK2): the gradient of the
  final momentum anisotropy along a random direction of the initial state,
  along the initial e-field (through T^{tau tau} = e and P(e)) and with
  respect to the eta/s and tau_pi coefficients against central differences:
  to 1e-6 on 64x64x16 over 50 steps, where the differences converge; within
  the differences' own scatter on the full 128x128x32 grid over 200 steps
  with checkpoints every 20, where they do not (see that test).
"""
import numpy as np
import pytest
import torch

from viral_amd import capi, ic, autograd
from viral_amd.desc import (make_desc, NFIELDS, FIELD_NAMES, EOS_S95P, ETA_TIMES_S, TPI_OVER_EP, F_EN, F_P, F_T00,
                            F_TEMP, F_VX, F_U0, F_PU0, F_PI0, F_BULK, FLAG_ACAUSAL, FLAG_ENEG, FLAG_SOLVER,
                            FLAG_FREEZEOUT, OK)
import metrics
import oracle_ctypes

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(nan | (a.view(np.int64) == b.view(np.int64))))


def closed_form_p(d):
    dref = d.copy()
    dref.fast_math = dref.eos_table = 0
    return lambda e: capi.eos_eval(dref, e)[0].reshape(np.shape(e))


# ------------------------------------------------------------------ Glauber grid

def glauber_desc(fast):
    return make_desc(40, 40, 16, dx=0.3, dy=0.3, deta=0.3, rk=3, eos=EOS_S95P, eta_law=ETA_TIMES_S, eta_coef=0.08,
                     tpi_law=TPI_OVER_EP, tpi_coef=5.0, fast_math=fast, eos_table=fast)


def test_glauber_grid_exact_bitwise():
    d = glauber_desc(0)
    tau, ts = 0.6, 0.002
    st = ic.glauber(d, tau, ts, b=7.0, eos_p=closed_form_p(d))
    print("case glauber 40x40x16 exact", flush=True)
    h = capi.Hydro(d)
    h.upload(st)
    o = oracle_ctypes.Oracle(d, oracle_ctypes.fast_lib())
    o.set_state(st)
    for s in range(3):
        assert h.step(tau + s * ts, ts) == o.step(tau + s * ts, ts) == OK
        for q in range(3):
            assert np.array_equal(h.iterations(q), o.iters(q)), (s, q)
        assert np.array_equal(h.flags(), o.flags()), s
    g, c = h.download(), o.state()
    for f in range(NFIELDS):
        assert same_bits(g[f], c[f]), FIELD_NAMES[f]


def test_glauber_grid_fast_table_100_steps_within_1e8():
    d = glauber_desc(1)
    dref = glauber_desc(0)
    tau, ts = 0.6, 0.002
    st = ic.glauber(d, tau, ts, b=7.0, eos_p=closed_form_p(d))
    print("case glauber 40x40x16 fast 100 steps", flush=True)
    h = capi.Hydro(d)
    h.upload(st)
    h.autotune(reps=1)
    h.upload(st)
    o = oracle_ctypes.Oracle(dref, oracle_ctypes.fast_lib())
    o.set_state(st)
    agree, total = 0, 0
    stop = FLAG_ACAUSAL | FLAG_ENEG | FLAG_SOLVER
    interior = metrics.interior_mask(d)
    for s in range(100):
        t = tau + s * ts
        assert h.step(t, ts) == o.step(t, ts) == OK
        for q in range(3):
            a, b = h.iterations(q)[interior], o.iters(q)[interior]
            agree += int(np.count_nonzero(a == b))
            total += a.size
        fg, fo = h.flags(), o.flags()
        assert np.array_equal(fg & stop, fo & stop), s
        assert np.array_equal(fg & FLAG_FREEZEOUT, fo & FLAG_FREEZEOUT), s
    errs = metrics.state_errors(d, h.download(), o.state())
    print(errs)
    print("Newton iteration counts equal in %.4f of cell-stages (fast vs exact)" % (agree / total))
    assert metrics.worst(errs) <= 1e-8, errs
    assert agree / total > 0.5


# ------------------------------------------------------------------ hot-spot ensemble

def hotspot_ensemble(fast, n=5, seed=21):
    rng = np.random.default_rng(seed)
    descs, states = [], []
    for q in range(n):
        side = int(rng.integers(24, 41))
        neta = int(rng.integers(8, 17))
        d = make_desc(side, side, neta, rk=3, eos=EOS_S95P, eta_law=ETA_TIMES_S, eta_coef=0.08, tpi_law=TPI_OVER_EP,
                      tpi_coef=5.0, fast_math=fast, eos_table=fast)
        descs.append(d)
        states.append(ic.hotspots(d, 0.6, 0.002, seed=500 + q, b=float(rng.uniform(0.0, 12.0)),
                                  eos_p=closed_form_p(d)))
    return descs, states


def run_batched(descs, states, nsteps, autotune=False):
    h = capi.Hydro(descs)
    h.set_launch(batched=1, graph=1)
    for ev, st in enumerate(states):
        h.upload(st, ev=ev)
    if autotune:
        assert h.step(0.6, 0.002) == OK
        h.autotune(reps=1)
        for ev, st in enumerate(states):
            h.upload(st, ev=ev)
    r, done = h.steps(0.6, 0.002, nsteps)
    assert (r, done) == (OK, nsteps)
    return h


def oracle_alone(d, st, nsteps):
    dref = d.copy()
    dref.fast_math = dref.eos_table = 0
    o = oracle_ctypes.Oracle(dref, oracle_ctypes.fast_lib())
    o.set_state(st)
    for s in range(nsteps):
        assert o.step(0.6 + s * 0.002, 0.002) == OK
    return o


def test_hotspot_ensemble_batched_graph_exact_bitwise():
    descs, states = hotspot_ensemble(0)
    print("case hot-spot ensemble exact", [(d.nx, d.nz) for d in descs], flush=True)
    h = run_batched(descs, states, 3)
    for ev, (d, st) in enumerate(zip(descs, states)):
        o = oracle_alone(d, st, 3)
        g, c = h.download(ev=ev), o.state()
        for f in range(NFIELDS):
            assert same_bits(g[f], c[f]), (ev, FIELD_NAMES[f])
        for q in range(3):
            assert np.array_equal(h.iterations(q, ev=ev), o.iters(q)), (ev, q)
        assert np.array_equal(h.flags(ev=ev), o.flags()), ev


def test_hotspot_ensemble_batched_fast_20_steps():
    descs, states = hotspot_ensemble(1)
    print("case hot-spot ensemble fast", flush=True)
    h = run_batched(descs, states, 20, autotune=True)
    print("launch", h.launch().as_dict())
    for ev, (d, st) in enumerate(zip(descs, states)):
        o = oracle_alone(d, st, 20)
        errs = metrics.state_errors(d, h.download(ev=ev), o.state())
        print(ev, errs)
        assert metrics.worst(errs) <= 1e-9, (ev, errs)


# ------------------------------------------------------------------ adjoint

TAU0, TS = 0.6, 0.002


def adjoint_case(n=(128, 128, 32)):
    d = make_desc(*n, rk=2, eos=EOS_S95P, eos_table=1, fast_math=1, eta_coef=0.08, tpi_coef=5.0)
    st = ic.glauber(d, TAU0, TS, b=8.0, eos_p=closed_form_p(d))
    return d, st


def directions(d, st, rng):
    """a random direction over the fields the step reads, scaled cell by cell
    (the event spans 1e-7 of its peak energy density: a perturbation scaled
    by a field's maximum would swamp the dilute cells): nonzero fields
    relative to themselves, the stresses and T^{tau i} to 1 % of the cell's
    e, u^i and prevu^i to 1 % of u^0; and the initial e-field of the fluid at
    rest, through T^{tau tau} = e and P = P(e)"""
    dx = np.zeros_like(st)
    for f in range(NFIELDS):
        if f == F_TEMP or F_VX <= f < F_VX + 3:
            continue
        if F_T00 < f < F_T00 + 4 or F_PI0 <= f <= F_BULK:
            base = 0.01 * st[F_EN]
        elif F_U0 < f < F_U0 + 4 or F_PU0 < f < F_PU0 + 4:
            base = 0.01 * st[F_U0]
        else:
            base = np.abs(st[f])
        dx[f] = rng.standard_normal(st.shape[1]) * base
    de = rng.standard_normal(st.shape[1]) * st[F_EN]
    dref = d.copy()
    dref.fast_math = dref.eos_table = 0
    ddx = np.zeros_like(st)
    ddx[F_EN] = de
    ddx[F_T00] = de
    ddx[F_P] = capi.eos_eval(dref, st[F_EN])[1] * de
    return [("random state direction", dx), ("initial e-field", ddx)]


def adjoint_vs_fd(d, st, nsteps, every, hs):
    """(name, adjoint, [central differences at each h]) for the state
    directions and the two transport coefficients, objective eps_2 at the
    final time"""
    h = capi.Hydro(d)
    h.upload(st)
    h.adjoint_begin(every)
    r, done = h.steps(TAU0, TS, nsteps)
    assert (r, done) == (OK, nsteps)
    tf = TAU0 + nsteps * TS
    fin = torch.from_numpy(h.download()).requires_grad_(True)
    autograd.momentum_anisotropy(d, fin, tf).backward()
    grad, pg = h.adjoint(fin.grad.numpy())
    h.adjoint_end()
    h.close()
    fwd = capi.Hydro(d)

    def eps2(dd, x):
        hh = fwd if dd is d else capi.Hydro(dd)
        hh.upload(x)
        r, done = hh.steps(TAU0, TS, nsteps)
        assert (r, done) == (OK, nsteps)
        return float(autograd.momentum_anisotropy(dd, torch.from_numpy(hh.download()), tf))

    out = []
    for name, vec in directions(d, st, np.random.default_rng(31)):
        fds = [(eps2(d, st + hh * vec) - eps2(d, st - hh * vec)) / (2 * hh) for hh in hs]
        out.append((name, float(np.sum(grad * vec)), fds))
    for q, name in enumerate(["eta_coef", "tpi_coef"]):
        c0 = getattr(d, name)
        fds = []
        for hh in hs:
            vals = []
            for sgn in (1, -1):
                dd = d.copy()
                setattr(dd, name, c0 * (1 + sgn * hh))
                vals.append(eps2(dd, st))
            fds.append((vals[0] - vals[1]) / (2 * hh * c0))
        out.append((name, float(pg[q]), fds))
    for name, ad, fds in out:
        print("%-24s adjoint % .10e  fd %s" % (name, ad, "  ".join("% .10e" % v for v in fds)), flush=True)
    return out


def test_adjoint_matches_fd_to_1e6_on_64x64x16():
    """the C5 event's physics (smooth Glauber b = 8 fm, s95p table, eta/s =
    0.08, tau_pi = 5 eta/(e+p), RK2) on a 64x64x16 grid, 50 steps,
    checkpoints every 10: central differences at h = 1e-7 (relative) are
    converged to ~1e-8 here, and the adjoint must match them to 1e-6 along
    every direction and for both coefficients"""
    d, st = adjoint_case((64, 64, 16))
    print("case adjoint 64x64x16 50 steps every 10", flush=True)
    for name, ad, (fd,) in adjoint_vs_fd(d, st, 50, 10, [1e-7]):
        assert abs(ad - fd) <= 1e-6 * abs(fd), (name, ad, fd)


def test_adjoint_c5_128x128x32_200_steps_within_fd_scatter():
    """The C5 configuration: 128x128x32, 200 RK2 steps, checkpoints every 20.
    Here central differences do not converge below ~1e-5: the fluid starts at
    rest, so the Rusanov face speeds |u^i/u^0| sit at the kink of |.| on the
    symmetry planes and every perturbation moves cells across it (the
    difference quotient's error falls like h, not h^2, until rounding takes
    over).  The test prints the differences at h = 1e-6, 1e-7, 3e-8 and
    requires the adjoint to lie within twice the spread of the two smallest
    steps around their mean (or 1e-6 of it, whichever is larger)."""
    d, st = adjoint_case()
    print("case adjoint 128x128x32 200 steps every 20", flush=True)
    for name, ad, fds in adjoint_vs_fd(d, st, 200, 20, [1e-6, 1e-7, 3e-8]):
        fd, scatter = 0.5 * (fds[1] + fds[2]), abs(fds[1] - fds[2])
        assert abs(ad - fd) <= max(1e-6 * abs(fd), 2 * scatter), (name, ad, fds)


# ------------------------------------------------------------------ SHASTA on hot spots

def shasta_hotspot_desc(n, neta, fast, fullstep=True):
    from viral_amd.desc import SHASTA, Q_DEFAULT, Q_SHASTA_FULLSTEP
    return make_desc(n, n, neta, rk=2, scheme=SHASTA, eos=EOS_S95P, eos_table=fast, fast_math=fast,
                     eta_law=ETA_TIMES_S, eta_coef=0.12, tpi_law=TPI_OVER_EP, tpi_coef=5.0,
                     quirks=Q_DEFAULT | (Q_SHASTA_FULLSTEP if fullstep else 0))


@pytest.mark.parametrize("rk", [2, 3])
def test_shasta_fullstep_hotspots_exact_bitwise(rk):
    """SHASTA with the full-step sub-step on hot spots (200 spots, width 0.4 fm,
    seed 3): exact mode bitwise against the oracle for 3 steps"""
    d = shasta_hotspot_desc(48, 20, 0)
    d.rk = rk
    st = ic.hotspots(d, 0.6, 0.002, seed=3, eos_p=closed_form_p(d))
    print("case shasta fullstep hot spots exact rk=%d" % rk, flush=True)
    h = capi.Hydro(d)
    h.upload(st)
    o = oracle_ctypes.Oracle(d, oracle_ctypes.fast_lib())
    o.set_state(st)
    for s in range(3):
        assert h.step(0.6 + s * 0.002, 0.002) == o.step(0.6 + s * 0.002, 0.002) == OK
    g, c = h.download(), o.state()
    for f in range(NFIELDS):
        assert same_bits(g[f], c[f]), FIELD_NAMES[f]
    for q in range(rk):
        assert np.array_equal(h.iterations(q), o.iters(q))
    assert np.array_equal(h.flags(), o.flags())


def test_shasta_fullstep_hotspots_400_steps():
    """the fluctuating-event configuration's 400 SHASTA steps (s95p table,
    eta/s = 0.12, RK2) on a 96x96x32 hot-spot grid with the full-step
    sub-step: every step succeeds and the state stays finite and positive
    (the reference's dtau/2 sub-step fails within a few steps on such states,
    tests/test_oracle_shasta_stability.py)"""
    d = shasta_hotspot_desc(96, 32, 1)
    st = ic.hotspots(d, 0.6, 0.002, seed=3, eos_p=closed_form_p(d))
    print("case shasta fullstep hot spots 400 steps", flush=True)
    h = capi.Hydro(d)
    h.upload(st)
    r, done = h.steps(0.6, 0.002, 400)
    assert (r, done) == (OK, 400), h.first_bad_cell()
    fin = h.download()
    I = metrics.interior_mask(d)
    assert np.all(np.isfinite(fin[:, I])) and fin[F_EN][I].min() > 0
