"""GPU: the hand-written adjoint against central finite differences.

J = <seed, final state> after a few steps; the adjoint's dJ/d(initial state)
and dJ/d(eta_coef, tpi_coef) are checked along random directions against
four-point central differences (fourth order, steps h and 2h, h = 1e-6
relative: at 1e-5 and above, kinks of the Rusanov speeds |u^i/u^0| and the
strongly nonlinear WENO weights within reach of the perturbation spoil them,
while at 1e-6 they agree with the adjoint to ~1e-10) of the GPU forward
and of the CPU oracle's forward (the reference gradient), to 1e-6 relative;
the torch.autograd harness is checked on the momentum anisotropy of a
Glauber-like event.
"""
import numpy as np
import pytest
import torch

from viral_amd import capi, ic, autograd
from viral_amd.desc import (make_desc, NFIELDS, F_T00, F_PI0, F_BULK, F_EN, F_P, F_U0, F_PU0, F_TEMP, F_VX,
                            ETA_S_OVER_4PI, TPI_OVER_P, EOS_S95P, OK)
import oracle_ctypes

pytestmark = pytest.mark.gpu

TAU, TS = 1.0, 0.005
# fields the step reads from its input state (Temp and the 3-velocity are outputs only)
READ = [f for f in range(NFIELDS) if f != F_TEMP and not (F_VX <= f < F_VX + 3)]
H = 1e-6   # relative step of the differences along a state direction
# relative step for the transport coefficients: their derivatives are small
# against J, so differences at 1e-6 carry ~1e-6 of rounding noise; at 3e-4 the
# four-point difference agrees with the adjoint to ~1e-8
HP = 3e-4
BAR = 1e-6


def fd4(J, h=H):
    """dJ/dt at t = 0 from J(+-h), J(+-2h): (8 (J(h) - J(-h)) - (J(2h) - J(-2h))) / (12 h)"""
    return (8.0 * (J(h) - J(-h)) - (J(2 * h) - J(-2 * h))) / (12.0 * h)


def close(ad, fd, bar=BAR):
    return abs(ad - fd) <= bar * abs(fd)


def run_gpu(d, st, nsteps, every=None, seed=None):
    h = capi.Hydro(d)
    h.upload(st)
    if every is not None:
        h.adjoint_begin(every)
    r, done = h.steps(TAU, TS, nsteps)
    assert (r, done) == (OK, nsteps)
    out = h.download()
    if seed is None:
        return out
    grad, pg = h.adjoint(seed)
    after = h.download()
    assert np.array_equal(after, out), "the adjoint must leave the state as it was"
    return out, grad, pg


def run_cpu(d, st, nsteps):
    o = oracle_ctypes.Oracle(d)
    o.set_state(st)
    t = TAU
    for _ in range(nsteps):
        assert o.step(t, TS) == OK
        t = t + TS
    return o.state()


def direction(st, rng):
    dx = np.zeros_like(st)
    for f in READ:
        scale = np.abs(st[f]).max()
        dx[f] = rng.standard_normal(st.shape[1]) * (scale if scale > 0 else 1e-3)
    return dx


def small_case(rk=2, fast=0, **kw):
    d = make_desc(10, 9, 7, rk=rk, fast_math=fast, **kw)
    if d.eos == EOS_S95P:
        dref = d.copy()
        dref.fast_math = dref.eos_table = 0
        st = ic.random_smooth(d, seed=3, tau=TAU, pi_frac=0.02, e_scale=30.0,
                              eos_p=lambda e: capi.eos_eval(dref, e)[0].reshape(np.shape(e)))
    else:
        st = ic.random_smooth(d, seed=3, tau=TAU, pi_frac=0.02)
    return d, st


@pytest.mark.parametrize("rk,every", [(1, 1), (2, 2), (3, 2)])
def test_adjoint_state_gradient_matches_fd(rk, every):
    d, st = small_case(rk=rk)
    rng = np.random.default_rng(rk)
    seed = rng.standard_normal(st.shape)
    print("case adjoint rk=%d every=%d" % (rk, every), flush=True)
    out, grad, pg = run_gpu(d, st, 3, every, seed)
    for trial in range(3):
        dx = direction(st, rng)
        fd = fd4(lambda h: np.sum(seed * run_gpu(d, st + h * dx, 3)))
        ad = np.sum(grad * dx)
        print(trial, ad, fd, abs(ad - fd) / abs(fd), flush=True)
        assert close(ad, fd), (ad, fd)


def test_adjoint_parameter_gradients_match_fd_and_cpu_reference():
    d, st = small_case(eta_law=ETA_S_OVER_4PI, eta_coef=2.0, tpi_law=TPI_OVER_P, tpi_coef=1.5)
    rng = np.random.default_rng(9)
    seed = rng.standard_normal(st.shape)
    print("case adjoint parameters", flush=True)
    out, grad, pg = run_gpu(d, st, 3, 2, seed)
    for q, name in enumerate(["eta_coef", "tpi_coef"]):
        c0 = getattr(d, name)

        def J(h, run):
            dd = d.copy()
            setattr(dd, name, c0 * (1 + h))
            return np.sum(seed * run(dd, st, 3))

        fd_gpu = fd4(lambda h: J(h, run_gpu), HP) / c0
        fd_cpu = fd4(lambda h: J(h, run_cpu), HP) / c0
        print(name, pg[q], fd_gpu, fd_cpu, flush=True)
        assert close(pg[q], fd_gpu), (name, pg[q], fd_gpu)
        assert close(pg[q], fd_cpu), (name, pg[q], fd_cpu)
    # the state gradient against the CPU oracle's finite differences
    dx = direction(st, rng)
    fd = fd4(lambda h: np.sum(seed * run_cpu(d, st + h * dx, 3)))
    ad = np.sum(grad * dx)
    print("state vs oracle", ad, fd, flush=True)
    assert close(ad, fd), (ad, fd)


def test_gubser_profile_adjoint_matches_fd():
    """the boost-invariant grid with the GUBSER profile's mirror boundary"""
    from viral_amd.desc import Q_DEFAULT, Q_GUBSER_MIRROR, ETA_TIMES_S, TPI_OVER_EP
    d = make_desc(12, 11, 1, lbi=1, quirks=Q_DEFAULT | Q_GUBSER_MIRROR, eta_law=ETA_TIMES_S, eta_coef=0.2,
                  tpi_law=TPI_OVER_EP, tpi_coef=5.0)
    st = ic.random_smooth(d, seed=13, tau=TAU, pi_frac=0.02)
    rng = np.random.default_rng(14)
    seed = rng.standard_normal(st.shape)
    print("case adjoint gubser mirror", flush=True)
    _, grad, pg = run_gpu(d, st, 3, 2, seed)
    for trial in range(2):
        dx = direction(st, rng)
        fd = fd4(lambda h: np.sum(seed * run_gpu(d, st + h * dx, 3)))
        ad = np.sum(grad * dx)
        print(trial, ad, fd, flush=True)
        assert close(ad, fd), (ad, fd)


def test_checkpoint_interval_does_not_change_the_gradient():
    d, st = small_case()
    seed = np.random.default_rng(4).standard_normal(st.shape)
    print("case adjoint checkpoints", flush=True)
    _, g1, p1 = run_gpu(d, st, 4, 1, seed)
    _, g3, p3 = run_gpu(d, st, 4, 3, seed)
    assert np.array_equal(g1, g3) and np.array_equal(p1, p3)


@pytest.mark.parametrize("rk,fast,variant", [(2, 0, "default"), (3, 1, "default"), (1, 0, "default"),
                                             (2, 0, "gubser_lbi"), (3, 0, "no_k0")])
def test_stored_stage_states_match_recompute(rk, fast, variant):
    """every == 0 keeps each stage's state instead of recomputing it: the
    forward state (every cell, two viral_steps calls) equals the unrecorded
    run's, and the gradient equals the checkpointed adjoint's, bit for bit;
    also on the boost-invariant grid with the GUBSER mirror (no eta ghosts in
    the stage slabs) and with the k0 broadcast off"""
    from viral_amd.desc import Q_DEFAULT, Q_GUBSER_MIRROR, Q_K0_BROADCAST, ETA_TIMES_S, TPI_OVER_EP
    if variant == "gubser_lbi":
        d = make_desc(12, 11, 1, lbi=1, rk=rk, fast_math=fast, quirks=Q_DEFAULT | Q_GUBSER_MIRROR,
                      eta_law=ETA_TIMES_S, eta_coef=0.2, tpi_law=TPI_OVER_EP, tpi_coef=5.0)
        st = ic.random_smooth(d, seed=13, tau=TAU, pi_frac=0.02)
    elif variant == "no_k0":
        d, st = small_case(rk=rk, fast=fast, quirks=Q_DEFAULT & ~Q_K0_BROADCAST)
    else:
        d, st = small_case(rk=rk, fast=fast)
    seed = np.random.default_rng(5).standard_normal(st.shape)
    print("case adjoint stored rk=%d fast=%d %s" % (rk, fast, variant), flush=True)
    plain = run_gpu(d, st, 4)
    _, g3, p3 = run_gpu(d, st, 4, 3, seed)
    h = capi.Hydro(d)
    h.upload(st)
    h.adjoint_begin(0)
    assert h.steps(TAU, TS, 1) == (OK, 1)
    assert h.steps(TAU + TS, TS, 3) == (OK, 3)
    out = h.download()
    assert np.array_equal(out.view(np.int64), plain.view(np.int64))
    g0, p0 = h.adjoint(seed)
    assert np.array_equal(h.download().view(np.int64), plain.view(np.int64))
    assert np.array_equal(g0, g3) and np.array_equal(p0, p3)
    h.adjoint_end()


@pytest.mark.parametrize("every,calls", [(3, (1, 3, 3)), (2, (4, 3)), (7, (2, 3, 2)), (1, (3, 2)), (4, (4,))])
def test_checkpoint_recording_across_calls(every, calls):
    """With checkpoints the recording keeps the current segment's stage
    states in a ring, and the adjoint reverses the last segment from it: over
    steps recorded in several viral_steps calls, with segment ends inside and
    at the end of calls, the forward state (every cell) equals the unrecorded
    run's, and two adjoint calls on one recording both give the stored-stage
    gradient bit for bit"""
    d, st = small_case(rk=2, fast=1)
    seed = np.random.default_rng(8).standard_normal(st.shape)
    M = sum(calls)
    print("case adjoint ring every=%d calls=%s" % (every, calls), flush=True)
    plain = run_gpu(d, st, M)
    _, g0, p0 = run_gpu(d, st, M, 0, seed)
    h = capi.Hydro(d)
    h.upload(st)
    h.adjoint_begin(every)
    t = TAU
    for c in calls:
        assert h.steps(t, TS, c) == (OK, c)
        for _ in range(c):
            t = t + TS
    assert np.array_equal(h.download().view(np.int64), plain.view(np.int64))
    for _ in range(2):
        g, pg = h.adjoint(seed)
        assert np.array_equal(h.download().view(np.int64), plain.view(np.int64))
        assert np.array_equal(g, g0) and np.array_equal(pg, p0)
    h.adjoint_end()


def test_fast_mode_adjoint_matches_fd():
    d, st = small_case(fast=1)
    rng = np.random.default_rng(5)
    seed = rng.standard_normal(st.shape)
    print("case adjoint fast", flush=True)
    _, grad, pg = run_gpu(d, st, 2, 2, seed)
    dx = direction(st, rng)
    fd = fd4(lambda h: np.sum(seed * run_gpu(d, st + h * dx, 2)))
    ad = np.sum(grad * dx)
    print("fast", ad, fd, flush=True)
    assert close(ad, fd), (ad, fd)


@pytest.mark.parametrize("fast", [0, 1], ids=["closed_form", "table"])
def test_s95p_adjoint_matches_fd(fast):
    """s95p EOS: the closed forms' derivatives (exact mode) and the table
    interpolant's (fast mode), through the root, the velocities, T(e) and
    the shear source's s(e); state and parameter gradients against central
    differences of the same GPU forward"""
    d, st = small_case(fast=fast, eos=EOS_S95P, eos_table=fast, eta_law=ETA_S_OVER_4PI, eta_coef=1.0,
                       tpi_law=TPI_OVER_P, tpi_coef=1.5)
    rng = np.random.default_rng(12 + fast)
    seed = rng.standard_normal(st.shape)
    print("case adjoint s95p fast=%d" % fast, flush=True)
    _, grad, pg = run_gpu(d, st, 3, 2, seed)
    for trial in range(2):
        dx = direction(st, rng)
        # the s95p step is nonlinear enough that two-point differences at
        # h = 1e-6 carry ~1e-6 truncation error; the four-point difference's
        # error falls as h^4
        fd = fd4(lambda h: np.sum(seed * run_gpu(d, st + h * dx, 3)))
        ad = np.sum(grad * dx)
        print(trial, ad, fd, abs(ad - fd) / abs(fd), flush=True)
        assert close(ad, fd), (ad, fd)
    for q, name in enumerate(["eta_coef", "tpi_coef"]):
        c0 = getattr(d, name)

        def J(h):
            dd = d.copy()
            setattr(dd, name, c0 * (1 + h))
            return np.sum(seed * run_gpu(dd, st, 3))

        fd = fd4(J, HP) / c0
        print(name, pg[q], fd, flush=True)
        assert close(pg[q], fd), (name, pg[q], fd)


def test_torch_harness_momentum_anisotropy():
    """d eps_2 / d(initial e-field, eta/s, tau_pi coefficient) through
    torch.autograd against central differences of the same pipeline."""
    d = make_desc(16, 14, 6, eta_coef=0.08, tpi_coef=5.0)
    base = ic.glauber(d, TAU, TS, b=8.0, e_peak_gev=10.0)
    e0 = torch.tensor(base[F_EN], dtype=torch.float64)
    rng = np.random.default_rng(6)
    e0 = e0 * (1 + 0.05 * torch.from_numpy(rng.standard_normal(e0.shape[0])))
    print("case torch harness", flush=True)

    def objective(e, c_eta, c_tpi):
        st = autograd.state_from_energy(d, e, TAU)
        fin = autograd.run(st, c_eta, c_tpi, d, TAU, TS, 4, every=2)
        return autograd.momentum_anisotropy(d, fin, TAU + 4 * TS)

    e = e0.clone().requires_grad_(True)
    ce = torch.tensor(0.08, dtype=torch.float64, requires_grad=True)
    ct = torch.tensor(5.0, dtype=torch.float64, requires_grad=True)
    J = objective(e, ce, ct)
    J.backward()
    with torch.no_grad():
        de = torch.from_numpy(rng.standard_normal(e0.shape[0])) * e0
        fd = float(fd4(lambda h: objective(e0 + h * de, ce, ct)))
        ad = float((e.grad * de).sum())
        print("e-field", ad, fd, flush=True)
        assert abs(ad - fd) <= BAR * abs(fd) + 1e-12
        for t, g in ((ce, ce.grad), (ct, ct.grad)):
            p = t.detach().clone()
            fd = float(fd4(lambda h: objective(e0, p * (1 + h) if t is ce else ce, p * (1 + h) if t is ct else ct), HP))
            fd /= float(p)
            print("param", float(g), fd, flush=True)
            assert abs(float(g) - fd) <= BAR * abs(fd) + 1e-12


@pytest.mark.parametrize("every", [0, 3])
def test_closing_a_recorded_context_frees_its_stage_states(every):
    """every == 0 keeps the stage states in slabs of device memory, every > 0
    the current segment's in a ring beside the checkpoints: closing the
    context (the autograd backward does) must free them, so repeated
    recorded runs leave the device's free memory where it was"""
    d = make_desc(24, 22, 16)
    st = ic.random_smooth(d, seed=3, tau=TAU, pi_frac=0.02)
    seed = np.random.default_rng(8).standard_normal(st.shape)
    print("case adjoint memory release every=%d" % every, flush=True)

    def once():
        x = torch.tensor(st, requires_grad=True)
        fin = autograd.run(x, torch.tensor(d.eta_coef, dtype=torch.float64),
                           torch.tensor(d.tpi_coef, dtype=torch.float64), d, TAU, TS, 6, every=every)
        (fin * torch.from_numpy(seed)).sum().backward()
        return x.grad

    torch.cuda.synchronize()
    once()
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(3):
        g = once()
        assert torch.isfinite(g).all()
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    # one recording holds 6 steps x 2 stages x 29 doubles per ghosted cell
    # (every = 3: a ring of 3 x 2 stage states and 3 checkpoints)
    slab = 6 * 2 * 29 * d.cells() * 8
    assert free0 - free1 < slab / 2, (free0, free1, slab)


# ------------------------------------------------------------------ Cartesian (TXYZ5)

def cart_case(rk=2, fast=0, **kw):
    from viral_amd.desc import CARTESIAN
    d = make_desc(10, 9, 8, coords=CARTESIAN, rk=rk, fast_math=fast, **kw)
    if d.eos == EOS_S95P:
        dref = d.copy()
        dref.fast_math = dref.eos_table = 0
        st = ic.random_smooth(d, seed=21, tau=TAU, pi_frac=0.02, e_scale=30.0,
                              eos_p=lambda e: capi.eos_eval(dref, e)[0].reshape(np.shape(e)))
    else:
        st = ic.random_smooth(d, seed=21, tau=TAU, pi_frac=0.02)
    return d, st


@pytest.mark.parametrize("rk,every", [(1, 1), (2, 2), (3, 2)])
def test_cartesian_adjoint_state_gradient_matches_fd(rk, every):
    """the Cartesian (TXYZ5) step in reverse: the 4-D root by the
    implicit-function theorem, the flux with the shear tensor built from its
    five evolved components, no geometric sources, the shear-slice quirk"""
    d, st = cart_case(rk=rk, eta_law=ETA_S_OVER_4PI, eta_coef=1.0, tpi_law=TPI_OVER_P, tpi_coef=1.5)
    rng = np.random.default_rng(30 + rk)
    seed = rng.standard_normal(st.shape)
    print("case cartesian adjoint rk=%d every=%d" % (rk, every), flush=True)
    _, grad, pg = run_gpu(d, st, 3, every, seed)
    for trial in range(2):
        dx = direction(st, rng)
        fd = fd4(lambda h: np.sum(seed * run_gpu(d, st + h * dx, 3)))
        ad = np.sum(grad * dx)
        print(trial, ad, fd, abs(ad - fd) / abs(fd), flush=True)
        assert close(ad, fd), (ad, fd)


@pytest.mark.parametrize("fast", [0, 1], ids=["exact", "fast_table"])
def test_cartesian_adjoint_s95p_state_and_coefficients(fast):
    """s95p EOS (fast mode: the table in the sources; the Cartesian root and
    its outputs keep the closed forms): state and transport-coefficient
    gradients against differences of the GPU forward, and of the CPU oracle
    (the reference gradient) in exact mode"""
    d, st = cart_case(fast=fast, eos=EOS_S95P, eos_table=fast, eta_law=ETA_S_OVER_4PI, eta_coef=1.0,
                      tpi_law=TPI_OVER_P, tpi_coef=1.5)
    rng = np.random.default_rng(40 + fast)
    seed = rng.standard_normal(st.shape)
    print("case cartesian adjoint s95p fast=%d" % fast, flush=True)
    _, grad, pg = run_gpu(d, st, 3, 2, seed)
    runs = [run_gpu] + ([run_cpu] if not fast else [])
    dx = direction(st, rng)
    ad = np.sum(grad * dx)
    for run in runs:
        fd = fd4(lambda h: np.sum(seed * run(d, st + h * dx, 3)))
        print("state", run.__name__, ad, fd, abs(ad - fd) / abs(fd), flush=True)
        assert close(ad, fd), (run.__name__, ad, fd)
    for q, name in enumerate(["eta_coef", "tpi_coef"]):
        c0 = getattr(d, name)
        for run in runs:
            def J(h):
                dd = d.copy()
                setattr(dd, name, c0 * (1 + h))
                return np.sum(seed * run(dd, st, 3))

            fd = fd4(J, HP) / c0
            print(name, run.__name__, pg[q], fd, flush=True)
            assert close(pg[q], fd), (name, run.__name__, pg[q], fd)


def test_cartesian_stored_stage_states_match_recompute():
    d, st = cart_case(rk=3)
    seed = np.random.default_rng(6).standard_normal(st.shape)
    print("case cartesian adjoint stored", flush=True)
    plain = run_gpu(d, st, 4)
    _, g3, p3 = run_gpu(d, st, 4, 3, seed)
    _, g1, p1 = run_gpu(d, st, 4, 1, seed)
    assert np.array_equal(g1, g3) and np.array_equal(p1, p3)
    h = capi.Hydro(d)
    h.upload(st)
    h.adjoint_begin(0)
    assert h.steps(TAU, TS, 4) == (OK, 4)
    assert np.array_equal(h.download().view(np.int64), plain.view(np.int64))
    g0, p0 = h.adjoint(seed)
    assert np.array_equal(g0, g3) and np.array_equal(p0, p3)
    h.adjoint_end()
