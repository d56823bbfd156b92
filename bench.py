"""Benchmark of the MI355X viscous hydro step (one GPU).

Headline workload: the batched-ensemble configuration -- 64 seeded
fluctuating events (200 Gaussian hot spots of width 0.4 fm from two
Woods-Saxon nuclei, b ~ U(0,12) fm) on uneven Milne grids (transverse side
~U{96..256} cells, eta cells ~U{32..64}, dx = deta = 0.1), viscous
Kurganov-Tadmor step with the s95p EOS (table), eta/s = 0.08,
tau_pi = 5 eta/(e+p), RK3, dtau = 0.002 fm/c, FP64, all events in one context.
A step is one full RK step of every event; `value` is interior cell-updates
per second of the batched launch (one launch per kernel phase over all
events, hipGraph replay, autotuned launch configuration), and `plain` is the
same ensemble stepped with per-event launches of the same kernels.  Inputs are
resident in HBM before timing starts.

Further legs (each its own JSON key): the smooth-Glauber 200x200x64 grid in
fast and exact arithmetic, the Gubser 64x64x8 grid, the 256x256x96 hot-spot
grid (KT and SHASTA), the adjoint of 200 steps on 128x128x32 with checkpoints
every 20 steps, a small-event ensemble, and the CPU restatement of the step
on the host's cores (one thread and all cores).

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 cell-updates/s: plain vs batched ensemble; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK_TFLOPS = 78.6
# Compulsory HBM bytes of an ideal fused step per interior cell-update: the
# per-cell state T^{tau mu} (4), pi (10), e, u (4), prevu (4) = 23 doubles,
# read and written once per stage where the stage needs them (P, T and v are
# recomputed).  RK2 60 reads + 38 writes, RK3 97 + 57, RK1 42 doubles.
IDEAL_BYTES_PER_UPDATE = {1: 336, 2: 784, 3: 1232}


def _progress(msg):
    """One line on stderr when a leg of the run ends (the JSON line goes to
    stdout at the very end)."""
    sys.stderr.write("[bench] %s\n" % msg)
    sys.stderr.flush()


def workload_desc(rk=3, fast=1):
    from viral_amd.desc import make_desc, EOS_S95P, ETA_TIMES_S, TPI_OVER_EP
    return make_desc(200, 200, 64, dx=0.1, dy=0.1, deta=0.1, rk=rk, eos=EOS_S95P,
                     eta_law=ETA_TIMES_S, eta_coef=0.08, tpi_law=TPI_OVER_EP, tpi_coef=5.0, fast_math=fast,
                     eos_table=fast)


def eos_p_vec(desc):
    """s95p pressure on the host for the initial state: the library's host
    EOS (viral_eos_eval) in its closed form."""
    from viral_amd import capi
    dref = desc.copy()
    dref.fast_math = dref.eos_table = 0
    return lambda e: capi.eos_eval(dref, e)[0].reshape(e.shape)


# Compulsory HBM bytes per launch of each kernel: every input field read
# once and every output written once over the cells/faces of one launch.
#   flux_x/_y/_eta, fast LDS tiles (per interior cell): 21 state fields read
#     (e, P, PI, u^mu, pi^{mu nu}, T^{tau mu}) + the cell's 15 partial results
#     written (x) or read, added to and written back (y, eta: the tiles sum
#     the three directions into one array) = 36 or 51 doubles; exact per-face
#     kernel: 21 state fields read + 15 face fluxes written per face
#   source (per interior cell): e, P, PI, u (4), prevu (4), pi (10) read,
#     15 sources written = 36 doubles = 288 B
#   update (per interior cell): the partial results (fast KT: the 15 summed
#     ones; exact KT and SHASTA: 15 per direction), 29 state fields written,
#     plus per stage kind the RK registers.  Exact mode: the first stage
#     writes Var0, Var1 (30); RK2 final reads Var0, Var1 (30); RK3 middle
#     reads Var0, Var1 and writes Var2 (30 + 15); RK3 final reads Var0, Var2
#     (30).  Fast mode keeps Var0 only: written by the first stage (15), read
#     by the later ones (15), which take the previous stage's Var from their
#     input state.  Every stage reads its input state's Var: T^{tau mu} (4)
#     when the sources are fused in (pi and PI are among the source's 21
#     inputs), the 15 evolved fields otherwise.  Sources in their own kernel:
#     15 sources + e, u, prevu (9) read.  + 4 B iteration count, 1 B flags
#     (+1 B read after the first stage).  Averaged over the stages.
def update_doubles(rk, fused, fast=True, kt=True):
    partial = 15 if (fast and kt) else 45
    base = partial + 29 + (21 if fused else 15 + 9)
    own = 4 if fused else 15
    if fast:
        kinds = [own + 15] + [own + 15] * (rk - 1)
    else:
        kinds = {1: [own + 30], 2: [own + 30, 30], 3: [own + 30, 30 + 15, 30]}[rk]
    return [base + k for k in kinds]


def kernel_bytes(name, d, fused):
    faces = [(d.nx + 1) * d.ny * d.nz, d.nx * (d.ny + 1) * d.nz, d.nx * d.ny * (d.nz + 1)]
    cells = d.nx * d.ny * d.nz
    if name.startswith("flux_"):
        q = {"flux_x": 0, "flux_y": 1, "flux_eta": 2}[name]
        if d.fast_math and d.scheme == 0:
            return cells * (288.0 if q == 0 else 408.0)
        return (cells if d.fast_math else faces[q]) * 288.0
    if name == "source":
        return cells * 288.0
    if name == "update":
        per = update_doubles(d.rk, fused, bool(d.fast_math), d.scheme == 0)
        avg = sum(8.0 * x + 5.0 + (1.0 if q else 0.0) for q, x in enumerate(per)) / len(per)
        return cells * avg
    if name == "k0_broadcast":
        return cells * 8.0 * 10
    return 0.0


# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE passes
# on the 200x200x64 RK3 fast-mode workload (tabulated s95p), each counter
# scaled by what it reads for a 1 GiB device copy (read x2.00, write x1.00);
# profiles/counters_kt_s95p_glauber_200x200x64_rk3.md lists the passes.
# The update kernel's figure is for its fused form (sources evaluated inside).
MEASURED_TRAFFIC = {
    "update_fused": 1.893e9,
    "flux_x": 1.007e9,
    "flux_y": 1.285e9,
    "flux_eta": 1.272e9,
    "k0_broadcast": 0.175e9,
}


# The headline ensemble's dominant kernel (the fused update of one stage, one
# batched launch over all 64 events), from rocprofv3 counter passes on the
# ensemble leg: HBM bytes per launch (FETCH_SIZE x 2.00 x 1024 + WRITE_SIZE x
# 1.00 x 1024, the 1 GiB copy's calibration) and FP64 VALU work per launch
# (64 lanes x (2 SQ_INSTS_VALU_FMA_F64 + SQ_INSTS_VALU_MUL_F64 +
# SQ_INSTS_VALU_ADD_F64 + SQ_INSTS_VALU_TRANS_F64)), VALU wave-instructions per
# launch and the VALU issue fraction (SQ_INSTS_VALU / (256 x GRBM_GUI_ACTIVE
# per XCD): the chip issues at most one FP64 wave-instruction per CU and
# cycle).  Averages over the six launches of the two steps the bench times
# per kernel; profiles/counters_ensemble_64_hotspot_events_rk3.md lists the
# passes.
ENSEMBLE_TRAFFIC = {"update": 87.37e9}
ENSEMBLE_COUNTERS = {
    "update": {"fp64_flops": 2.554e11, "valu_wave_instructions": 4.603e9, "valu_issue_frac": 0.277,
               "source": "rocprofv3 --pmc passes on bench.py --legs c4 (profiles/counters_ensemble_64_hotspot_"
                         "events_rk3.md)"},
}


def ensemble_counters(kernel, avg_s):
    c = ENSEMBLE_COUNTERS.get(kernel)
    if not c:
        return {"valu_frac": None}
    return {"valu_frac": c["fp64_flops"] / avg_s / 1e12 / FP64_VALU_PEAK_TFLOPS,
            "fp64_tflops": c["fp64_flops"] / avg_s / 1e12, "valu_issue_frac": c["valu_issue_frac"],
            "counters": c["source"]}


def measured_traffic(name, d, fused):
    if (d.nx, d.ny, d.nz, d.rk, d.fast_math, d.eos_table) == (200, 200, 64, 3, 1, 1):
        if name == "update":
            return MEASURED_TRAFFIC.get("update_fused") if fused else None
        return MEASURED_TRAFFIC.get(name)
    return None


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_all_core():
    """The oracle built for the host it runs on (-O3 -march=native -fopenmp,
    -ffp-contract=off so it rounds as the test build does) into a temporary
    directory; the prebuilt portable OpenMP build (x86-64-v3) where no
    compiler is at hand."""
    import subprocess
    import tempfile
    src = os.path.join(ROOT, "oracle", "oracle.c")
    out = os.path.join(tempfile.mkdtemp(prefix="viral_oracle_"), "liboracle_native.so")
    flags = ["-O3", "-march=native", "-fopenmp", "-ffp-contract=off"]
    r = subprocess.run(["gcc"] + flags + ["-fPIC", "-shared", "-o", out, src, "-lm"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT)
    if r.returncode == 0:
        return out, " ".join(flags)
    return os.path.join(ROOT, "oracle", "liboracle_omp.so"), "-O3 -march=x86-64-v3 -fopenmp -ffp-contract=off"


def cpu_baseline(seconds_target=12.0, max_steps=3):
    """The CPU restatement of the reference step (oracle/oracle.c, the
    reference's loop structure and operation order) on a bounded sample of
    the headline workload: the central 16 eta slices of the ensemble's first
    event (hot spots, s95p closed form, eta/s = 0.08, RK3), timed first on one
    thread, then with OpenMP over the host cores this process may use."""
    import ctypes
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ctypes
    from viral_amd import capi, ic
    descs, bs = ensemble_descs(1)
    d = descs[0].copy()
    d.fast_math = d.eos_table = 0
    d.nz = 16
    st = ic.hotspots(d, 0.6, 0.002, seed=1000, b=bs[0], eos_p=lambda x: capi.eos_eval(d, x)[0].reshape(x.shape))
    path, flags = _oracle_all_core()
    L = oracle_ctypes.load(path)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    cells = d.nx * d.ny * d.nz

    L.omp_set_num_threads.argtypes = [ctypes.c_int]

    def timed(nthreads):
        L.omp_set_num_threads(nthreads)
        o = oracle_ctypes.Oracle(d, L)
        o.set_state(st)
        t0 = time.perf_counter()
        steps = 0
        while True:
            assert o.step(0.6 + steps * 0.002, 0.002) == 0
            steps += 1
            if time.perf_counter() - t0 > seconds_target or steps >= max_steps:
                break
        dt = time.perf_counter() - t0
        return cells * steps / dt, steps

    one, n1 = timed(1)
    _progress("CPU one thread: %.3g cell-updates/s" % one)
    allc, na = timed(threads)
    sample = ("%d x %d x 16 cells (the central 16 eta slices of the ensemble's first event, hot spots, s95p closed "
              "form, RK3, dtau=0.002), CPU oracle built with %s" % (d.nx, d.ny, flags))
    return {"value": allc, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": "%d RK3 step(s) on %d threads of %s" % (na, threads, sample),
            "single_thread": {"value": one, "unit": "cell-updates/s", "cores": 1, "steps": n1},
            "nproc": os.cpu_count(), "cpu_model": _cpu_model()}


def ensemble_descs(n_events, seed=4, side=(96, 256), eta=(32, 64)):
    """Uneven ensemble: square transverse side ~U{side} cells, eta cells
    ~U{eta}, b ~U(0,12) fm; viscous KT, s95p EOS (table), eta/s=0.08,
    tau_pi=5eta/(e+p), RK3, dx=deta=0.1."""
    import numpy as np
    from viral_amd.desc import make_desc, EOS_S95P, ETA_TIMES_S, TPI_OVER_EP
    rng = np.random.default_rng(seed)
    descs, bs = [], []
    for _ in range(n_events):
        side_ = int(rng.integers(side[0], side[1] + 1))
        neta = int(rng.integers(eta[0], eta[1] + 1))
        descs.append(make_desc(side_, side_, neta, rk=3, eos=EOS_S95P, eta_law=ETA_TIMES_S, eta_coef=0.08,
                               tpi_law=TPI_OVER_EP, tpi_coef=5.0, fast_math=1, eos_table=1))
        bs.append(float(rng.uniform(0.0, 12.0)))
    return descs, bs


def bench_ensemble(n_events, steps, warmup=2, tau0=0.6, ts=0.002, side=(96, 256), eta=(32, 64), kernel_times=False,
                   modes=("plain_default", "batched", "plain")):
    """The same uneven hot-spot ensemble in one context stepped three ways,
    each synchronising once per step:
      plain_default  one launch per event and kernel phase, the library's
                     default launch options
      plain          one launch per event and kernel phase with the kernel
                     options the autotuner picks for the batched launch
                     (flux block sizes and tile shapes, fused sources)
      batched        one launch per kernel phase over all events, hipGraph
                     replay of the whole RK step, autotuned (batched_over_plain
                     is batched / plain)
    kernel_times: afterwards, two batched steps without the graph with HIP
    events around every launch (the kernels the graph replays), for the
    per-kernel average durations."""
    from concurrent.futures import ThreadPoolExecutor
    from viral_amd import capi, ic
    descs, bs = ensemble_descs(n_events, side=side, eta=eta)
    cells = sum(d.nx * d.ny * d.nz for d in descs)

    def make_state(e):
        d = descs[e]
        dref = d.copy()
        dref.fast_math = dref.eos_table = 0
        return ic.hotspots(d, tau0, ts, seed=1000 + e, b=bs[e], eos_p=lambda x: capi.eos_eval(dref, x)[0].reshape(x.shape))

    t_ic = time.perf_counter()
    h = capi.Hydro(descs)
    with ThreadPoolExecutor(4) as pool:
        for e, st in enumerate(pool.map(make_state, range(n_events))):
            h.upload(st, ev=e)
            del st
    t_ic = time.perf_counter() - t_ic
    clock = [tau0]

    def run(n):
        for _ in range(n):
            r = h.step(clock[0], ts)
            assert r == 0, r
            clock[0] += ts

    def timed(n):
        run(warmup)
        t0 = time.perf_counter()
        run(n)
        return time.perf_counter() - t0

    out = {"workload": "%d hotspot events (200 spots, width 0.4 fm, b~U(0,12) fm), transverse side ~U{%d..%d}, "
                       "eta cells ~U{%d..%d}, viscous KT, s95p table, eta/s=0.08, tau_pi=5eta/(e+p), RK3, dtau=%g"
                       % (n_events, side[0], side[1], eta[0], eta[1], ts),
           "events": n_events, "interior_cells": cells, "steps": steps, "warmup": warmup,
           "ic_seconds": round(t_ic, 1)}
    if "plain_default" in modes:
        h.set_launch(batched=0, graph=0, blocks_per_cu=0, fusion=0, xcd=0)
        dt = timed(steps)
        out["plain_default"] = {"value": cells * steps / dt, "ms_per_step": dt * 1000.0 / steps,
                                "launch": h.launch().as_dict()}
        _progress("%d events: plain_default %.3g cell-updates/s" % (n_events, out["plain_default"]["value"]))
    h.set_launch(batched=1, graph=1)
    h.autotune(reps=1)
    tuned = h.launch().as_dict()
    out["launch"] = tuned
    dt = timed(steps)
    out["batched"] = {"value": cells * steps / dt, "ms_per_step": dt * 1000.0 / steps}
    _progress("%d events: batched %.3g cell-updates/s" % (n_events, out["batched"]["value"]))
    if "plain" in modes:
        h.set_launch(batched=0, graph=0, blocks_per_cu=0, xcd=0)
        dt = timed(steps)
        out["plain"] = {"value": cells * steps / dt, "ms_per_step": dt * 1000.0 / steps,
                        "launch": h.launch().as_dict()}
        out["batched_over_plain"] = out["batched"]["value"] / out["plain"]["value"]
        _progress("%d events: plain %.3g cell-updates/s" % (n_events, out["plain"]["value"]))
    if "plain_default" in modes:
        out["batched_over_plain_default"] = out["batched"]["value"] / out["plain_default"]["value"]
    if kernel_times:
        h.set_launch(**{k: v for k, v in tuned.items() if k != "graph"}, graph=0)
        run(1)
        h.timers(True)
        run(2)
        rep = h.timer_report()
        h.timers(False)
        out["kernel_times"] = rep
        out["kernel_launch"] = h.launch().as_dict()
    h.close()
    return out


def kernel_roofline(rep, d_rk, cells, fused, traffic=None, counters=None):
    """Roofline figure of the dominant kernel of a timed run (`rep`: kernel ->
    (total ms, launches)): algorithmic bytes per launch from the ideal fused
    step's bytes per cell-update (one stage's share for a per-stage kernel,
    times the interior cells one launch covers), over the kernel's HIP-event
    average launch time."""
    dom = max(rep, key=lambda k: rep[k][0])
    tot_ms, launches = rep[dom]
    avg_s = tot_ms / 1000.0 / max(launches, 1)
    per_stage = IDEAL_BYTES_PER_UPDATE[d_rk] / d_rk
    algo = per_stage * cells
    achieved = algo / avg_s / 1e9
    out = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "avg_launch_ms": avg_s * 1000.0,
           "launches": launches, "algorithmic_bytes_per_launch": algo,
           "basis": "ideal fused step, %d B per RK%d cell-update (23 state doubles per cell), one stage's share "
                    "(%.1f B) per cell per launch" % (IDEAL_BYTES_PER_UPDATE[d_rk], d_rk, d_rk and per_stage)}
    if traffic:
        out["traffic_over_algorithmic"] = traffic / algo
    if counters:
        out.update(counters(dom, avg_s))
    return out


# Energy floor of the fluctuating-event states, relative to their peak
# (10 GeV/fm^3): 1e-2, i.e. 0.1 GeV/fm^3.  At 1e-7 the hot spots expand into
# near-vacuum and a dilute cell turns acausal (KT after 259 steps; SHASTA with
# the full-step sub-step after 36, at 1e-3 after 338); at 1e-2 both schemes
# run the configuration's 400 steps.
E_FLOOR_C3 = 1e-2


def bench_c3(steps, tau0=0.6, ts=0.002):
    """The fluctuating-event configuration: 200 Gaussian hot spots of width
    0.4 fm from two Woods-Saxon nuclei (b ~ U(0,10) fm, seed 3) on a Milne
    256x256x96 grid, s95p EOS (table in fast mode), eta/s = 0.12,
    tau_pi = 5 eta/(e+p), RK2, `steps` steps (400 as configured), timed from
    the state after two warm-up steps: the SHASTA step (the configuration's
    scheme) with the full-step sub-step (VIRAL_Q_SHASTA_FULLSTEP; the
    reference's dtau/2 makes the grid-scale checkerboard grow 2.5x per step in
    3-D, so its step fails within 5 steps on this state; the library's
    default launch), and the viscous KT step (autotuned) on the same state,
    with the update kernel's HBM figure."""
    import torch
    from viral_amd import capi, ic
    from viral_amd.desc import make_desc, EOS_S95P, ETA_TIMES_S, TPI_OVER_EP, SHASTA, Q_DEFAULT, Q_SHASTA_FULLSTEP
    out = {"workload": "200 hot spots (width 0.4 fm, b~U(0,10) fm, seed 3, energy floor %g of the peak), Milne "
                       "256x256x96, s95p table, eta/s=0.12, tau_pi=5eta/(e+p), RK2, dtau=%g, fast arithmetic, %d steps; "
                       "SHASTA with the full-step sub-step, and KT" % (E_FLOOR_C3, ts, steps)}
    for scheme in ("shasta", "kt"):
        d = make_desc(256, 256, 96, dx=0.1, dy=0.1, deta=0.1, rk=2, eos=EOS_S95P, eta_law=ETA_TIMES_S, eta_coef=0.12,
                      tpi_law=TPI_OVER_EP, tpi_coef=5.0, fast_math=1, eos_table=1,
                      scheme=SHASTA if scheme == "shasta" else 0,
                      quirks=Q_DEFAULT | (Q_SHASTA_FULLSTEP if scheme == "shasta" else 0))
        dref = d.copy()
        dref.fast_math = dref.eos_table = 0
        st = ic.hotspots(d, tau0, ts, seed=3, e_floor=E_FLOOR_C3,
                         eos_p=lambda x: capi.eos_eval(dref, x)[0].reshape(x.shape))
        h = capi.Hydro(d)
        h.upload(st)
        tau = tau0
        for _ in range(2):
            r = h.step(tau, ts)
            assert r == 0, r
            tau += ts
        if scheme == "kt":  # SHASTA keeps the library's default launch
            h.autotune(reps=1)
        h.timers(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            r = h.step(tau, ts)
            assert r == 0, (scheme, r, h.first_bad_cell())
            tau += ts
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        rep = h.timer_report()
        h.timers(False)
        cells = d.nx * d.ny * d.nz
        leg = {"value": cells * steps / dt, "unit": "cell-updates/s", "ms_per_step": dt * 1e3 / steps, "steps": steps,
               "launch": h.launch().as_dict(),
               "kernel_ms_per_launch": {k: round(v[0] / v[1], 4) for k, v in rep.items() if v[1]}}
        leg["roofline"] = kernel_roofline(rep, d.rk, cells, leg["launch"]["fusion"] > 0)
        out[scheme] = leg
        h.close()
        _progress("256x256x96 %s: %.3g cell-updates/s" % (scheme, leg["value"]))
    return out


def bench_adjoint(steps, every=20, tau0=0.6, ts=0.002, eos="s95p", fast=1):
    """Reverse mode through `steps` viscous KT steps of a smooth Glauber event
    (b=8 fm, peak 30 GeV/fm^3) on a 128x128x32 Milne grid, eta/s=0.08,
    tau_pi=5eta/(e+p), RK2, checkpoints every `every` steps: the gradient of
    the final momentum anisotropy eps_2 with respect to the initial state,
    eta/s and the tau_pi coefficient.  eos "s95p" with fast=1 reads the s95p
    table (the adjoint differentiates its interpolant); "conformal" is e = 3p.
    Reports the time of the recorded forward run and of the adjoint call
    (adjoint_ms: the C call, with the seed's upload and the gradient's
    download, 2 x 29 fields over PCIe from and into page-locked host
    buffers; adjoint_device_ms: its device part,
    HIP events around the recompute and reverse kernels) with checkpoints
    every `every` steps (the adjoint recomputes the stage states between
    them), and as a labelled extra with every stage state kept during the
    forward run (every = 0, no recompute: 63 GB of HBM for 200 RK2 steps)."""
    import numpy as np
    import torch
    from viral_amd import capi, ic, autograd
    from viral_amd.desc import make_desc, F_EN, F_T00, F_P, EOS_S95P, EOS_CONFORMAL
    s95 = eos == "s95p"
    d = make_desc(128, 128, 32, rk=2, eta_coef=0.08, tpi_coef=5.0, fast_math=fast,
                  eos=EOS_S95P if s95 else EOS_CONFORMAL, eos_table=int(s95 and fast))
    st = ic.glauber(d, tau0, ts, b=8.0, eos_p=eos_p_vec(d) if s95 else None)
    h = capi.Hydro(d)
    h.upload(st)
    r = h.step(tau0, ts)   # warm-up (kernel load), then restart from the initial state
    h.upload(st)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r, done = h.steps(tau0, ts, steps)
    t_plain = time.perf_counter() - t0
    assert r == 0, (r, done)
    runs = {}
    seed = eps2 = None
    for ev in (0, every):
        h.upload(st)
        h.adjoint_begin(ev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r, done = h.steps(tau0, ts, steps)
        t_fwd = time.perf_counter() - t0
        assert r == 0, (r, done)
        if seed is None:
            fin = torch.from_numpy(h.download()).requires_grad_(True)
            eps2 = autograd.momentum_anisotropy(d, fin, tau0 + steps * ts)
            eps2.backward()
            # seed and gradient in page-locked host buffers (filled before the timing)
            seed = capi.pinned_empty(fin.grad.shape)
            seed[...] = fin.grad.numpy()
        gbuf = capi.pinned_empty(seed.shape)
        h.timers(True)
        t0 = time.perf_counter()
        grad, pg = h.adjoint(seed, out=gbuf)
        t_adj = time.perf_counter() - t0
        dev_ms = h.timer_report()["adjoint_device"][0]
        h.timers(False)
        h.adjoint_end()
        runs[ev] = (t_fwd, t_adj, grad, pg, dev_ms)
        _progress("adjoint every=%d: %.0f ms (forward %.0f ms)" % (ev, t_adj * 1e3, t_plain * 1e3))
    h.close()
    dref = d.copy()
    dref.fast_math = dref.eos_table = 0
    dpde = capi.eos_eval(dref, st[F_EN])[1]
    t_fwd, t_adj, grad, pg, dev = runs[0]
    c_fwd, c_adj, c_grad, c_pg, c_dev = runs[every]
    assert np.array_equal(grad, c_grad) and np.array_equal(pg, c_pg), "stored and recomputed stage states disagree"
    slab_gb = 29 * 8 * d.cells() * 2 * steps / 1e9
    return {"workload": "adjoint of %d viscous KT steps, smooth Glauber b=8 fm, Milne 128x128x32, %s EOS, "
                        "eta/s=0.08, tau_pi=5eta/(e+p), RK2, dtau=%g, checkpoints every %d steps, %s arithmetic; "
                        "d eps_2 / d(initial state, eta/s, tau_pi coefficient)"
                        % (steps, "s95p (table)" if s95 and fast else eos, ts, every, "fast" if fast else "exact"),
            "steps": steps, "every": every, "forward_ms": t_plain * 1e3, "forward_recorded_ms": c_fwd * 1e3,
            "adjoint_ms": c_adj * 1e3, "adjoint_over_forward": c_adj / t_plain,
            "adjoint_device_ms": c_dev, "adjoint_device_over_forward": c_dev / (t_plain * 1e3),
            "stored_stages": {"note": "every stage state kept in HBM during the recorded forward (every = 0, no "
                                      "recompute), %.1f GB" % slab_gb,
                              "forward_recorded_ms": t_fwd * 1e3, "adjoint_ms": t_adj * 1e3,
                              "adjoint_over_forward": t_adj / t_plain, "adjoint_device_ms": dev,
                              "adjoint_device_over_forward": dev / (t_plain * 1e3), "same_gradient_bits": True},
            "eps2": float(eps2.detach()), "d_eps2_d_eta_coef": float(pg[0]), "d_eps2_d_tpi_coef": float(pg[1]),
            "d_eps2_d_initial_e_norm": float(np.linalg.norm(grad[F_EN] + grad[F_T00] + grad[F_P] * dpde)),
            "d_eps2_d_initial_e_note": "the initial e-field of the fluid at rest enters the state as En, T^{tau tau} "
                                       "= e and P(e) (s95p closed form): dJ/de = dJ/dEn + dJ/dT^tt + dJ/dP dP/de"}


def bench_c2(steps, warmup, exact_steps):
    """The smooth-Glauber grid: viscous KT, s95p EOS, Glauber b=7 fm, Milne
    200x200x64, dx=deta=0.1, tau0=0.6, dtau=0.002, eta/s=0.08,
    tau_pi=5eta/(e+p), RK3.  Fast arithmetic (s95p table, autotuned, HIP
    events around every launch) and exact arithmetic (the oracle's operation
    order, bitwise equal to it)."""
    import torch
    from viral_amd import capi, ic
    out = {"workload": "viscous KT, s95p EOS, Glauber b=7 fm, Milne 200x200x64, dx=deta=0.1, tau0=0.6, dtau=0.002, "
                       "eta/s=0.08, tau_pi=5eta/(e+p), RK3"}
    for mode, n in (("fast", steps), ("exact", exact_steps)):
        if n <= 0:
            continue
        d = workload_desc(fast=1 if mode == "fast" else 0)
        tau0, ts = 0.6, 0.002
        st = ic.glauber(d, tau0, ts, b=7.0, eos_p=eos_p_vec(d))
        h = capi.Hydro(d)
        h.upload(st)
        tau = tau0
        for _ in range(warmup if mode == "fast" else 1):
            assert h.step(tau, ts) == 0
            tau += ts
        tiles = h.autotune(reps=2) if d.fast_math else None
        launch = h.launch().as_dict()
        h.timers(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            assert h.step(tau, ts) == 0
            tau += ts
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        rep = h.timer_report()
        h.timers(False)
        cells = d.nx * d.ny * d.nz
        fused = launch["fusion"] > 0
        leg = {"value": cells * n / dt, "unit": "cell-updates/s", "ms_per_step": dt * 1000.0 / n, "steps": n,
               "launch": launch, "flux_tiles": tiles,
               "kernel_ms_per_launch": {k: round(v[0] / v[1], 4) for k, v in rep.items() if v[1]}}
        rl = kernel_roofline(rep, d.rk, cells, fused, traffic=measured_traffic(max(rep, key=lambda k: rep[k][0]), d,
                                                                              fused))
        dom = rl["kernel"]
        rl["kernel_model"] = {"algorithmic_bytes_per_launch": kernel_bytes(dom, d, fused),
                              "note": "every input field of the kernel read once, every output written once"}
        leg["roofline"] = rl
        leg["step_hbm"] = {"ideal_bytes_per_cell_update": IDEAL_BYTES_PER_UPDATE[d.rk],
                           "achieved_GBs": leg["value"] * IDEAL_BYTES_PER_UPDATE[d.rk] / 1e9,
                           "frac": leg["value"] * IDEAL_BYTES_PER_UPDATE[d.rk] / 1e9 / HBM_PEAK_GBS}
        out[mode] = leg
        h.close()
        _progress("200x200x64 %s: %.3g cell-updates/s" % (mode, leg["value"]))
    return out


def bench_c1(steps=100):
    """The Gubser grid: ideal Gubser initial flow (q = 1 fm^-1) at tau0 = 1,
    conformal EOS, eta/s = 0.2, tau_pi = 5 eta/(e+p), Milne 64x64x8,
    dx = deta = 0.1, dtau = 0.01, RK2; all `steps` steps in one viral_steps
    call (hipGraph replay), fast and exact arithmetic."""
    import torch
    from viral_amd import capi, ic
    from viral_amd.desc import make_desc
    out = {"workload": "Gubser flow, conformal EOS, eta/s=0.2, tau_pi=5eta/(e+p), Milne 64x64x8, dx=deta=0.1, "
                       "tau0=1, dtau=0.01, RK2, %d steps" % steps}
    for mode in ("fast", "exact"):
        d = make_desc(64, 64, 8, fast_math=1 if mode == "fast" else 0)
        st = ic.gubser(d, 1.0, 0.01)
        h = capi.Hydro(d)
        h.set_launch(graph=1)
        h.upload(st)
        r, done = h.steps(1.0, 0.01, 2)
        assert r == 0, (r, done)
        h.upload(st)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r, done = h.steps(1.0, 0.01, steps)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        assert r == 0, (r, done)
        out[mode] = {"value": d.nx * d.ny * d.nz * steps / dt, "unit": "cell-updates/s",
                     "ms_per_step": dt * 1000.0 / steps}
        h.close()
    return out


LEGS = ("c4", "c2", "c1", "c3", "c5", "small", "cpu")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps of the headline ensemble (each way)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--legs", default=",".join(LEGS),
                    help="comma list of legs to run: c4 (headline ensemble), c2 (200x200x64 Glauber), c1 (Gubser "
                         "64x64x8), c3 (256x256x96 hot spots, KT and SHASTA), c5 (adjoint), small (small-event "
                         "ensemble), cpu (CPU baseline)")
    ap.add_argument("--ensemble-events", type=int, default=64)
    ap.add_argument("--ensemble-modes", default="plain_default,batched,plain",
                    help="ways the headline ensemble is stepped (profiling runs take batched alone)")
    ap.add_argument("--c2-steps", type=int, default=20)
    ap.add_argument("--exact-steps", type=int, default=3, help="timed steps of the exact-arithmetic c2 leg")
    ap.add_argument("--c3-steps", type=int, default=400)
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-eos", choices=["s95p", "conformal"], default="s95p")
    ap.add_argument("--adjoint-exact", action="store_true", help="run the adjoint leg in exact arithmetic")
    ap.add_argument("--small-events", type=int, default=128)
    ap.add_argument("--copy-probe", action="store_true",
                    help="at the end, run one plain 1 GiB device copy so a counter run can scale "
                         "FETCH_SIZE/WRITE_SIZE by a kernel of known traffic")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (events shard by event, a "
              "single event by x-y slabs; see the design notes). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    legs = set(x for x in args.legs.split(",") if x)
    if args.no_cpu:
        legs.discard("cpu")

    import torch
    assert torch.cuda.is_available(), "needs a GPU"
    torch.cuda.set_device(0)

    out = {"metric": METRIC, "value": None, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded hot-spot and optical-Glauber initial states generated on the host)"}
    if "c4" in legs:
        ens = bench_ensemble(args.ensemble_events, args.steps, warmup=args.warmup, kernel_times=True,
                             modes=tuple(args.ensemble_modes.split(",")))
        out["value"] = ens["batched"]["value"]
        out["ms_per_step"] = ens["batched"]["ms_per_step"]
        out["config"] = {"workload": ens["workload"], "events": ens["events"], "interior_cells": ens["interior_cells"],
                         "launch": ens["launch"], "parallelism": "single GPU"}
        for k in ("plain", "plain_default", "batched_over_plain", "batched_over_plain_default"):
            if k in ens:
                out[k] = ens[k]
        rep = ens["kernel_times"]
        out["roofline"] = kernel_roofline(rep, 3, ens["interior_cells"], ens["launch"]["fusion"] > 0,
                                          traffic=ENSEMBLE_TRAFFIC.get(max(rep, key=lambda k: rep[k][0])),
                                          counters=ensemble_counters)
        out["roofline"]["timing"] = ("HIP events around every launch of two batched steps replayed without the "
                                     "graph after the timed steps (launch %s)" % ens["kernel_launch"])
        tot = sum(v[0] for v in rep.values())
        out["kernel_time_share"] = {k: round(v[0] / tot, 4) for k, v in rep.items() if v[1]}
        out["ensemble"] = {k: ens[k] for k in ("ic_seconds", "steps", "warmup")}
    if "c2" in legs:
        out["glauber_200x200x64"] = bench_c2(args.c2_steps, args.warmup, args.exact_steps)
    if "c1" in legs:
        out["gubser_64x64x8"] = bench_c1()
    if "c3" in legs and args.c3_steps > 0:
        out["grid_256x256x96"] = bench_c3(args.c3_steps)
    if "c5" in legs and args.adjoint_steps > 0:
        adj = bench_adjoint(args.adjoint_steps, eos=args.adjoint_eos, fast=0 if args.adjoint_exact else 1)
        out["adjoint"] = adj
        out["adjoint_over_forward"] = adj["adjoint_over_forward"]
    if "small" in legs and args.small_events > 0:
        # many small events: the launch-bound regime the batched launch is for
        out["ensemble_small"] = bench_ensemble(args.small_events, 4 * args.steps, warmup=args.warmup,
                                               side=(24, 64), eta=(8, 24))
    if "cpu" in legs:
        out["cpu_baseline"] = cpu_baseline()
    if args.copy_probe:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    print(json.dumps(out))


if __name__ == "__main__":
    main()
