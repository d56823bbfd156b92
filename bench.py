"""Benchmark of the MI355X viscous hydro step (one GPU).

Workload (N=1): the second benchmark configuration -- viscous Kurganov-Tadmor
step with the s95p EOS on a smooth optical-Glauber event (Woods-Saxon
R=6.62 fm, a=0.546 fm, b=7 fm, peak e=30 GeV/fm^3, eta plateau 1.5 with
Gaussian fall-off 1.0), Milne grid 200x200x64, dx=deta=0.1, tau0=0.6 fm/c,
eta/s=0.08, tau_pi=5 eta/(e+p), RK3 with dtau=0.002 fm/c as the reference's
s95p build (NEWEST10/Makeeos:5, maindef.h:255), FP64.  A step is one full RK
step of the whole grid; inputs are resident in HBM before timing starts.

Prints one JSON line.  `value` is interior cell-updates per second.
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
#     written = 36 doubles = 288 B; exact per-face kernel: the same 36 doubles
#     per face (15 face fluxes written)
#   source (per interior cell): e, P, PI, u (4), prevu (4), pi (10) read,
#     15 sources written = 36 doubles = 288 B
#   update (per interior cell): 45 partial results, 29 state fields written,
#     plus per stage kind the RK registers: first stage writes Var0, Var1
#     (30); RK2 final reads Var0, Var1 (30); RK3 middle reads Var0, Var1 and
#     writes Var2 (30 + 15); RK3 final reads Var0, Var2 (30); + 4 B
#     iteration count, 1 B flags (+1 B read after the first stage).
#     Sources in their own kernel: 15 sources + e, u, prevu (9) read, and the
#     first stage reads the 15 evolved fields.  Sources fused in: the
#     source's 21 inputs read, and the first stage reads T^{tau mu} (4; pi
#     and PI are among the 21).  Averaged over the stages.
def update_doubles(rk, fused):
    base = 45 + 29 + (21 if fused else 15 + 9)
    first = (4 if fused else 15) + 30
    kinds = {1: [first], 2: [first, 30], 3: [first, 30 + 15, 30]}[rk]
    return [base + k for k in kinds]


def kernel_bytes(name, d, fused):
    faces = [(d.nx + 1) * d.ny * d.nz, d.nx * (d.ny + 1) * d.nz, d.nx * d.ny * (d.nz + 1)]
    cells = d.nx * d.ny * d.nz
    if name.startswith("flux_"):
        q = {"flux_x": 0, "flux_y": 1, "flux_eta": 2}[name]
        return (cells if d.fast_math else faces[q]) * 288.0
    if name == "source":
        return cells * 288.0
    if name == "update":
        per = update_doubles(d.rk, fused)
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
    "update_fused": 3.784e9,
    "flux_x": 1.114e9,
    "flux_y": 1.098e9,
    "flux_eta": 0.965e9,
    "k0_broadcast": 0.168e9,
}


def measured_traffic(name, d, fused):
    if (d.nx, d.ny, d.nz, d.rk, d.fast_math, d.eos_table) == (200, 200, 64, 3, 1, 1):
        if name == "update":
            return MEASURED_TRAFFIC.get("update_fused") if fused else None
        return MEASURED_TRAFFIC.get(name)
    return None


def cpu_baseline(seconds_target=20.0):
    """The CPU oracle (a one-thread restatement of the reference step) on a
    bounded sample of the same workload: the central 16-slice eta slab of the
    200x200 Glauber grid, one RK3 step."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ctypes
    from viral_amd import ic
    d = workload_desc(fast=0)
    d.nz = 16
    st = ic.glauber(d, 0.6, 0.002, b=7.0, eos_p=eos_p_vec(d))
    o = oracle_ctypes.Oracle(d)
    o.set_state(st)
    t0 = time.perf_counter()
    steps = 0
    while True:
        o.step(0.6 + steps * 0.002, 0.002)
        steps += 1
        if time.perf_counter() - t0 > seconds_target or steps >= 3:
            break
    dt = time.perf_counter() - t0
    cells = d.nx * d.ny * d.nz
    return {"value": cells * steps / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "%d RK3 step(s) of the central 200x200x16 eta slab of the same Glauber s95p grid, "
                      "CPU oracle (C, -O2, one thread)" % steps}


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


def bench_ensemble(n_events, steps, tau0=0.6, ts=0.002, side=(96, 256), eta=(32, 64)):
    """Plain (one launch per event and kernel phase) vs batched (one launch per
    phase over all events, hipGraph replay, autotuned launch configuration)
    on the same uneven hotspot ensemble in one context; both synchronise once
    per step."""
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

    def timed(n):
        t = [tau0]
        r = h.step(t[0], ts)
        assert r == 0, r
        t[0] += ts
        t0 = time.perf_counter()
        for _ in range(n):
            r = h.step(t[0], ts)
            assert r == 0, r
            t[0] += ts
        return time.perf_counter() - t0

    out = {"workload": "%d hotspot events (200 spots, width 0.4 fm, b~U(0,12) fm), transverse side ~U{%d..%d}, "
                       "eta cells ~U{%d..%d}, viscous KT, s95p table, eta/s=0.08, RK3, dtau=%g"
                       % (n_events, side[0], side[1], eta[0], eta[1], ts),
           "events": n_events, "interior_cells": cells, "ic_seconds": round(t_ic, 1)}
    h.set_launch(batched=0, graph=0, blocks_per_cu=0, fusion=0, xcd=0)
    dt = timed(steps)
    out["plain"] = {"value": cells * steps / dt, "ms_per_step": dt * 1000.0 / steps}
    h.set_launch(batched=1, graph=1)
    h.autotune(reps=1)
    out["launch"] = h.launch().as_dict()
    dt = timed(steps)
    out["batched"] = {"value": cells * steps / dt, "ms_per_step": dt * 1000.0 / steps}
    out["batched_over_plain"] = out["batched"]["value"] / out["plain"]["value"]
    h.close()
    return out


def bench_c3(steps, tau0=0.6, ts=0.002):
    """The 256x256x96 grid of the fluctuating-event configuration (s95p EOS,
    table in fast mode, eta/s = 0.12, tau_pi = 5 eta/(e+p), RK2): the viscous
    KT step (autotuned) on 200 Gaussian hot spots of width 0.4 fm from two
    Woods-Saxon nuclei (b ~ U(0,10) fm, seed 3), with the update kernel's HBM
    figure; and the SHASTA step on the smooth Glauber state (b = 7 fm) of the
    same grid.  SHASTA's step breaks down in the dilute periphery of both
    states (e <= 0 after 4 steps on the hot spots, 10 on the Glauber state,
    exact arithmetic included), so its leg runs at most 2 + 5 steps."""
    import torch
    from viral_amd import capi, ic
    from viral_amd.desc import make_desc, EOS_S95P, ETA_TIMES_S, TPI_OVER_EP, SHASTA
    out = {"workload": "256x256x96 Milne grid, s95p table, eta/s=0.12, tau_pi=5eta/(e+p), RK2, dtau=%g, fast "
                       "arithmetic; KT on 200 hot spots (width 0.4 fm, b~U(0,10) fm, seed 3), SHASTA on the smooth "
                       "Glauber state (b=7 fm)" % ts}
    for scheme in ("kt", "shasta"):
        d = make_desc(256, 256, 96, dx=0.1, dy=0.1, deta=0.1, rk=2, eos=EOS_S95P, eta_law=ETA_TIMES_S, eta_coef=0.12,
                      tpi_law=TPI_OVER_EP, tpi_coef=5.0, fast_math=1, eos_table=1,
                      scheme=SHASTA if scheme == "shasta" else 0)
        dref = d.copy()
        dref.fast_math = dref.eos_table = 0
        eosp = (lambda x: capi.eos_eval(dref, x)[0].reshape(x.shape))
        if scheme == "kt":
            st = ic.hotspots(d, tau0, ts, seed=3, eos_p=eosp)
        else:
            st = ic.glauber(d, tau0, ts, b=7.0, eos_p=eosp)
            steps = min(steps, 5)
        h = capi.Hydro(d)
        h.upload(st)
        tau = tau0
        for _ in range(2):
            r = h.step(tau, ts)
            assert r == 0, r
            tau += ts
        if scheme == "kt":
            h.autotune(reps=1)
        h.timers(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            r = h.step(tau, ts)
            assert r == 0, r
            tau += ts
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        rep = h.timer_report()
        h.timers(False)
        cells = d.nx * d.ny * d.nz
        leg = {"value": cells * steps / dt, "unit": "cell-updates/s", "ms_per_step": dt * 1e3 / steps, "steps": steps,
               "launch": h.launch().as_dict(),
               "kernel_ms_per_launch": {k: round(v[0] / v[1], 4) for k, v in rep.items() if v[1]}}
        if scheme == "kt":
            fused = leg["launch"]["fusion"] > 0
            ms, nl = rep["update"]
            leg["update_hbm"] = {"algorithmic_GBs": kernel_bytes("update", d, fused) / (ms / nl * 1e-3) / 1e9,
                                 "frac": kernel_bytes("update", d, fused) / (ms / nl * 1e-3) / 1e9 / HBM_PEAK_GBS}
        out[scheme] = leg
        h.close()
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
    download, 2 x 29 fields over PCIe; adjoint_device_ms: its device part,
    HIP events around the recompute and reverse kernels),
    first with every stage state kept during the forward run (every = 0, no
    recompute: 63 GB of HBM for 200 RK2 steps), then with checkpoints every
    `every` steps (the adjoint recomputes the stage states between them)."""
    import numpy as np
    import torch
    from viral_amd import capi, ic, autograd
    from viral_amd.desc import make_desc, F_EN, EOS_S95P, EOS_CONFORMAL
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
            seed = fin.grad.numpy()
        h.timers(True)
        t0 = time.perf_counter()
        grad, pg = h.adjoint(seed)
        t_adj = time.perf_counter() - t0
        dev_ms = h.timer_report()["adjoint_device"][0]
        h.timers(False)
        h.adjoint_end()
        runs[ev] = (t_fwd, t_adj, grad, pg, dev_ms)
    h.close()
    t_fwd, t_adj, grad, pg, dev = runs[0]
    c_fwd, c_adj, c_grad, c_pg, c_dev = runs[every]
    assert np.array_equal(grad, c_grad) and np.array_equal(pg, c_pg), "stored and recomputed stage states disagree"
    return {"workload": "adjoint of %d viscous KT steps, smooth Glauber b=8 fm, Milne 128x128x32, %s EOS, "
                        "eta/s=0.08, tau_pi=5eta/(e+p), RK2, dtau=%g, every stage state kept in HBM, %s arithmetic; "
                        "d eps_2 / d(initial state, eta/s, tau_pi coefficient)"
                        % (steps, "s95p (table)" if s95 and fast else eos, ts, "fast" if fast else "exact"),
            "steps": steps, "forward_ms": t_plain * 1e3, "forward_recorded_ms": t_fwd * 1e3,
            "adjoint_ms": t_adj * 1e3, "adjoint_over_forward": t_adj / t_plain,
            "adjoint_device_ms": dev, "adjoint_device_over_forward": dev / (t_plain * 1e3),
            "checkpointed": {"every": every, "forward_recorded_ms": c_fwd * 1e3, "adjoint_ms": c_adj * 1e3,
                             "adjoint_over_forward": c_adj / t_plain, "adjoint_device_ms": c_dev,
                             "adjoint_device_over_forward": c_dev / (t_plain * 1e3), "same_gradient_bits": True},
            "eps2": float(eps2), "d_eps2_d_eta_coef": float(pg[0]), "d_eps2_d_tpi_coef": float(pg[1]),
            "d_eps2_d_e_norm": float(np.linalg.norm(grad[F_EN]))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--ensemble-events", type=int, default=64,
                    help="events of the batched-ensemble leg (0 skips it)")
    ap.add_argument("--ensemble-steps", type=int, default=5)
    ap.add_argument("--small-events", type=int, default=128,
                    help="events of the small-event ensemble leg (transverse side 24..64, eta 8..24; 0 skips it)")
    ap.add_argument("--adjoint-steps", type=int, default=200,
                    help="steps of the adjoint leg (0 skips it)")
    ap.add_argument("--adjoint-eos", choices=["s95p", "conformal"], default="s95p")
    ap.add_argument("--adjoint-exact", action="store_true", help="run the adjoint leg in exact arithmetic")
    ap.add_argument("--c3-steps", type=int, default=10,
                    help="timed steps of the 256x256x96 hot-spot grid legs (KT and SHASTA); 0 skips them")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the timed steps, run one plain 1 GiB device copy so a counter run can scale "
                         "FETCH_SIZE/WRITE_SIZE by a kernel of known traffic")
    ap.add_argument("--exact", action="store_true",
                    help="time the exact-arithmetic mode (bitwise equal to the CPU oracle) instead of fast mode")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (events shard by event, a "
              "single event by x-y slabs; see the design notes). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from viral_amd import capi, ic

    assert torch.cuda.is_available(), "needs a GPU"
    torch.cuda.set_device(0)
    d = workload_desc(fast=0 if args.exact else 1)
    tau0, ts = 0.6, 0.002
    st = ic.glauber(d, tau0, ts, b=7.0, eos_p=eos_p_vec(d))
    h = capi.Hydro(d)
    h.upload(st)
    tau = tau0
    for _ in range(args.warmup):
        r = h.step(tau, ts)
        assert r == 0, r
        tau += ts
    tiles = h.autotune(reps=2) if d.fast_math else None   # launch configuration, on this state
    launch = h.launch().as_dict()
    torch.cuda.synchronize()
    h.timers(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = h.step(tau, ts)
        assert r == 0, r
        tau += ts
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    rep = h.timer_report()
    h.timers(False)
    if args.copy_probe:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    cells = d.nx * d.ny * d.nz
    value = cells * args.steps / dt
    ms_step = dt * 1000.0 / args.steps

    # dominant kernel and its roofline figure
    dom = max(rep, key=lambda k: rep[k][0])
    tot_ms, launches = rep[dom]
    avg_s = tot_ms / 1000.0 / max(launches, 1)
    fused = launch["fusion"] > 0
    bytes_launch = kernel_bytes(dom, d, fused)
    achieved = bytes_launch / avg_s / 1e9
    kernel_share = {k: round(v[0] / max(sum(x[0] for x in rep.values()), 1e-12), 4) for k, v in rep.items() if v[1]}
    bytes_per_update = {1: 336, 2: 784, 3: 1232}[d.rk]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (optical-Glauber initial state generated on the host)",
        "config": {"workload": "viscous KT, s95p EOS, Glauber b=7 fm, Milne 200x200x64, dx=deta=0.1, "
                               "tau0=0.6, dtau=0.002, eta/s=0.08, tau_pi=5eta/(e+p), RK3",
                   "grid": [d.nx, d.ny, d.nz], "rk": d.rk, "fast_math": d.fast_math, "eos_table": d.eos_table,
                   "flux_tiles": tiles, "launch": launch,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom, d, fused),
                     "avg_launch_ms": avg_s * 1000.0, "launches": launches,
                     "algorithmic_bytes_per_launch": bytes_launch},
        "step_hbm": {"algorithmic_bytes_per_cell_update": bytes_per_update,
                     "achieved_GBs": value * bytes_per_update / 1e9,
                     "frac": value * bytes_per_update / 1e9 / HBM_PEAK_GBS},
        "kernel_time_share": kernel_share,
    }
    h.close()
    if args.ensemble_events > 0:
        out["ensemble"] = bench_ensemble(args.ensemble_events, args.ensemble_steps)
    if args.small_events > 0:
        # many small events: the launch-bound regime the batched launch is for
        out["ensemble_small"] = bench_ensemble(args.small_events, 4 * args.ensemble_steps, side=(24, 64), eta=(8, 24))
    if args.c3_steps > 0:
        out["grid_256x256x96"] = bench_c3(args.c3_steps)
    if args.adjoint_steps > 0:
        out["adjoint"] = bench_adjoint(args.adjoint_steps, eos=args.adjoint_eos, fast=0 if args.adjoint_exact else 1)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
