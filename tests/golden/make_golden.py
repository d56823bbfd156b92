"""Writes tests/golden/ref_boxes.npz from the reference's own code.

Runs the oracle/_ref harness binaries (the reference's NEWEST10 step headers
compiled in place, see oracle/build_ref.sh) on seeded states and keeps, per
profile, a 9x9x9 box of the input state (9x9x1 for the boost-invariant
profile) together with the reference's CalcSource and KT Result on the
3x3x3 core of the box.  The core cells depend only on the box (the KT
stencil reaches three cells, the velocity gradients two), so the oracle can
recompute them on a 3x3x3-interior grid whose ghosts are the box halo.

It also writes tests/golden/ref_mirror.npz: the GUBSER profile's boundary()
run by the reference's own code on its 6x6 rank grid (the harness under
MPICH's mpiexec, mode 6), keeping u^1 and u^2 on the four edge bands (three
ghost rows and the three interior rows they mirror) before and after the
fill, the left bands from rank (0,0) and the right ones from rank (5,5).

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
import refdata  # noqa: E402

CASES = [("ref_kt", 21, 1.3, 0.01, (17, 22, 40)),
         ("ref_kt", 22, 0.9, 0.004, (8, 30, 90)),
         ("ref_kt_ginit", 23, 0.6, 0.002, (12, 10, 50)),
         ("ref_kt_gubser", 24, 1.0, 0.005, (40, 61, 0))]


def main():
    out = {}
    for q, (name, seed, tau, ts, (ci, cj, ck)) in enumerate(CASES):
        r = refdata.run(name, seed, tau, ts)
        d = r["desc"]
        gx, gy, gz = d.ghost_dims()
        st = r["state"].reshape((-1, gx, gy, gz))
        S = r["source"].reshape((15, gx, gy, gz))
        R = r["result"].reshape((15, gx, gy, gz))
        si = slice(ci - 3, ci + 6)
        sj = slice(cj - 3, cj + 6)
        if d.lbi:
            sk, core_k = slice(0, 1), slice(0, 1)
        else:
            sk, core_k = slice(ck - 3, ck + 6), slice(ck, ck + 3)
        core = (slice(ci, ci + 3), slice(cj, cj + 3), core_k)
        p = "c%d_" % q
        out[p + "profile"] = np.array(name)
        out[p + "tau_ts"] = np.array([r["tau"], r["ts"]])
        out[p + "state"] = st[:, si, sj, sk].copy()
        out[p + "source"] = S[(slice(None),) + core].copy()
        out[p + "result"] = R[(slice(None),) + core].copy()
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "ref_boxes.npz"), **out)
    print("wrote", len(CASES), "cases")
    mirror_fixture()


def mirror_bands(a, d):
    """the four 6-row edge bands of a (gx, gy) field over the interior
    columns: x-left, x-right (rows i), y-left, y-right (columns j)"""
    gx, gy, _ = d.ghost_dims()
    a = a.reshape(gx, gy)
    B = 3
    return dict(xl=a[0:2 * B, B:gy - B].copy(), xr=a[gx - 2 * B:gx, B:gy - B].copy(),
                yl=a[B:gx - B, 0:2 * B].copy(), yr=a[B:gx - B, gy - 2 * B:gy].copy())


def mirror_fixture():
    from viral_amd.desc import F_U1, F_U2
    ranks = refdata.run_ranks("ref_kt_gubser", 31, 1.0, 0.005, 36)
    lo, hi = ranks[(0, 0)], ranks[(5, 5)]
    d = lo["desc"]
    out = {"tau_ts": np.array([lo["tau"], lo["ts"]])}
    for name, f in (("u1", F_U1), ("u2", F_U2)):
        before = mirror_bands(lo["state"][f], d)
        after_lo = mirror_bands(lo["after_boundary"][f], d)
        after_hi = mirror_bands(hi["after_boundary"][f], d)
        for side in ("xl", "xr", "yl", "yr"):
            out["%s_%s_before" % (name, side)] = before[side]
            src = after_lo if side.endswith("l") else after_hi
            out["%s_%s_after" % (name, side)] = src[side]
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "ref_mirror.npz"), **out)
    print("wrote the GUBSER mirror bands")


if __name__ == "__main__":
    main()
