"""Writes tests/golden/ref_boxes.npz from the reference's own code.

Runs the oracle/_ref harness binaries (the reference's NEWEST10 step headers
compiled in place, see oracle/build_ref.sh) on seeded states and keeps, per
profile, a 9x9x9 box of the input state (9x9x1 for the boost-invariant
profile) together with the reference's CalcSource and KT Result on the
3x3x3 core of the box.  The core cells depend only on the box (the KT
stencil reaches three cells, the velocity gradients two), so the oracle can
recompute them on a 3x3x3-interior grid whose ghosts are the box halo.

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


if __name__ == "__main__":
    main()
