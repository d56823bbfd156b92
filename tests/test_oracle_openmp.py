"""The all-core build of the oracle (-O3 -march=native -fopenmp, the CPU
baseline's build) gives the serial test build's results bit for bit: state,
Newton iteration counts, flags and first failing cell, on the Milne KT (RK3),
s95p, SHASTA, Cartesian and bulk paths, with four threads."""
import os
import subprocess

import numpy as np
import pytest

import oracle_ctypes
from viral_amd import ic
from viral_amd.desc import make_desc, EOS_S95P, SHASTA, CARTESIAN, OK

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CASES = {"kt_rk3": dict(nx=26, ny=24, nz=10, rk=3), "s95p": dict(nx=20, ny=18, nz=9, eos=EOS_S95P),
         "shasta": dict(nx=20, ny=18, nz=9, scheme=SHASTA), "cartesian": dict(nx=18, ny=16, nz=9, rk=3, coords=CARTESIAN),
         "bulk": dict(nx=18, ny=16, nz=9, bulk=1)}


@pytest.fixture(scope="module")
def omp_lib(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("oracle_omp") / "liboracle_omp.so")
    r = subprocess.run(["gcc", "-O3", "-march=native", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", "-o", path,
                        os.path.join(ROOT, "oracle", "oracle.c"), "-lm"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        pytest.skip("no OpenMP build here: " + r.stdout.decode()[-300:])
    return oracle_ctypes.load(path)


@pytest.mark.parametrize("name", sorted(CASES))
def test_openmp_build_is_bitwise_serial(name, omp_lib, monkeypatch):
    monkeypatch.setenv("OMP_NUM_THREADS", "4")
    d = make_desc(**CASES[name])
    Ls = oracle_ctypes.lib()
    eosp = (lambda e: np.array([Ls.or_eos_p(d, float(x)) for x in np.atleast_1d(e)]).reshape(np.shape(e))) \
        if d.eos else None
    st = ic.random_smooth(d, seed=5, tau=0.9, e_scale=30.0 if d.eos else 4.0, eos_p=eosp,
                          bulk_frac=0.02 if d.bulk else 0.0)
    res = []
    for L in (Ls, omp_lib):
        o = oracle_ctypes.Oracle(d, L)
        o.set_state(st)
        for s in range(2):
            assert o.step(0.9 + s * 0.004, 0.004) == OK
        res.append((o.state(), [o.iters(q) for q in range(3)], o.flags(), o.first_bad()))
    a, b = res
    assert np.array_equal(a[0].view(np.int64), b[0].view(np.int64))
    assert all(np.array_equal(x, y) for x, y in zip(a[1], b[1]))
    assert np.array_equal(a[2], b[2]) and a[3] == b[3]
