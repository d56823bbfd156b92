"""The host driver (viral_amd/viral_run) on the CPU: its initial states and
first output file (--dry-run touches no device).

- default profile: the reference's own initvar (NEWEST10/init.h:150-214, run
  by the oracle/_ref harness, mode 3) bit for bit, except Temp, which the
  library's EOS gives as sqrt(sqrt(e/c)) against the reference's
  pow(e/c, 0.25) (within 1 ulp);
- gubser: the semi-analytic viscous Gubser state of tests/gubser_viscous.py
  (an independent integration of the same equations);
- glauber: viral_amd.ic.glauber;
- the first output file: the oracle's write.h records of the dumped state.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_ctypes
import refdata
from viral_amd import ic
from viral_amd.desc import make_desc, NFIELDS, F_TEMP, EOS_S95P, ETA_TIMES_S, TPI_OVER_EP, Q_DEFAULT, \
    Q_GUBSER_MIRROR

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "viral_amd", "viral_run")


def run(tmp_path, *args):
    if not os.path.exists(EXE):
        pytest.skip("viral_amd/viral_run not built")
    dump = tmp_path / "init.bin"
    r = subprocess.run([EXE, "--dry-run", "--out", str(tmp_path), "--dump-state", str(dump)] + list(args),
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    return np.fromfile(dump, dtype=np.float64).reshape(NFIELDS, -1), r.stdout


def test_default_profile_initvar_matches_reference(tmp_path):
    if not refdata.available("ref_kt"):
        pytest.skip("oracle/_ref/ref_kt not built (reference tree absent)")
    st, _ = run(tmp_path, "--profile", "default")
    ref = refdata.run("ref_kt", 1, 0.6, 0.002, mode=3)
    assert st.shape == ref["state"].shape
    for f in range(NFIELDS):
        if f == F_TEMP:
            assert np.all(np.abs(st[f] - ref["state"][f]) <= 2.3e-16 * np.abs(ref["state"][f]))
        else:
            assert np.array_equal(st[f], ref["state"][f]), f


def test_gubser_profile_matches_semi_analytic_solution(tmp_path):
    from gubser_viscous import ViscousGubser
    st, _ = run(tmp_path, "--profile", "gubser", "--nx", "60", "--ny", "60")
    d = make_desc(60, 60, 1, lbi=1, dx=0.05, dy=0.05, eta_law=ETA_TIMES_S, eta_coef=0.2, tpi_law=TPI_OVER_EP,
                  tpi_coef=5.0, quirks=Q_DEFAULT | Q_GUBSER_MIRROR)
    want = ViscousGubser(eta_s=0.2).state(d, 1.0, 0.005)
    scale = np.abs(want).max(axis=1, keepdims=True) + 1e-300
    err = np.abs(st - want) / scale
    assert err.max() <= 1e-9, (err.max(axis=1))


def test_glauber_init_matches_python_generator(tmp_path):
    st, _ = run(tmp_path, "--profile", "ginit", "--nx", "30", "--ny", "30", "--nz", "21", "--tau0", "0.6")
    d = make_desc(30, 30, 21, eos=EOS_S95P)
    L = oracle_ctypes.lib()
    want = ic.glauber(d, 0.6, 0.002, b=7.0, e_peak_gev=30.0,
                      eos_p=lambda e: np.vectorize(lambda x: L.or_eos_p(d, x))(e))
    scale = np.abs(want).max(axis=1, keepdims=True) + 1e-300
    err = np.abs(st - want) / scale
    # T is the driver's s95p T(e), which the generator leaves at zero
    err[F_TEMP] = 0
    assert err.max() <= 1e-12, err.max(axis=1)


@pytest.mark.parametrize("args,freq,xy", [(["--profile", "default", "--nx", "20", "--ny", "10", "--nz", "9"], 1, 0),
                                          (["--profile", "gubser", "--nx", "40", "--ny", "30"], 2, 1)])
def test_first_output_file_is_the_records_of_the_state(tmp_path, args, freq, xy):
    st, _ = run(tmp_path, *args)
    files = [f for f in os.listdir(tmp_path) if f.startswith("tau")]
    assert len(files) == 1
    raw = np.fromfile(tmp_path / files[0], dtype=np.float64)
    nx, ny = int(args[args.index("--nx") + 1]), int(args[args.index("--ny") + 1])
    nz = 1 if xy else int(args[args.index("--nz") + 1])
    d = make_desc(nx, ny, nz, lbi=xy)
    o = oracle_ctypes.Oracle(d)
    o.set_state(st)
    assert np.array_equal(raw, o.output(freq, 1, xy))
    assert files[0] == ("tau0.600fm.bin" if not xy else "tau1.000fm.bin")


def test_strides_that_do_not_divide_are_refused(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("viral_amd/viral_run not built")
    r = subprocess.run([EXE, "--dry-run", "--out", str(tmp_path), "--nx", "21", "--freq", "2"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    assert r.returncode == 2 and "do not divide" in r.stdout
