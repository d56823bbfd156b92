"""GPU: the host driver (viral_amd/viral_run) end to end.

The driver runs a few steps of a profile, writing the reference's output
files on its print schedule (NEWEST10/main.cpp:176-186) and printing the
per-step CheckRoot and MaxTempGev lines.  The same initial state (the
driver's --dump-state) is stepped through the Python binding of the same C
ABI; every file the driver wrote must hold exactly the records of that state
at that tau, and the printed diagnostics must be those of the state.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from viral_amd import capi
from viral_amd.desc import make_desc, NFIELDS, ETA_S_OVER_4PI, TPI_OVER_P, ETA_TIMES_S, TPI_OVER_EP, Q_DEFAULT, \
    Q_GUBSER_MIRROR, OK

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "viral_amd", "viral_run")

CASES = {
    "default": (["--profile", "default", "--nx", "24", "--ny", "20", "--nz", "11", "--max-steps", "6",
                 "--pfreq", "0.004"],
                make_desc(24, 20, 11, eta_law=ETA_S_OVER_4PI, eta_coef=1.0, tpi_law=TPI_OVER_P, tpi_coef=1.5),
                0.6, 0.002, 1, 0),
    "gubser": (["--profile", "gubser", "--nx", "60", "--ny", "60", "--max-steps", "5", "--pfreq", "0.01"],
               make_desc(60, 60, 1, lbi=1, dx=0.05, dy=0.05, eta_law=ETA_TIMES_S, eta_coef=0.2,
                         tpi_law=TPI_OVER_EP, tpi_coef=5.0, quirks=Q_DEFAULT | Q_GUBSER_MIRROR),
               1.0, 0.005, 2, 1),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_driver_files_and_reports_match_the_binding(tmp_path, name):
    args, d, tau0, ts, freq, xy = CASES[name]
    assert os.path.exists(EXE), "viral_amd/viral_run not built"
    print("case driver %s" % name, flush=True)
    dump = tmp_path / "init.bin"
    r = subprocess.run([EXE, "--out", str(tmp_path), "--dump-state", str(dump)] + args, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-3000:]
    files = sorted(f for f in os.listdir(tmp_path) if f.startswith("tau"))
    st = np.fromfile(dump, dtype=np.float64).reshape(NFIELDS, -1)
    h = capi.Hydro(d)
    h.upload(st)
    nsteps = int(args[args.index("--max-steps") + 1])
    taus = [tau0]
    t = tau0
    seen = set()
    name0 = "tau%2.3ffm.bin" % tau0
    assert name0 in files
    assert np.array_equal(np.fromfile(tmp_path / name0, dtype=np.float64), h.output(freq, 1, xy))
    seen.add(name0)
    maxtemps = [float(v) for v in re.findall(r"TempMax is (\d+\.\d+) MeV", r.stdout)]
    roots = [float(v) for v in re.findall(r"ROOT CHECK -- MAX (\S+) &", r.stdout)]
    assert len(maxtemps) == nsteps + 1 and len(roots) == nsteps + 1
    for s in range(nsteps):
        assert h.step(t, ts) == OK
        t = t + ts
        taus.append(t)
        fn = "tau%2.3ffm.bin" % t
        if fn in files:
            assert np.array_equal(np.fromfile(tmp_path / fn, dtype=np.float64), h.output(freq, 1, xy)), fn
            seen.add(fn)
        assert abs(maxtemps[s + 1] - round(1000 * h.max_temp_gev(), 4)) <= 1.01e-4
        mx, avg, at = h.check_root(t)
        assert abs(roots[s + 1] - mx[0]) <= 1e-6 * abs(mx[0]) + 1e-300
    assert seen == set(files), (seen, files)
    # the first step's output and at least one print-interval output
    assert "tau%2.3ffm.bin" % (tau0 + ts) in seen and len(seen) >= 3


@pytest.mark.parametrize("name", ["bjorken", "bulktest"])
def test_driver_temp_file_matches_the_oracle(tmp_path, name):
    """The BJORKEN and BULKTEST profiles' temp.txt (NEWEST10/main.cpp:124-141):
    one "tau  value" line before every step, value the centre cell's T in GeV
    (BJORKEN) or PI (BULKTEST), printed as C++ streams print doubles (6
    significant digits).  The CPU oracle stepped from the same initial state
    must give the same lines; the state at the last line is the step count's."""
    import oracle_ctypes
    from viral_amd.desc import F_EN, F_BULK, GEVFM, ZETA_CONST, TPIB_CONST, TPI_CONST
    assert os.path.exists(EXE), "viral_amd/viral_run not built"
    print("case driver temp.txt %s" % name, flush=True)
    nsteps = 40
    dump = tmp_path / "init.bin"
    temp = tmp_path / "temp.txt"
    r = subprocess.run([EXE, "--profile", name, "--max-steps", str(nsteps), "--write", "0", "--checks", "0",
                        "--dump-state", str(dump), "--temp-file", str(temp)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-3000:]
    lines = temp.read_text().splitlines()
    assert len(lines) == nsteps
    kw = dict(eta_law=ETA_TIMES_S, eta_coef=0.2, tpi_law=TPI_CONST, tpi_coef=0.005)
    if name == "bulktest":
        kw.update(bulk=1, zeta_law=ZETA_CONST, zeta_coef=1.0, tpib_law=TPIB_CONST, tpib_coef=0.1)
    d = make_desc(10, 10, 41, **kw)
    st = np.fromfile(dump, dtype=np.float64).reshape(NFIELDS, -1)
    o = oracle_ctypes.Oracle(d)
    o.set_state(st)
    centre = np.ravel_multi_index((8, 8, 20), d.ghost_dims())
    L = oracle_ctypes.lib()
    t, ts = 0.6, 0.01
    for s in range(nsteps):
        if name == "bjorken":
            v = (1000 * (L.or_eos_T(d, float(o.state()[F_EN][centre])) * GEVFM)) / 1000
        else:
            v = float(o.state()[F_BULK][centre])
        assert lines[s] == "%g  %g" % (t, v), (s, lines[s], t, v)
        assert o.step(t, ts) == OK
        t = t + ts
