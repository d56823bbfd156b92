"""The C-ABI library loads and exports every entry point include/viral.h
declares (no compute call: this runs without a GPU)."""
import ctypes
import os
import re

from viral_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "viral.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(viral_[a-z_0-9]+)\s*\(", src)))


def test_header_lists_match():
    assert declared_symbols() == sorted(capi.EXPORTS)


def test_library_exports_all_declared_symbols():
    lib = capi.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_desc_default_and_dims_without_gpu():
    from viral_amd.desc import ViralDesc, make_desc
    lib = capi.load()
    d = ViralDesc()
    lib.viral_desc_default(ctypes.byref(d), 64, 64, 8)
    ref = make_desc(64, 64, 8)
    for f, _ in ViralDesc._fields_:
        assert getattr(d, f) == getattr(ref, f), f
    assert lib.viral_cells(ctypes.byref(d)) == 70 * 70 * 14
    d.lbi, d.nz = 1, 1
    gx, gy, gz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib.viral_ghost_dims(ctypes.byref(d), ctypes.byref(gx), ctypes.byref(gy), ctypes.byref(gz))
    assert (gx.value, gy.value, gz.value) == (70, 70, 1)


def test_narrow_mirror_grid_refused_without_gpu():
    """Under the Gubser mirror quirk ghost layer h of U1, U2 copies the
    (2-h)-th cell from the face; with fewer than 3 interior cells that cell is
    a ghost the same fill writes, so viral_create refuses the grid (it checks
    the descriptors before touching the device)."""
    from viral_amd.desc import make_desc, Q_DEFAULT, Q_GUBSER_MIRROR, ERR_ARG
    lib = capi.load()
    h = ctypes.c_void_p()
    for nx, ny in ((2, 9), (9, 1)):
        d = make_desc(nx, ny, 1, lbi=1, quirks=Q_DEFAULT | Q_GUBSER_MIRROR)
        assert lib.viral_create(ctypes.byref(d), 1, ctypes.byref(h)) == ERR_ARG
        assert not h.value
