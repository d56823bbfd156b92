"""GPU: viral_check_root_noalloc, the allocation-free CheckRoot, against
viral_check_root through the C ABI: equal maxima, averages and cells, bit for
bit, on two calls in a row (the second on a smaller grid, so it reads fewer
block partials than the first left behind), on Milne and Cartesian events;
the calls leave the device's free memory unchanged."""
import ctypes as C

import numpy as np
import pytest
import torch

from viral_amd import capi, ic
from viral_amd.desc import make_desc, CARTESIAN, OK

pytestmark = pytest.mark.gpu


def call(h, name, tau=0.904):
    """(max, avg, ijk, nstat) of viral_check_root_noalloc or viral_check_root"""
    mx, avg = np.zeros(9), np.zeros(9)
    ijk = np.zeros(27, dtype=np.int32)
    ns = C.c_int()
    fn = h.L.viral_check_root_noalloc if name == "noalloc" else h.L.viral_check_root
    r = fn(h.h, 0, tau, mx.ctypes.data_as(C.POINTER(C.c_double)), avg.ctypes.data_as(C.POINTER(C.c_double)),
           ijk.ctypes.data_as(C.POINTER(C.c_int)), C.byref(ns))
    assert r == OK, h.L.viral_last_error(h.h)
    return mx, avg, ijk, ns.value


def same(a, b):
    ns = a[3]
    assert b[3] == ns
    return (np.array_equal(a[0][:ns].view(np.int64), b[0][:ns].view(np.int64)) and
            np.array_equal(a[1][:ns].view(np.int64), b[1][:ns].view(np.int64)) and
            np.array_equal(a[2][:3 * ns], b[2][:3 * ns]))


def stepped(d, seed):
    h = capi.Hydro(d)
    h.upload(ic.random_smooth(d, seed=seed, tau=0.9))
    assert h.step(0.9, 0.004) == OK
    return h


@pytest.mark.parametrize("coords", ["milne", "cartesian"])
def test_noalloc_check_root_equals_check_root_on_two_calls(coords):
    kw = dict(coords=CARTESIAN) if coords == "cartesian" else {}
    big = make_desc(60, 58, 40, **kw)     # 139,200 interior cells: 512 blocks of partials
    small = make_desc(9, 8, 5, **kw)      # 360 interior cells: 2 blocks
    print("case check_root_noalloc %s" % coords, flush=True)
    hb, hs = stepped(big, 3), stepped(small, 4)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    na_b = call(hb, "noalloc")
    na_s = call(hs, "noalloc")    # the second call, on the smaller grid
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0, "viral_check_root_noalloc allocated device memory"
    assert na_b[3] == na_s[3] == (6 if coords == "milne" else 9)
    assert same(na_b, call(hb, "alloc"))
    assert same(na_s, call(hs, "alloc"))
    # the big grid again, after the small one overwrote the front of the scratch
    assert same(na_b, call(hb, "noalloc"))
