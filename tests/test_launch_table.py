"""Host-side checks of the launch layer's tile arithmetic (no GPU): tile
counts of every kernel phase against the cells and faces each phase covers,
and the XCD tile order being a permutation."""
import math

import numpy as np
import pytest

from viral_amd import capi
from viral_amd.desc import make_desc, SHASTA, Q_DEFAULT, Q_GHOST_EIGEN

BLK = 256


def blocks(n):
    return (n + BLK - 1) // BLK


@pytest.mark.parametrize("nx,ny,nz,lbi", [(64, 64, 8, 0), (17, 9, 5, 0), (12, 10, 1, 1), (200, 200, 64, 0)])
def test_cell_and_ghost_phases(nx, ny, nz, lbi):
    d = make_desc(nx, ny, nz, lbi=lbi)
    cells = nx * ny * nz
    assert capi.tile_count(d, capi.PHASE_CELLS) == blocks(cells)
    assert capi.tile_count(d, capi.PHASE_K0) == (0 if lbi else blocks(cells))
    B = 3
    ghosts = 2 * B * (ny * nz + nx * nz + (0 if lbi else nx * ny))
    assert capi.tile_count(d, capi.PHASE_GHOST) == 28 * blocks(ghosts)
    d.quirks = Q_GHOST_EIGEN  # no k0 broadcast
    assert capi.tile_count(d, capi.PHASE_K0) == 0


def test_exact_and_shasta_flux_phases():
    d = make_desc(17, 9, 5)
    faces = [(18 * 9 * 5), (17 * 10 * 5), (17 * 9 * 6)]
    for q in range(3):
        assert capi.tile_count(d, capi.PHASE_FLUX_X + q) == blocks(faces[q])
    d.scheme = SHASTA
    for q in range(3):
        assert capi.tile_count(d, capi.PHASE_FLUX_X + q) == blocks(17 * 9 * 5)
    d2 = make_desc(12, 10, 1, lbi=1)
    assert capi.tile_count(d2, capi.PHASE_FLUX_ETA) == 0


@pytest.mark.parametrize("tp,tk", [(64, 16), (32, 8), (70, 14), (8, 1), (256, 4)])
def test_fast_flux_tiles_cover_every_face_once(tp, tk):
    """A tile holds tk lines and tp cells of each, of which the tp-5 middle
    positions own a face; tiles along a line step by tp-5 faces."""
    d = make_desc(200, 200, 64, fast_math=1)
    n = [200, 200, 64]
    for q in range(3):
        other = [n[a] for a in range(3) if a != q]
        lines = other[0] * other[1]
        faces = n[q] + 1
        segs = math.ceil(faces / (tp - 5))
        assert capi.tile_count(d, capi.PHASE_FLUX_X + q, tp, tk) == math.ceil(lines / tk) * segs
        # the faces the segments own cover [lo-1, hi-1] exactly once
        owned = np.concatenate([np.arange(s * (tp - 5), min((s + 1) * (tp - 5), faces)) for s in range(segs)])
        assert np.array_equal(owned, np.arange(faces))
    assert capi.tile_count(d, capi.PHASE_FLUX_X, 7, 4) == -1


@pytest.mark.parametrize("total", [1, 7, 8, 9, 63, 64, 1000, 4097])
def test_xcd_order_is_a_permutation_with_contiguous_xcd_ranges(total):
    m = np.array([capi.xcd_tile(w, total) for w in range(total)])
    assert sorted(m.tolist()) == list(range(total))
    for x in range(8):
        mine = m[np.arange(total) % 8 == x]
        if len(mine):
            assert np.array_equal(mine, np.arange(mine[0], mine[0] + len(mine)))
    assert capi.xcd_tile(total, total) == -1
