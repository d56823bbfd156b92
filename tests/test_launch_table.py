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
    # the ghost-fill kernel covers the x and y faces; the update and k0
    # kernels write the eta faces of the cells they produce
    ghosts = 2 * B * (ny * nz + nx * nz)
    assert capi.tile_count(d, capi.PHASE_GHOST) == 28 * blocks(ghosts)
    d.quirks = Q_GHOST_EIGEN  # no k0 broadcast
    assert capi.tile_count(d, capi.PHASE_K0) == 0


def test_exact_and_shasta_flux_phases():
    d = make_desc(17, 9, 5)
    faces = [(18 * 9 * 5), (17 * 10 * 5), (17 * 9 * 6)]
    for q in range(3):
        assert capi.tile_count(d, capi.PHASE_FLUX_X + q) == blocks(faces[q])
    # SHASTA runs on the same LDS line tiles as fast KT, in both modes
    d.scheme = SHASTA
    n = [17, 9, 5]
    for q in range(3):
        other = [n[a] for a in range(3) if a != q]
        for tp, tk in ((64, 16), (8, 2)):
            assert capi.tile_count(d, capi.PHASE_FLUX_X + q, tp, tk) == \
                math.ceil(other[0] * other[1] / tk) * math.ceil(n[q] / (tp - 6))
    assert capi.tile_count(d, capi.PHASE_FLUX_X, 7, 4) == -1
    d2 = make_desc(12, 10, 1, lbi=1)
    assert capi.tile_count(d2, capi.PHASE_FLUX_ETA) == 0


@pytest.mark.parametrize("tp,tk", [(64, 16), (32, 8), (70, 14), (8, 1), (256, 4)])
def test_fast_flux_tiles_cover_every_cell_once(tp, tk):
    """A tile holds tk lines and tp cells of each.  The tp-5 middle positions
    hold a face; the tp-6 cells with both faces in the tile write their
    partial result, so tiles along a line step by tp-6 cells and consecutive
    tiles share one face."""
    d = make_desc(200, 200, 64, fast_math=1)
    n = [200, 200, 64]
    for q in range(3):
        other = [n[a] for a in range(3) if a != q]
        lines = other[0] * other[1]
        cpt = tp - 6
        segs = math.ceil(n[q] / cpt)
        assert capi.tile_count(d, capi.PHASE_FLUX_X + q, tp, tk) == math.ceil(lines / tk) * segs
        # tile s starts its faces at cell lo-1+s*cpt (position 2) and owns the
        # cells at positions 3 .. tp-4 of the line that lie in [lo, hi)
        owned = []
        for s in range(segs):
            c0 = s * cpt - 1
            owned += [c0 + ps - 2 for ps in range(3, tp - 3) if 0 <= c0 + ps - 2 < n[q]]
        assert owned == list(range(n[q]))
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
