"""Reference-patch grid: closed form for one thread and the counts of the
threaded tiling (ind_initialize, subDivideTight, src/vbm3d.cpp:471-489)."""
import numpy as np
import pytest

import fp64_model as M

# refs per frame: nb = 1 single-thread grid; the tiled counts are those of the
# reference's crop grid with a 2*16 px border.
COUNTS = [
    ((176, 144, 8, 6, 1), 696), ((176, 144, 7, 3, 1), 2726), ((960, 540, 8, 6, 1), 14400),
    ((960, 540, 7, 3, 1), 57101), ((1920, 1080, 8, 6, 1), 57600), ((1920, 1080, 7, 3, 1), 229401),
    ((1920, 1080, 8, 4, 1), 128851), ((640, 360, 8, 6, 1), 6420), ((640, 360, 7, 3, 1), 25228),
    ((176, 144, 8, 6, 2), 816), ((176, 144, 8, 6, 8), 1380), ((176, 144, 8, 6, 32), 3240),
    ((1920, 1080, 8, 6, 2), 58680), ((1920, 1080, 8, 6, 8), 62496), ((1920, 1080, 8, 6, 32), 73656),
]


@pytest.mark.parametrize("args,count", COUNTS)
def test_grid_counts(oracle, args, count):
    W, H, k, p, nb = args
    xs, ys, cr = oracle.grid(W, H, k, p, 16, nb)
    assert len(xs) == count


@pytest.mark.parametrize("W,H,k,p", [(176, 144, 8, 6), (33, 21, 7, 3), (64, 8, 8, 5)])
def test_single_thread_grid_closed_form(oracle, W, H, k, p):
    xs, ys, _ = oracle.grid(W, H, k, p, 16, 1)
    rows, cols = M.axis_grid(H - k, p), M.axis_grid(W - k, p)
    np.testing.assert_array_equal(ys, np.repeat(rows, len(cols)))
    np.testing.assert_array_equal(xs, np.tile(cols, len(rows)))


@pytest.mark.parametrize("nb", [2, 4, 8, 16, 32])
def test_tiled_grid_structure(oracle, nb):
    W, H, k, p = 960, 540, 8, 6
    xs, ys, cr = oracle.grid(W, H, k, p, 16, nb)
    assert cr.max() == nb - 1 and np.all(np.diff(cr.astype(int)) >= 0)
    assert xs.max() == W - k and ys.max() == H - k and xs.min() == 0 and ys.min() == 0
    # every pixel is covered by some reference patch
    cov = np.zeros((H, W), bool)
    for x, y in zip(xs, ys):
        cov[y:y + k, x:x + k] = True
    assert cov.all()


def test_invalid_grids(oracle):
    with pytest.raises(ValueError):
        oracle.grid(7, 20, 8, 6, 16, 1)
    with pytest.raises(ValueError):
        oracle.grid(176, 144, 8, 6, 16, 3)
