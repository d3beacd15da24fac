"""The push-order sort key, restated in the test helpers (CPU only).  The GPU
sort tests compare the device's permutation with a stable argsort of this
key, so these checks pin the restatement: a bijection of the cells with 8 x 8
blocks contiguous, supertiles along a curve of edge steps that visits every
supertile once, and, on a power-of-two square, the classic Hilbert curve."""
import numpy as np
import pytest

from cases import gilbert_order, supertile_key


def _hilbert_d2xy(n, d):
    """Classic Hilbert curve on an n x n grid (n a power of two): the point at
    distance d, by the quadrant-rotation recurrence (Butz; the form found in
    most references), independent of the rectangle-splitting construction."""
    x = y = 0
    s, t = 1, d
    while s < n:
        rx = 1 & (t // 2)
        ry = 1 & (t ^ rx)
        if ry == 0:
            if rx == 1:
                x, y = s - 1 - x, s - 1 - y
            x, y = y, x
        x += s * rx
        y += s * ry
        t //= 4
        s *= 2
    return x, y


@pytest.mark.parametrize("n", [2, 4, 8, 16, 32])
def test_power_of_two_square_is_the_hilbert_curve(n):
    assert gilbert_order(n, n) == [_hilbert_d2xy(n, d) for d in range(n * n)]


@pytest.mark.parametrize("w,h", [(1, 1), (1, 7), (9, 1), (5, 5), (128, 65), (65, 128), (25, 3),
                                 (3, 25), (38, 33), (256, 256), (6, 10), (10, 6)])
def test_curve_visits_every_cell_by_edge_steps(w, h):
    pts = gilbert_order(w, h)
    assert sorted(pts) == [(x, y) for x in range(w) for y in range(h)]
    steps = np.abs(np.diff(np.array(pts), axis=0)).sum(axis=1)
    diag = int(np.sum(steps == 2)) if len(pts) > 1 else 0
    assert np.all((steps == 1) | (steps == 2)) and diag <= 1
    if (w, h) in ((256, 256), (128, 65), (6, 10), (10, 6)):
        assert diag == 0


def test_key_is_a_bijection_with_contiguous_blocks():
    for nx, ny in ((64, 64), (40, 37), (1024, 514), (8, 300), (200, 20)):
        k = supertile_key(np.arange(nx * ny), nx, ny)
        assert len(np.unique(k)) == nx * ny
        assert k.max() < ((nx + 7) // 8) * ((ny + 7) // 8) * 64
    nx = ny = 64
    c = np.arange(nx * ny)
    k = supertile_key(c, nx, ny)
    blk = ((c // ny) >> 3) * 8 + ((c % ny) >> 3)
    runs = blk[np.argsort(k)].reshape(-1, 64)
    assert np.all(runs == runs[:, :1])
