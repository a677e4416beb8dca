"""The irradiance's blend grid (CPU, no GPU): at a fixed LOD the trilinear
value (1 - delta) B0 + delta B1 of two pyramid levels, B1 the exact 2x2
reduction of B0's level, is bilinear between nodes at fine coordinates
2k, 2k + 1/2 and 2k + 1 along each axis (where the fine level's and the
coarse level's bilinear pieces meet), so a bilinear lookup of the node
values (REPEAT border, as the kernel stores them) reproduces it.  The cell
selection is the device's: r = x0 - 2 floor(x0 / 2), lower node
3 floor(x0 / 2) + [r >= 1/2] + [r >= 1], fraction 2r, 2r - 1 or r - 1.

A mistake in the node positions, the cell boundaries or the fractions
(e.g. nodes evenly spaced at 2/3, or the coarse level's centres at 2k)
breaks the agreement by far more than the tolerance."""
import numpy as np
import pytest


def _bilinear(level, x, y):
    h, w = level.shape
    fx, fy = np.floor(x), np.floor(y)
    ax, ay = x - fx, y - fy
    i, j = fx.astype(int), fy.astype(int)
    at = lambda a, b: level[b % h, a % w]
    return (1 - ay) * ((1 - ax) * at(i, j) + ax * at(i + 1, j)) + ay * ((1 - ax) * at(i, j + 1) + ax * at(i + 1, j + 1))


def _cell(x0):
    k = np.floor(0.5 * x0)
    r = x0 - 2.0 * k
    c1, c2 = r >= 0.5, r >= 1.0
    t = np.where(c2, r - 1.0, np.where(c1, 2.0 * r - 1.0, 2.0 * r))
    return (3 * k + c1 + c2).astype(int), t


@pytest.mark.parametrize("w0,h0,delta", [(16, 8, 0.77), (8, 8, 0.25), (2, 2, 0.5), (32, 4, 0.001)])
def test_blend_grid_reproduces_the_two_level_lookup(w0, h0, delta):
    rng = np.random.default_rng(w0 * 100 + h0)
    fine = rng.uniform(0.0, 10.0, (h0, w0))
    coarse = 0.25 * (fine[0::2, 0::2] + fine[1::2, 0::2] + fine[0::2, 1::2] + fine[1::2, 1::2])
    blend = lambda x, y: (1 - delta) * _bilinear(fine, x, y) + delta * _bilinear(coarse, 0.5 * x - 0.25, 0.5 * y - 0.25)
    nx, ny = 3 * w0 // 2, 3 * h0 // 2
    pos = lambda n: 2.0 * (n // 3) + 0.5 * (n % 3)
    c, r = np.meshgrid(np.arange(nx + 2), np.arange(ny + 2))
    grid = blend(pos((c - 1) % nx), pos((r - 1) % ny))  # padded, REPEAT border
    x0 = rng.uniform(-0.5, w0 - 0.5, 20000)
    y0 = rng.uniform(-0.5, h0 - 0.5, 20000)
    x0[:4] = [-0.5, 0.0, 0.5, w0 - 0.5]
    n, tx = _cell(x0)
    m, ty = _cell(y0)
    assert n.min() >= -1 and n.max() + 1 <= nx  # inside the padded grid
    g = lambda a, b: grid[b + 1, a + 1]
    got = (1 - ty) * ((1 - tx) * g(n, m) + tx * g(n + 1, m)) + ty * ((1 - tx) * g(n, m + 1) + tx * g(n + 1, m + 1))
    np.testing.assert_allclose(got, blend(x0, y0), rtol=0, atol=1e-13)
    # Evenly spaced nodes (3 per 2 texels at 2/3) are not exact: the check has teeth.
    ex = x0 * 1.5
    even = _bilinear(grid[1:-1, 1:-1], ex, np.clip(y0 * 1.5, 0, ny - 1))
    assert np.abs(even - blend(x0, y0)).max() > 1e-3
