"""CPU checks of the workload helpers in sgnx.cases that the gradient checks
lean on: seeded finite-difference directions scale with the grid's lengths,
and the margin's gauge stations lie inside its uplift (no GPU needed)."""
import dataclasses

import numpy as np

from sgnx import cases
from sgnx.solver import node_coords


def test_smooth_direction_scales_with_the_grid():
    """The same seeded direction on a grid k times larger, with scale k, is
    the same field (the Gaussians' centres, offsets and widths all scale):
    a direction built for metres serves a grid of kilometres."""
    c = cases.inversion(seed=1, Nx=12, Ny=6)
    gauges = c.meta["gauges"]
    for k in (10.0, 1000.0):
        big = dataclasses.replace(c, Lx=k * c.Lx, Ly=k * c.Ly)
        a = cases.smooth_direction(c, np.random.default_rng(5), near=gauges)
        b = cases.smooth_direction(big, np.random.default_rng(5), near=gauges, scale=k)
        assert np.abs(a).max() > 1e-3
        np.testing.assert_allclose(b, a, rtol=0, atol=1e-13)
    # scale 1 is the default
    a = cases.smooth_direction(c, np.random.default_rng(7), near=gauges)
    b = cases.smooth_direction(c, np.random.default_rng(7), near=gauges, scale=1.0)
    assert np.array_equal(a, b)


def test_margin_gauges_lie_inside_the_uplift():
    c = cases.margin(seed=2, Nx=40, Ny=20)
    mp = c.meta["margin"]
    g = cases.margin_gauges(c, 12, seed=4)
    assert g.dtype == np.int64 and g.size == 12
    assert np.all(np.diff(g) > 0) and g[0] >= 0 and g[-1] < c.nd
    x, y = node_coords(c.P, c.Nx, c.Ny, c.Lx, c.Ly)
    r2 = ((x[g] - mp["xs"] * c.Lx) / mp["sx"]) ** 2 + ((y[g] - mp["ys"] * c.Ly) / mp["sy"]) ** 2
    assert np.all(r2 < 1.0)
    # the surface is raised there: eta0 = amp exp(-r2) > amp / e
    eta0 = c.U0[:c.nd] - c.depth["Hnod"]
    assert np.all(eta0[g] > mp["amp"] * np.exp(-1.0) * (1 - 1e-12))
    assert np.array_equal(cases.margin_gauges(c, 12, seed=4), g)
    assert not np.array_equal(cases.margin_gauges(c, 12, seed=5), g)
