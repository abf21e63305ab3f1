"""The kernels' terrain march (isect_terrain: skips the positions at which a crossing provably cannot lie, by a
first-order Lipschitz bound and a second-order bound on the height gap) returns exactly what the reference's
march over every position returns (GPURendererKernel.java:1470-1506), bit for bit, on millions of rays of every
kind: random, skimming the surface, climbing off it (sun and bounce rays from surface points), vertical, along
the axes and far away; in fast and exact math (df_check_terrain_march runs both marches on the GPU)."""

import numpy as np
import pytest

import dayflower_amd as dfa

pytestmark = pytest.mark.gpu


def _height(x, z):
    # object-space height of the reference's field: y = 10 sin(x / 10) sin(z / 10)
    return 10.0 * np.sin(x / 10.0) * np.sin(z / 10.0)


def _unit(v):
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    # random origins around the surface, random directions
    o = np.stack([rng.uniform(-80, 80, n), rng.uniform(-15, 15, n), rng.uniform(-80, 80, n)], 1)
    out.append(np.hstack([o, _unit(rng.normal(size=(n, 3)))]))
    # surface points (slightly above, on, below), directions skimming the tangent plane or leaving it
    x, z = rng.uniform(-80, 80, n), rng.uniform(-80, 80, n)
    y = _height(x, z) + rng.choice([1e-4, 1e-3, 1e-2, 0.0, -1e-3], n)
    hx = np.cos(x / 10) * np.sin(z / 10)
    hz = np.sin(x / 10) * np.cos(z / 10)
    nrm = _unit(np.stack([-hx, np.ones(n), -hz], 1))
    tang = _unit(np.cross(nrm, rng.normal(size=(n, 3))))
    lift = rng.choice([0.0, 1e-4, 1e-3, 1e-2, 0.05, 0.3, -1e-3], n)[:, None]
    d = _unit(tang + lift * nrm)
    out.append(np.hstack([np.stack([x, y, z], 1), d]))
    # the sun at 20 degrees and a low sun from surface points
    for elev in (20.0, 3.0):
        az = rng.uniform(0, 2 * np.pi)
        sd = np.array([np.cos(np.radians(elev)) * np.sin(az), np.sin(np.radians(elev)),
                       np.cos(np.radians(elev)) * np.cos(az)])
        out.append(np.hstack([np.stack([x, y + 1e-3, z], 1), np.tile(sd, (n, 1))]))
    # vertical, axis-aligned and degenerate directions, far origins, unnormalised directions
    k = n // 8
    axes = np.array([[0, 1, 0], [0, -1, 0], [1, 0, 0], [0, 0, 1], [-1, 0, 0], [0, 0, -1], [1, 0, 1e-7], [0, 1e-6, 1]],
                    np.float64)
    o = np.stack([rng.uniform(-80, 80, k * 8), rng.uniform(-12, 12, k * 8), rng.uniform(-80, 80, k * 8)], 1)
    out.append(np.hstack([o, np.repeat(axes, k, 0)]))
    o = np.stack([rng.uniform(-5e3, 5e3, k), rng.uniform(-30, 30, k), rng.uniform(-5e3, 5e3, k)], 1)
    out.append(np.hstack([o, rng.normal(size=(k, 3)) * rng.uniform(0.2, 5.0, (k, 1))]))
    return np.vstack(out).astype(np.float32)


@pytest.mark.parametrize("math,n", [(dfa.MATH_FAST, 1 << 20), (dfa.MATH_EXACT, 1 << 16)])
def test_terrain_march_skips_no_crossing(gpu, math, n):
    rays = _rays(n, 11 + math)
    t_skip, t_plain, evals = dfa.check_terrain_march(rays, math)
    bad = np.flatnonzero(t_skip.view(np.uint32) != t_plain.view(np.uint32))
    assert bad.size == 0, (bad.size, rays[bad[:5]].tolist(), t_skip[bad[:5]].tolist(), t_plain[bad[:5]].tolist())
    hits = np.count_nonzero(t_plain < 1e4)
    assert hits > rays.shape[0] // 10  # the families above do cross the field
    # far fewer heights than the reference's up to 926 per ray
    assert evals.mean() < 60, evals.mean()
