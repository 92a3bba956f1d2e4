"""GPU: the reduction on independent ray and plane grids (wrt_radon_reduction_rays), through
the C-ABI.  On identical full-turn grids it is wrt_radon_reduction bit for bit; on the
throughput ray shape (128 slices x 180 angles in [0, pi) x 257 shifts -> 64 azimuths in
[0, 2 pi) x 32 theta x 129 offsets) it matches the oracle to 1e-12 of the largest value
(the GPU forms each azimuth's rows before the spline, the oracle splines each ray angle
and combines the values: the same number up to rounding); an off-centre Gaussian meets its
closed-form plane integrals."""
import math

import numpy as np
import pytest

import orc
import wrt
from golden_inputs import gauss_planes, gauss_rays

pytestmark = pytest.mark.gpu


def test_identical_grids_equal_reduction_bitwise(ctx):
    nz, ns, nphi, nth = 33, 33, 16, 8
    vol = wrt.phantom(nz, False, 10, 2, 0.0)
    z = wrt.zslice_grid(nz)
    s, phi = wrt.ray_grid_init(ns, nphi)
    rays, _ = ctx.ray_grid(vol, z, s, phi)
    th, _ = orc.theta_grid(nth)
    got, st = ctx.radon_reduction_rays(rays, s, False, s, phi, th)
    want, _ = ctx.radon_reduction(rays, s, phi, th)
    assert np.array_equal(got, want)
    assert st["outputs"] == ns * nphi * nth


def test_throughput_ray_shape_vs_oracle(ctx):
    n = 128
    z = wrt.zslice_grid(128)
    rs, phi2 = wrt.ray_grid_init(257, 360)
    rphi = np.ascontiguousarray(phi2[:180])
    vol = wrt.phantom(n, False, 10, 3, 0.0)
    rays, _ = ctx.ray_grid(vol, z, rs, rphi)
    s, phi = wrt.ray_grid_init(129, 64)
    _, th, _, _ = ctx.radon_grid_init(64, 32, 129)
    got, st = ctx.radon_reduction_rays(rays, rs, True, s, phi, th)
    assert np.all(np.isfinite(got))
    scale = np.max(np.abs(got))
    rng = np.random.default_rng(6)
    for _ in range(16):
        i, j, k = rng.integers(len(s)), rng.integers(len(phi)), rng.integers(len(th))
        want = orc.radon_reduction_rays_one(rays, rs, True, s[i], phi[j], th[k])
        assert abs(got[i, j, k] - want) <= 1e-12 * scale, (i, j, k)


def test_gaussian_closed_form(ctx):
    z = wrt.zslice_grid(128)
    rs, _ = wrt.ray_grid_init(257, 4)
    ang = (math.pi / 180) * np.arange(180)
    rays = np.ascontiguousarray(gauss_rays(z, rs, ang))
    p = np.linspace(-1, 1, 65)
    rng = np.random.default_rng(12)
    phis = rng.uniform(-math.pi, 3 * math.pi, 20)
    th, _ = orc.theta_grid(8)
    got, _ = ctx.radon_reduction_rays(rays, rs, True, p, phis, th)
    want = gauss_planes(p, phis, th)
    assert np.max(np.abs(got - want)) / np.max(want) < 2e-4


def test_invalid_arguments(ctx):
    rs = np.linspace(-1, 1, 9)
    rays = np.zeros((8, 9, 6))
    th = np.array([0.5])
    ctx.radon_reduction_rays(rays, rs, True, rs, [0.1, 4.0], th)  # valid
    skew = rs.copy()
    skew[0] = -0.99
    for args in ((rays, skew, True), (rays[:, :, :3].copy(), rs, True),
                 (rays[:2].copy(), rs, False)):
        with pytest.raises(wrt.WrtError) as e:
            ctx.radon_reduction_rays(*args, rs, [0.1], th)
        assert e.value.code == wrt.WRT_E_ARG
    with pytest.raises(wrt.WrtError) as e:
        ctx.radon_reduction_rays(rays, rs[::-1].copy(), False, rs, [0.1], th)
    assert e.value.code == wrt.WRT_E_ARG
