"""GPU tests of the fast plane-integral form (FP32 values / BF16 bricks, FP32 Kahan sums,
exact FP64 index decisions) against the plain FP64 form and the oracle.

Tolerances: FP32 relative L2 <= 1e-5 and max |error| <= 1e-5 * max|R| (the target of the
form); BF16 bricks relative L2 <= 2e-3 (BF16 keeps 8 significand bits per voxel)."""
import numpy as np
import pytest

import orc
import wrt

pytestmark = pytest.mark.gpu


def _phantom_case(ctx, n, nphi, nth, ns, seed=0, sigma=0.0):
    vol = wrt.phantom(n, True, 10, seed, sigma)
    phi, th, w, s = ctx.radon_grid_init(nphi, nth, ns)
    return vol, phi, th, s


@pytest.mark.parametrize("n", [64, 48, 33, 96])
def test_fast_f32_matches_plain(ctx, n):
    vol, phi, th, s = _phantom_case(ctx, n, 32, 16, n + 1, seed=n, sigma=0.01)
    ref, _ = ctx.radon_grid(vol, phi, th, s, prec="f64")
    got, st = ctx.radon_grid(vol, phi, th, s, prec="f32")
    assert orc.rel_l2(got, ref) <= 1e-5, orc.rel_l2(got, ref)
    assert orc.max_rel(got, ref) <= 1e-5, orc.max_rel(got, ref)
    assert st["logical_samples"] == 16.0 * n * n * ref.size


def test_fast_f32_c0_full_grid(ctx):
    vol, phi, th, s = _phantom_case(ctx, 64, 64, 32, 129)
    ref, _ = ctx.radon_grid(vol, phi, th, s, prec="f64")
    got, _ = ctx.radon_grid(vol, phi, th, s, prec="f32")
    assert orc.rel_l2(got, ref) <= 1e-5
    # structured directions (phi multiple of pi/2) where index decisions hinge on rounding
    for ip in (0, 16, 32, 48):
        assert orc.rel_l2(got[:, ip, :], ref[:, ip, :]) <= 1e-5
    rng = np.random.default_rng(3)
    idx = rng.choice(ref.size, 64, replace=False)
    iS, iP, iT = np.unravel_index(idx, ref.shape)
    want = orc.radon_planes(vol, phi[iP], th[iT], s[iS])
    assert orc.rel_l2(got.ravel()[idx], want) <= 1e-5


def test_fast_f32_random_volume_and_odd_grids(ctx):
    # a volume with no zero bricks and values of both signs everywhere in the cube
    rng = np.random.default_rng(5)
    n = 40
    vol = rng.standard_normal((n, n, n))
    phi = rng.uniform(0, 2 * np.pi, 7)
    th = np.arccos(rng.uniform(-1, 1, 5))
    s = np.sort(rng.uniform(-1, 1, 11))
    ref, _ = ctx.radon_grid(vol, phi, th, s, prec="f64")
    got, _ = ctx.radon_grid(vol, phi, th, s, prec="f32")
    assert orc.rel_l2(got, ref) <= 1e-5


def test_fast_bf16_bricks(ctx):
    vol, phi, th, s = _phantom_case(ctx, 64, 32, 16, 65, seed=2, sigma=0.01)
    ref, _ = ctx.radon_grid(vol, phi, th, s, prec="f64")
    got, _ = ctx.radon_grid(vol, phi, th, s, prec="bf16")
    err = orc.rel_l2(got, ref)
    assert err <= 2e-3, err


def test_fast_job_repeatable(ctx):
    vol, phi, th, s = _phantom_case(ctx, 64, 16, 8, 65, seed=4)
    j = ctx.job("radon", "f32", n=64, phi=phi, theta=th, shift=s)
    j.load(vol)
    j.run()
    a = j.fetch()
    j.run()
    b = j.fetch()
    j.close()
    assert orc.rel_l2(a, b) <= 1e-12
