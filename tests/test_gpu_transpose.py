"""GPU tests of the FP32 exact transposes (wrt_radon_grid_T / wrt_ray_grid_T with WRT_F32).

Checks, each against the plain FP64 transposes (themselves pinned against the oracle and by
the FP64 dot-product test in test_gpu_parity.py):
- Gaussian plane/ray values: relative L2 <= 1e-5 and max |error| <= 1e-5 * max|ref|; the
  plane form sums in 64-bit fixed point (quantum below 1e-12 of max|g|) and rounds once to
  FP32, so it is held to 1e-6;
- one-hot values on structured planes (axis-aligned and diagonal directions, shifts on cell
  faces): every voxel then holds delta^2 * (number of the plane's samples in its cell), so a
  single flipped index decision is an error of at least a quarter of the largest value;
- the dot-product test of the FP32 pairs, <A f, g> = <f, A^T g>, to 1e-6 relative to
  ||A f|| ||g||;
- n not a multiple of the 32^3 brick, n = 200 (the face half bricks then run on a second
  stream, concurrently with the interior ones), and the ray form at and above the LDS plane
  limit (n = 143, FP64 plane);
- the plane form gives the same bits on every launch (integer sums), and stays exact on a
  shift grid whose shifts crowd into one cell's width (the fixed point's bound).
"""
import numpy as np
import pytest

import orc
import wrt

pytestmark = pytest.mark.gpu


def _check(got, ref, tol=1e-5):
    assert orc.rel_l2(got, ref) <= tol, orc.rel_l2(got, ref)
    assert np.max(np.abs(got - ref)) <= tol * np.max(np.abs(ref)), \
        np.max(np.abs(got - ref)) / np.max(np.abs(ref))


@pytest.mark.parametrize("n,nphi,nth", [(16, 12, 6), (40, 16, 8), (64, 32, 16), (96, 24, 12),
                                         (200, 8, 4)])
def test_radon_T_f32_matches_plain(ctx, n, nphi, nth):
    phi, th, w, s = ctx.radon_grid_init(nphi, nth, n + 1)
    g = wrt.gaussian((len(s), nphi, nth), seed=n)
    ref, _ = ctx.radon_grid_T(g, n, phi, th, s, prec="f64")
    got, st = ctx.radon_grid_T(g, n, phi, th, s, prec="f32")
    _check(got, ref, 1e-6)
    assert got.shape == (n, n, n)
    assert st["logical_samples"] == 16.0 * n * n * g.size


@pytest.mark.parametrize("n", [32, 40, 64, 200])
def test_radon_T_f32_one_hot_structured(ctx, n):
    phi = np.array([0.0, np.pi / 4, np.pi / 2, np.pi, 1.5 * np.pi, 0.3])
    _, th_gl, _, _ = ctx.radon_grid_init(4, 6, 3)
    th = np.concatenate([th_gl, [np.pi / 2, np.pi / 4, 0.0]])
    s = np.unique(np.concatenate([-1.0 + np.arange(0, n + 1, 3) * (2.0 / n), [-1.0, 1.0, 0.0]]))
    rng = np.random.default_rng(n)
    for trial in range(3):
        g = np.zeros((len(s), len(phi), len(th)))
        # a few planes of distinct, exactly representable weights
        for k, (i, j, m) in enumerate(zip(rng.integers(len(s), size=4), rng.integers(len(phi), size=4),
                                          rng.integers(len(th), size=4))):
            g[i, j, m] += 2.0 ** k
        ref, _ = ctx.radon_grid_T(g, n, phi, th, s, prec="f64")
        got, _ = ctx.radon_grid_T(g, n, phi, th, s, prec="f32")
        if np.max(np.abs(ref)) == 0:
            assert np.max(np.abs(got)) == 0
            continue
        _check(got, ref, 1e-6)


@pytest.mark.parametrize("n", [33, 64])
def test_radon_f32_dot_product(ctx, n):
    phi, th, w, s = ctx.radon_grid_init(32, 16, n + 1)
    f = wrt.gaussian((n, n, n), seed=1)
    g = wrt.gaussian((len(s), len(phi), len(th)), seed=2)
    Af, _ = ctx.radon_grid(f, phi, th, s, prec="f32")
    ATg, _ = ctx.radon_grid_T(g, n, phi, th, s, prec="f32")
    lhs, rhs = float(np.sum(Af * g)), float(np.sum(f * ATg))
    assert abs(lhs - rhs) <= 1e-6 * np.linalg.norm(Af) * np.linalg.norm(g), (lhs, rhs)


@pytest.mark.parametrize("n", [17, 33, 48, 143, 150, 190])
def test_ray_T_f32_matches_plain(ctx, n):
    # n - 1 a power of two (exact scaling) and not; 143: the largest LDS plane (FP64);
    # 150, 190: none
    s, phi = wrt.ray_grid_init(n + 1 if n < 100 else 65, 16)
    z = np.concatenate([wrt.zslice_grid(12), [1.5, -1.0000001]])
    g = wrt.gaussian((len(z), len(s), len(phi)), seed=n)
    ref, _ = ctx.ray_grid_T(g, n, z, s, phi, prec="f64")
    got, st = ctx.ray_grid_T(g, n, z, s, phi, prec="f32")
    _check(got, ref)
    assert st["logical_samples"] == (4.0 * n + 1) * g.size


@pytest.mark.parametrize("n", [17, 40])
def test_ray_T_f32_one_hot(ctx, n):
    s, phi = wrt.ray_grid_init(n + 1, 8)  # phi multiples of pi/4: the index quirk's rays
    z = wrt.zslice_grid(5)
    rng = np.random.default_rng(n)
    g = np.zeros((len(z), len(s), len(phi)))
    for k in range(4):
        g[rng.integers(len(z)), rng.integers(len(s)), rng.integers(len(phi))] += 2.0 ** k
    ref, _ = ctx.ray_grid_T(g, n, z, s, phi, prec="f64")
    got, _ = ctx.ray_grid_T(g, n, z, s, phi, prec="f32")
    _check(got, ref, 1e-6)


@pytest.mark.parametrize("n", [33, 64])
def test_ray_f32_dot_product(ctx, n):
    s, phi = wrt.ray_grid_init(n + 1, 24)
    z = wrt.zslice_grid(n // 2)
    f = wrt.gaussian((n, n, n), seed=3)
    g = wrt.gaussian((len(z), len(s), len(phi)), seed=4)
    Af, _ = ctx.ray_grid(f, z, s, phi, prec="f32")
    ATg, _ = ctx.ray_grid_T(g, n, z, s, phi, prec="f32")
    lhs, rhs = float(np.sum(Af * g)), float(np.sum(f * ATg))
    assert abs(lhs - rhs) <= 1e-6 * np.linalg.norm(Af) * np.linalg.norm(g), (lhs, rhs)


def test_transpose_f32_jobs_repeat(ctx):
    # a resident job repeats its result on every launch: the same bits (64-bit integer sums
    # in any order)
    n = 48
    phi, th, w, s = ctx.radon_grid_init(16, 8, n + 1)
    j = ctx.job("radon_T", "f32", n=n, phi=phi, theta=th, shift=s)
    j.load(wrt.gaussian((len(s), len(phi), len(th)), seed=7))
    j.run()
    a = j.fetch()
    j.run()
    b = j.fetch()
    j.close()
    assert np.array_equal(a, b)


def test_radon_T_f32_crowded_shifts(ctx):
    # 41 shifts within 0.02 (many planes of one direction through each cell) and a wide range
    # of magnitudes: the fixed point's scale follows both
    n = 40
    s = np.unique(np.concatenate([np.linspace(-0.01, 0.01, 41), np.linspace(-1, 1, 21)]))
    phi, th, _, _ = ctx.radon_grid_init(12, 6, 3)
    g = wrt.gaussian((len(s), len(phi), len(th)), seed=5) * np.logspace(-3, 3, len(s))[:, None, None]
    ref, _ = ctx.radon_grid_T(g, n, phi, th, s, prec="f64")
    got, _ = ctx.radon_grid_T(g, n, phi, th, s, prec="f32")
    _check(got, ref, 1e-6)
