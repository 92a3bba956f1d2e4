"""GPU tests of the plane integrals on shift grids in any order.

plane_integral (radon_grid/src/integral.c:26-59) takes any shift, so a caller's grid may be
descending, shuffled or hold repeated values.  The fast brick kernel bisects an ascending
copy of the grid and writes each plane to the caller's row; these tests check every fast
form (FP32 and BF16 forward, the exact FP32 transpose, the persistent batch) against the
plain FP64 form on the same grid, and against the ascending grid with the rows permuted.

Tolerances: FP32 relative L2 and max relative error <= 1e-5 against FP64; BF16 relative L2
<= 2e-3; the same form on a permuted grid: relative L2 <= 1e-7 (the FP64 accumulators'
atomic order) and bitwise for the transpose (exact integer sums)."""
import numpy as np
import pytest

import orc
import wrt

pytestmark = pytest.mark.gpu


def _grids(ctx, n, seed):
    phi, th, w, s = ctx.radon_grid_init(16, 8, n + 1)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(s))
    shuffled = s[perm]
    repeated = np.r_[s[::-1], s[5:12]]  # descending, then seven repeats
    return phi, th, s, {"descending": (s[::-1], np.arange(len(s))[::-1]),
                        "shuffled": (shuffled, perm),
                        "repeated": (repeated, np.r_[np.arange(len(s))[::-1], np.arange(5, 12)])}


@pytest.mark.parametrize("n", [64, 45])
def test_fast_forward_any_shift_order(ctx, n):
    vol = wrt.phantom(n, True, 10, n, 0.01)
    phi, th, s, cases = _grids(ctx, n, n)
    asc32, _ = ctx.radon_grid(vol, phi, th, s, prec="f32")
    for name, (grid, rows) in cases.items():
        ref, _ = ctx.radon_grid(vol, phi, th, grid, prec="f64")
        for prec, tol in (("f32", 1e-5), ("bf16", 2e-3)):
            got, _ = ctx.radon_grid(vol, phi, th, grid, prec=prec)
            err = orc.rel_l2(got, ref)
            assert err <= tol, (name, prec, err)
            if prec == "f32":
                assert orc.max_rel(got, ref) <= 1e-5, (name, orc.max_rel(got, ref))
                assert orc.rel_l2(got, asc32[rows]) <= 1e-7, name


def test_fast_transpose_any_shift_order(ctx):
    n = 48
    phi, th, s, cases = _grids(ctx, n, 7)
    g_asc = wrt.gaussian((len(s), len(phi), len(th)), 21)
    asc, _ = ctx.radon_grid_T(g_asc, n, phi, th, s, prec="f32")
    for name, (grid, rows) in cases.items():
        if name == "repeated":
            continue  # rows repeat: no permutation of the ascending case
        g = g_asc[rows]  # row r of the permuted grid holds shift s[rows[r]]
        got, _ = ctx.radon_grid_T(g, n, phi, th, grid, prec="f32")
        assert np.array_equal(got, asc), name
        ref, _ = ctx.radon_grid_T(g, n, phi, th, grid, prec="f64")
        assert orc.rel_l2(got, ref) <= 1e-5, (name, orc.rel_l2(got, ref))
    grid, _ = cases["repeated"]
    g = wrt.gaussian((len(grid), len(phi), len(th)), 22)
    got, _ = ctx.radon_grid_T(g, n, phi, th, grid, prec="f32")
    ref, _ = ctx.radon_grid_T(g, n, phi, th, grid, prec="f64")
    assert orc.rel_l2(got, ref) <= 1e-5
    # the pair is still an exact transpose: <A f, g> = <f, A^T g>
    f = wrt.gaussian((n, n, n), 23)
    Af, _ = ctx.radon_grid(f, phi, th, grid, prec="f64")
    lhs, rhs = float(np.dot(Af.ravel(), g.ravel())), float(np.dot(f.ravel(), ref.ravel()))
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(Af) * np.linalg.norm(g)


def test_batch_any_shift_order(ctx):
    n = 40
    phi, th, s, cases = _grids(ctx, n, 3)
    jobs, want = [], []
    for k, (grid, _) in enumerate(cases.values()):
        j = ctx.job("radon", "f32", n=n, phi=phi, theta=th, shift=grid)
        j.load(wrt.phantom(n, True, 10, 50 + k, 0.01))
        j.run()
        want.append(j.fetch())
        jobs.append(j)
    ctx.batch_run(jobs)
    for j, w in zip(jobs, want):
        assert orc.rel_l2(j.fetch(), w) <= 1e-7
        j.close()
