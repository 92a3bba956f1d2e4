"""GPU parity of the analytic programs (radon_analytic, ray_analytic), FP64.

- Against the reference itself: the fixtures in tests/golden/ref_analytic.npz hold the output
  of the reference's radon_analytic/src/integral.c and ray_analytic/integral.c compiled with
  its example test functions; the GPU must match them within 1e-10 (relative L2 and max).
- Against the oracle (pinned bit for bit to those fixtures) on whole grids, on the phantom's
  ellipsoids, at ngrid = 129, and on empty and far-away test functions.
The device's exp/sin may differ from the C library's in the last bit, and the plane sums run
in another order, so agreement is to ~1e-15, not bit for bit."""
import math
import os

import numpy as np
import pytest

import orc
import wrt
from golden_inputs import EXAMPLES, analytic_zslices, example_terms

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _close(got, ref, tol=1e-10):
    got, ref = np.asarray(got).ravel(), np.asarray(ref).ravel()
    assert orc.rel_l2(got, ref) <= tol, orc.rel_l2(got, ref)
    assert orc.max_rel(got, ref) <= tol, orc.max_rel(got, ref)


@pytest.mark.parametrize("name", EXAMPLES)
def test_radon_analytic_matches_reference_fixtures(ctx, name):
    f = np.load(os.path.join(G, "ref_analytic.npz"), allow_pickle=False)
    terms = wrt.fn_example(name)
    for key in ("g17", "g33"):
        k = f"{name}_plane_{key}"
        g = int(f[k + "_g"])
        phi, theta, shift = f[k + "_in"]
        # one plane per (phi, theta, shift) triple: a 1 x 1 x 1 grid each would be slow, so
        # run the distinct values as a grid and pick the triples out
        up, ut, us = np.unique(phi), np.unique(theta), np.unique(shift)
        out, st = ctx.radon_analytic(terms, g, up, ut, us)
        got = out[np.searchsorted(us, shift), np.searchsorted(up, phi), np.searchsorted(ut, theta)]
        _close(got, f[k])


@pytest.mark.parametrize("name", EXAMPLES)
def test_ray_analytic_matches_reference_fixtures(ctx, name):
    f = np.load(os.path.join(G, "ref_analytic.npz"), allow_pickle=False)
    terms = wrt.fn_example(name)
    for key in ("g17", "g64r"):
        k = f"{name}_ray_{key}"
        g = int(f[k + "_g"])
        z, s, phi = f[k + "_in"]
        if key == "g17":
            uz, us, up = np.unique(z), np.unique(s), np.unique(phi)
            out, st = ctx.ray_analytic(terms, g, uz, us, up)
            got = out[np.searchsorted(uz, z), np.searchsorted(us, s), np.searchsorted(up, phi)]
            assert st["logical_samples"] == (4.0 * g + 1.0) * out.size
        else:
            got = np.array([ctx.ray_analytic(terms, g, [a], [b], [c])[0][0, 0, 0]
                            for a, b, c in zip(z[:16], s[:16], phi[:16])])
            f_ref = f[k][:16]
            _close(got, f_ref)
            continue
        _close(got, f[k])


@pytest.mark.parametrize("name", EXAMPLES)
def test_radon_analytic_grid_vs_oracle(ctx, name):
    terms = wrt.fn_example(name)
    phi, th, w, s = ctx.radon_grid_init(12, 6, 17)
    got, st = ctx.radon_analytic(terms, 25, phi, th, s)
    want = orc.analytic_radon_grid(terms, 25, phi, th, s)
    _close(got, want)
    assert st["logical_samples"] == 16.0 * 25 * 25 * got.size
    assert np.count_nonzero(want) > want.size // 3


def test_radon_analytic_ngrid_129_and_ellipsoids(ctx):
    terms = wrt.fn_example("exponential_bump")
    phi, th, w, s = ctx.radon_grid_init(4, 2, 5)
    got, _ = ctx.radon_analytic(terms, 129, phi, th, s)
    _close(got, orc.analytic_radon_grid(terms, 129, phi, th, s))
    E = wrt.fn_phantom(10, seed=2)
    phi, th, w, s = ctx.radon_grid_init(8, 4, 9)
    got, _ = ctx.radon_analytic(E, 33, phi, th, s)
    _close(got, orc.analytic_radon_grid(E, 33, phi, th, s))


@pytest.mark.parametrize("name", EXAMPLES)
def test_ray_analytic_grid_vs_oracle(ctx, name):
    terms = wrt.fn_example(name)
    s, phi = wrt.ray_grid_init(17, 16)
    z = wrt.analytic_zslice_grid(33)
    got, _ = ctx.ray_analytic(terms, 33, z, s, phi)
    want = orc.analytic_ray_grid(terms, 33, z, s, phi)
    _close(got, want)
    E = wrt.fn_phantom(10, seed=3)
    got, _ = ctx.ray_analytic(E, 33, z, s, phi)
    _close(got, orc.analytic_ray_grid(E, 33, z, s, phi))


def test_analytic_empty_and_far_functions(ctx):
    phi, th, w, s = ctx.radon_grid_init(4, 2, 5)
    out, _ = ctx.radon_analytic([], 17, phi, th, s)
    assert np.all(out == 0.0)
    far = [(0, (50.0, 0.0, 0.0), (0.5, 1e-2))]
    out, _ = ctx.radon_analytic(far, 17, phi, th, s)
    assert np.array_equal(out, orc.analytic_radon_grid(far, 17, phi, th, s))
    z = wrt.analytic_zslice_grid(9)
    s2, p2 = wrt.ray_grid_init(9, 8)
    out, _ = ctx.ray_analytic(far, 9, z, s2, p2)
    assert np.all(out == 0.0)


def test_analytic_invalid_terms(ctx):
    phi, th, w, s = ctx.radon_grid_init(4, 2, 5)
    bad = [(7, (0.0, 0.0, 0.0), (0.5, 1e-2))]
    skew = [(2, (0.0, 0.0, 0.0), [1.0, 0.3, 0.2, 0.1] + [1, 0.1, 0, 0, 1, 0, 0, 0, 1])]
    for terms, g in ((bad, 17), (skew, 17), (example_terms("exponential_bump"), 1)):
        with pytest.raises(wrt.WrtError) as e:
            ctx.radon_analytic(terms, g, phi, th, s)
        assert e.value.code == wrt.WRT_E_ARG
