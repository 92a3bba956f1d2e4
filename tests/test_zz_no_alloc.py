"""GPU test of wrt_radon_grid_noalloc, the entry point that allocates nothing: through the
C-ABI, two calls in a row (the second on a smaller volume and grid) against
wrt_radon_grid's FP32 form on the same inputs, and the limit of its fixed scratch.

Equal bit for bit to the FP32 form walked in the same 8 x 8 bands (both add the same FP32
section sums into exact 64-bit fixed-point plane sums and round those to FP32), and within
relative L2 1e-5 of the plain FP64 form."""
import ctypes as C

import numpy as np
import pytest

import orc
import wrt

pytestmark = pytest.mark.gpu


def _case(ctx, n, nphi, nth, ns, seed):
    vol = wrt.phantom(n, True, 10, seed, 0.01)
    phi, th, _, s = ctx.radon_grid_init(nphi, nth, ns)
    return vol, phi, th, s


def test_noalloc_matches_fp32_form_twice(ctx, monkeypatch):
    # 160^3 on 80 x 40 x 161, then 48^3 on a shuffled 25-shift grid: the second call must
    # not see the first call's bricks, counts or sums.  The per-job form walks the 8 x 8
    # bands the entry point walks (its autotuner may pick another band shape, which sums
    # each section's lanes in another FP32 order).
    monkeypatch.setenv("WRT_BAND_ROWS", "8")
    rng = np.random.default_rng(1)
    cases = [_case(ctx, 160, 80, 40, 161, 3), _case(ctx, 48, 24, 12, 25, 4)]
    vol, phi, th, s = cases[1]
    cases[1] = (vol, phi, th, s[rng.permutation(len(s))])
    for vol, phi, th, s in cases:
        got, st = ctx.radon_grid_noalloc(vol, phi, th, s)
        want, _ = ctx.radon_grid(vol, phi, th, s, prec="f32")
        assert np.array_equal(got, want), orc.max_rel(got, want)
        ref, _ = ctx.radon_grid(vol, phi, th, s, prec="f64")
        assert orc.rel_l2(got, ref) <= 1e-5
        assert st["kernel_seconds"] > 0
        assert st["logical_samples"] == 16.0 * vol.shape[0] ** 2 * got.size


def test_noalloc_rejects_beyond_scratch(ctx, monkeypatch):
    # 256^3 FP32 bricks alone fill the 64 MiB scratch
    n = 256
    vol = np.zeros((n, n, n))
    phi, th, _, s = ctx.radon_grid_init(4, 2, 5)
    out = np.full((5, 4, 2), 7.0)
    rc = wrt.lib().wrt_radon_grid_noalloc(ctx.h, vol.ctypes.data, n, phi.ctypes.data, 4,
                                          th.ctypes.data, 2, s.ctypes.data, 5,
                                          out.ctypes.data, None)
    assert rc == wrt.WRT_E_SCRATCH
    assert np.all(out == 7.0)  # nothing written
    # and the entry point still works after the refusal
    monkeypatch.setenv("WRT_BAND_ROWS", "8")
    vol, phi, th, s = _case(ctx, 32, 8, 4, 17, 5)
    got, _ = ctx.radon_grid_noalloc(vol, phi, th, s)
    want, _ = ctx.radon_grid(vol, phi, th, s, prec="f32")
    assert np.array_equal(got, want)
