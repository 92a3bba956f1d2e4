"""GPU parity of the Fourier-slice (Chang-type) inversions (wrt_fourier_inversion and
wrt_fourier_inversion2d, FP64, radon_inversion/chang3d.m:3-30, chang2d.m:3-26) against the oracle's restatement (numpy rtft3d + the exact
adjoint sum in C, pinned in test_oracle.py by a brute-force DFT and a closed-form Gaussian):
relative L2 <= 1e-11 and max error <= 1e-10 of the output's scale.  The product evaluates the
adjoint sum by Horner's scheme, the oracle term by term, so the two differ by rounding only."""
import math

import numpy as np
import pytest

import orc
import wrt
from golden_inputs import (example_terms, fourier_case, fourier_case2d, gauss2d_on_grid,
                           gauss_on_grid)

pytestmark = pytest.mark.gpu

# (nphi, ntheta, nshift, ngrid, rsupp, padding): ntotal = padding * floor((nshift - 1) / (2 rsupp))
CASES = [
    (8, 4, 17, 8, 1.0, 4.0),     # ntotal 32
    (6, 5, 33, 12, 1.0, 2.0),    # ntotal 32
    (12, 6, 21, 10, 1.0, 4.0),   # ntotal 40
    (8, 4, 33, 8, 1.0, 0.5),     # ntotal 8 < nshift: fft(x, n) keeps the first 8 samples
    (10, 3, 13, 6, 0.75, 4.0),   # rsupp 0.75: dshift 1/8, ntotal 32
    (1, 1, 9, 4, 1.0, 2.0),      # one direction
]


def _max_scale_err(got, want):
    return np.max(np.abs(got - want)) / np.max(np.abs(want))


@pytest.mark.parametrize("case", CASES, ids=[str(c) for c in CASES])
def test_inversion_matches_oracle_on_random_data(ctx, case):
    nphi, ntheta, nshift, ngrid, rsupp, padding = case
    print("fourier_inversion random", case, flush=True)
    rng = np.random.default_rng(nphi * 1000 + nshift)
    radon = rng.standard_normal((nshift, nphi, ntheta))
    phi = (2 * math.pi / nphi) * np.arange(nphi)
    x, w = np.polynomial.legendre.leggauss(ntheta)
    theta = np.arccos(x)
    got, st = ctx.fourier_inversion(radon, phi, theta, w, ngrid, w0=1.5, rsupp=rsupp,
                                    padding=padding)
    want = orc.chang3d(radon, phi, theta, w, ngrid, w0=1.5, rsupp=rsupp, padding=padding)
    assert orc.rel_l2(got, want) <= 1e-11, orc.rel_l2(got, want)
    assert _max_scale_err(got, want) <= 1e-10, _max_scale_err(got, want)
    ntotal = int(padding * math.floor((nshift - 1) / (2 * rsupp)))
    assert st["outputs"] == ngrid ** 3
    assert st["logical_samples"] == ngrid ** 3 * nphi * ntheta * ntotal


def test_inversion_reconstructs_an_off_centre_gaussian(ctx):
    # closed-form plane integrals in, the Gaussian out up to the discretisation error (the same
    # bound the oracle meets at this size), and the oracle's values to rounding
    print("fourier_inversion gaussian 48x24x65 -> 24^3", flush=True)
    radon, phi, theta, w = fourier_case(48, 24, 65)
    got, st = ctx.fourier_inversion(radon, phi, theta, w, 24)
    want = gauss_on_grid(24)
    assert np.max(np.abs(got - want)) <= 5e-4
    assert orc.rel_l2(got, want) <= 5e-4
    ref = orc.chang3d(radon, phi, theta, w, 24)
    assert orc.rel_l2(got, ref) <= 1e-11
    assert st["kernel_seconds"] > 0


def test_analytic_plane_integrals_then_inversion(ctx):
    # the reference's chain: radon_analytic of the exponential bump on the inversion's grid,
    # then chang3d; the reconstruction is the bump itself (relative L2 0.012 at this size)
    print("radon_analytic -> fourier_inversion 32x16x65 -> 16^3", flush=True)
    terms = example_terms("exponential_bump")
    phi = orc.phi_grid(32)
    theta, w = orc.theta_grid(16)
    s = orc.shift_grid(65)
    radon, _ = ctx.radon_analytic(terms, 65, phi, theta, s)
    got, _ = ctx.fourier_inversion(radon, phi, theta, w, 16)
    x = (np.arange(16) - 8) * (2.0 / 16)
    want = np.array([[[orc.fn_eval(terms, a, b, c) for c in x] for b in x] for a in x])
    assert orc.rel_l2(got, want) <= 0.02, orc.rel_l2(got, want)
    assert np.max(np.abs(got - want)) <= 2e-3
    assert orc.rel_l2(got, orc.chang3d(radon, phi, theta, w, 16)) <= 1e-11


def test_inversion_rejects_bad_arguments(ctx):
    radon = np.zeros((17, 4, 2))
    phi = np.arange(4) * (math.pi / 2)
    theta, w = np.array([1.0, 2.0]), np.array([1.0, 1.0])
    for kw in [dict(ngrid=7), dict(ngrid=8, padding=0.3), dict(ngrid=8, w0=0.0),
               dict(ngrid=8, rsupp=-1.0), dict(ngrid=8, padding=1.125)]:
        with pytest.raises(wrt.WrtError):
            ctx.fourier_inversion(radon, phi, theta, w, **kw)


def test_chang3d_reads_the_reference_file_format(ctx, tmp_path):
    # Radon data written as the reference's "%lf, %lf, %lf, %lf" file (s -> phi -> theta, theta
    # ascending as rtft3d.m's grid has it), read back and inverted by the mirror of chang3d
    nphi, ntheta, nshift = 32, 16, 65
    print("chang3d from a text file", (nphi, ntheta, nshift), flush=True)
    radon, phi, theta, w = fourier_case(nphi, ntheta, nshift)
    radon, theta, w = radon[:, :, ::-1], theta[::-1], w[::-1]  # cos(theta) descending
    s = -1.0 + np.arange(nshift) * (2.0 / (nshift - 1))
    path = tmp_path / "radon.txt"
    wrt.write_text_grid(path, s, phi, theta, radon)
    got = ctx.chang3d(16, 2.0, str(path), nphi, ntheta, nshift)
    text = wrt.read_text(path, radon.size, 4).reshape(radon.shape)
    want, _ = ctx.fourier_inversion(text, phi, theta, w, 16, w0=2.0)
    assert orc.rel_l2(got, want) <= 1e-13
    assert np.max(np.abs(2.0 * got - gauss_on_grid(16))) <= 0.01


CASES2D = [
    (16, 17, 8, 1.0, 4.0),      # ntotal 32
    (12, 33, 12, 1.0, 2.0),     # ntotal 32
    (64, 65, 32, 1.0, 4.0),     # ntotal 128
    (7, 33, 6, 1.0, 0.5),       # ntotal 8 < nshift
    (20, 13, 10, 0.75, 8.0),    # rsupp 0.75, ntotal 64
    (1, 9, 4, 1.0, 2.0),        # one direction
]


@pytest.mark.parametrize("case", CASES2D, ids=[str(c) for c in CASES2D])
def test_inversion2d_matches_oracle_on_random_data(ctx, case):
    nphi, nshift, ngrid, rsupp, padding = case
    print("fourier_inversion2d random", case, flush=True)
    rng = np.random.default_rng(nphi * 1000 + nshift + 7)
    radon = rng.standard_normal((nshift, nphi))
    phi = (2 * math.pi / nphi) * np.arange(nphi)
    got, st = ctx.fourier_inversion2d(radon, phi, ngrid, w0=0.75, rsupp=rsupp, padding=padding)
    want = orc.chang2d(radon, phi, ngrid, w0=0.75, rsupp=rsupp, padding=padding)
    assert orc.rel_l2(got, want) <= 1e-11, orc.rel_l2(got, want)
    assert _max_scale_err(got, want) <= 1e-10, _max_scale_err(got, want)
    assert st["outputs"] == ngrid ** 2


def test_inversion2d_reconstructs_an_off_centre_gaussian(ctx):
    print("fourier_inversion2d gaussian 256 phi x 257 s -> 128^2, padding 16", flush=True)
    radon, phi = fourier_case2d(256, 257)
    got, _ = ctx.fourier_inversion2d(radon, phi, 128, padding=16.0)
    want = gauss2d_on_grid(128)
    assert np.max(np.abs(got - want)) <= 4e-4
    assert orc.rel_l2(got, want) <= 3e-3


def test_chang2d_reads_the_reference_file_format(ctx, tmp_path):
    nphi, nshift = 64, 129
    print("chang2d from a text file", (nphi, nshift), flush=True)
    radon, phi = fourier_case2d(nphi, nshift)
    s = -1.0 + np.arange(nshift) * (2.0 / (nshift - 1))
    S, P = np.meshgrid(s, phi, indexing="ij")
    path = tmp_path / "radon2d.txt"
    with open(path, "w") as f:
        for a, b, v in zip(S.ravel(), P.ravel(), radon.ravel()):
            f.write("%f, %f, %f\n" % (a, b, v))
    got = ctx.chang2d(32, 2.0, str(path), nphi, nshift, 1.0, 16.0)
    text = wrt.read_text(path, radon.size, 3).reshape(radon.shape)
    assert np.max(np.abs(text - radon)) <= 5e-7
    want, _ = ctx.fourier_inversion2d(text, phi, 32, w0=2.0, padding=16.0)
    assert orc.rel_l2(got, want) <= 1e-13
    assert np.max(np.abs(2.0 * got - gauss2d_on_grid(32))) <= 1e-3


def test_inversion2d_rejects_bad_arguments(ctx):
    radon = np.zeros((17, 4))
    phi = np.arange(4) * (math.pi / 2)
    for kw in [dict(ngrid=7), dict(ngrid=8, padding=0.3), dict(ngrid=8, w0=0.0),
               dict(ngrid=8, rsupp=0.0), dict(ngrid=8, padding=1.125)]:
        with pytest.raises(wrt.WrtError):
            ctx.fourier_inversion2d(radon, phi, **kw)
