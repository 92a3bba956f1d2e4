"""GPU parity of the spectral Laplacians (wrt_laplace3d / wrt_laplace2d, FP64) against the
numpy restatement of laplace3dfft.m / laplace2dfft.m (oracle/spectral.py, itself pinned by
closed forms in test_oracle.py): relative L2 and max error <= 1e-12 of the output's scale."""
import importlib.util
import os

import numpy as np
import pytest

import orc
import wrt

pytestmark = pytest.mark.gpu
_path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                     "spectral.py")
_spec = importlib.util.spec_from_file_location("spectral", _path)
spectral = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(spectral)


@pytest.mark.parametrize("n,period", [(16, 2.0), (33, 2.0), (64, 2.5), (48, 1.0)])
def test_laplace3d_matches_restatement(ctx, n, period):
    u = wrt.gaussian((n, n, n), seed=n)
    got, st = ctx.laplace(u, period)
    want = spectral.laplace_fft(u, period)
    assert orc.rel_l2(got, want) <= 1e-12, orc.rel_l2(got, want)
    assert orc.max_rel(got, want) <= 1e-12, orc.max_rel(got, want)
    assert st["outputs"] == n ** 3


def test_laplace3d_of_a_backprojection(ctx):
    # the use it has in the reference: the Laplacian of an adjoint (backprojection) volume
    n = 32
    phi, th, w, s = ctx.radon_grid_init(16, 8, n + 1)
    radon, _ = ctx.radon_grid(wrt.phantom(n, True, 10, 5, 0.0), phi, th, s)
    adj, _ = ctx.radon_adjoint(radon, phi, th, w, s, n)
    got, _ = ctx.laplace(adj, 2.0)
    want = spectral.laplace_fft(adj, 2.0)
    assert orc.rel_l2(got, want) <= 1e-12
    assert orc.max_rel(got, want) <= 1e-12


@pytest.mark.parametrize("n", [31, 64, 200])
def test_laplace2d_matches_restatement(ctx, n):
    u = wrt.gaussian((n, n), seed=n)
    got, _ = ctx.laplace(u, 2.0)
    want = spectral.laplace_fft(u, 2.0)
    assert orc.rel_l2(got, want) <= 1e-12
    assert orc.max_rel(got, want) <= 1e-12


def test_laplace_plan_reuse_and_errors(ctx):
    for n in (16, 16, 20, 16):  # same size reuses the plan, a new size replaces it
        u = wrt.gaussian((n, n, n), seed=1)
        got, _ = ctx.laplace(u, 2.0)
        assert orc.rel_l2(got, spectral.laplace_fft(u, 2.0)) <= 1e-12
    for period in (0.0, -1.0, float("nan")):
        with pytest.raises(wrt.WrtError) as e:
            ctx.laplace(np.zeros((4, 4, 4)), period)
        assert e.value.code == wrt.WRT_E_ARG
