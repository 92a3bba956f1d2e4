"""GPU pins of the backprojection's orientation and output order and of the reduction's z
stencil, through the C-ABI, by the closed forms of tests/closed_forms.py (the CPU side of the
same checks is tests/test_oracle_pins.py).

Tolerances: FP64 forms 1e-12 absolute (values O(1)); the FP32 backprojection 1e-5 of the
largest value (FP32 cubics and directions)."""
import math

import numpy as np
import pytest

import closed_forms as cf

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [0, 1, 2])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_adjoint_axis_data_gives_coordinate(ctx, k, prec):
    n = 24
    phi, th, w, s = ctx.radon_grid_init(12, 6, 33)
    g = cf.adjoint_axis_data(s, phi, th, k)
    adj, _ = ctx.radon_adjoint(g, phi, th, w, s, n, prec=prec)
    want, inside = cf.adjoint_axis_expected(n, k)
    err = np.max(np.abs(adj[inside] - want[inside]))
    tol = 1e-12 if prec == "f64" else 1e-5 * 4 * math.pi / 3
    assert err < tol, (prec, k, err)


def test_reduction_constant_rays_top_cell(ctx):
    nz, ns, nphi = 17, 33, 4
    phi, th, _, s = ctx.radon_grid_init(nphi, 8, ns)
    th = np.r_[0.0, th, math.pi / 2]
    c = 0.625
    R, _ = ctx.radon_reduction(cf.constant_rays(nz, ns, nphi, c), s, phi, th)
    want = cf.reduction_constant(c, s, th, nz)
    for k in range(nphi):
        assert np.max(np.abs(R[:, k, :] - want)) < 1e-12


def test_reduction_cubic_rays_forward_stencil(ctx):
    nz, ns, nphi = 17, 33, 4
    phi, th, _, s = ctx.radon_grid_init(nphi, 8, ns)
    th = np.r_[0.0, th, math.pi / 2]
    R, _ = ctx.radon_reduction(cf.cubic_rays(nz, ns, nphi), s, phi, th)
    want = cf.reduction_cubic(s, th, nz)
    for k in range(nphi):
        assert np.max(np.abs(R[:, k, :] - want)) < 1e-12
