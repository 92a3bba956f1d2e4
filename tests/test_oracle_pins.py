"""CPU pins of the oracle's backprojection (orientation, sign, [z][x][y] output order) and
reduction (forward-biased z stencil, zeroed third slice in the top cell) by closed forms
(tests/closed_forms.py), and planted mistakes that those checks must catch.

Tolerances: 1e-12 absolute against closed forms of size O(1): the forms are exact up to
the rounding of the splines, the Lagrange interpolation and the sums."""
import math

import numpy as np
import pytest

import closed_forms as cf
import orc


def _adjoint_grid(nphi=12, nth=6, ns=17):
    phi = orc.phi_grid(nphi)
    th, w = orc.theta_grid(nth)
    s = orc.shift_grid(ns)
    return phi, th, w, s


@pytest.mark.parametrize("k", [0, 1, 2])
def test_adjoint_axis_data_gives_coordinate(k):
    n = 8
    phi, th, w, s = _adjoint_grid()
    g = cf.adjoint_axis_data(s, phi, th, k)
    adj = orc.radon_adjoint(g, phi, th, w, s, n)
    want, inside = cf.adjoint_axis_expected(n, k)
    assert inside.sum() > 100
    err = np.max(np.abs(adj[inside] - want[inside]))
    assert err < 1e-12, err


def test_adjoint_axis_checks_catch_planted_mistakes():
    # answers with a wrong orientation or output order fail the closed form on some axis
    n = 8
    phi, th, w, s = _adjoint_grid()
    planted = {
        "output [x][y][z]": lambda a: a.transpose(1, 2, 0),
        "x and y swapped": lambda a: a.transpose(0, 2, 1),
        "z axis flipped": lambda a: a[::-1, :, :],
        "sign of the direction": lambda a: -a,
    }
    for name, mutate in planted.items():
        worst = 0.0
        for k in range(3):
            adj = mutate(orc.radon_adjoint(cf.adjoint_axis_data(s, phi, th, k), phi, th, w, s, n))
            want, inside = cf.adjoint_axis_expected(n, k)
            worst = max(worst, np.max(np.abs(adj[inside] - want[inside])))
        assert worst > 0.1, name


def _reduction_grid(nz=17, ns=33, nphi=3):
    s = orc.shift_grid(ns)
    phi = orc.phi_grid(nphi)
    th, _ = orc.theta_grid(8)
    # exact theta = 0 and pi/2 planes beside the Gauss-Legendre ones
    th = np.r_[0.0, th, math.pi / 2]
    return nz, s, phi, th


def test_reduction_constant_rays_top_cell():
    nz, s, phi, th = _reduction_grid()
    c = 0.625
    R = orc.radon_reduction(cf.constant_rays(nz, len(s), len(phi), c), s, phi, th)
    want = cf.reduction_constant(c, s, th, nz)
    for k in range(len(phi)):
        assert np.max(np.abs(R[:, k, :] - want)) < 1e-12
    # theta = 0: z = p on the whole plane, so the top cell's factor 1 - l2(p) is visible alone
    top = cf.reduction_constant(c, s, th, nz, zero_top=False)
    assert np.max(np.abs(want - top)) > 1e-3


def test_reduction_cubic_rays_forward_stencil():
    nz, s, phi, th = _reduction_grid()
    R = orc.radon_reduction(cf.cubic_rays(nz, len(s), len(phi)), s, phi, th)
    want = cf.reduction_cubic(s, th, nz)
    for k in range(len(phi)):
        assert np.max(np.abs(R[:, k, :] - want)) < 1e-12
    # on the theta = pi/2 planes away from the top cell the forward stencil's remainder is
    # negative and a centred stencil's positive
    mid = (np.abs(s) > 0.7) & (np.abs(s) < 0.99)  # |p| = 1 touches the ball at a node only
    fwd = want[mid, -1]
    cen = cf.reduction_cubic(s[mid], th[-1:], nz, centred=True)[:, 0]
    assert np.all(fwd < -1e-6) and np.all(cen > 1e-6)


def test_reduction_checks_catch_planted_mistakes():
    # restatements of the z interpolation with a planted mistake fail the closed forms
    nz, s, phi, th = _reduction_grid()
    c = 0.625
    good_c = cf.reduction_constant(c, s, th, nz)
    good_z = cf.reduction_cubic(s, th, nz)
    no_zero = cf.reduction_constant(c, s, th, nz, zero_top=False)
    centred = cf.reduction_cubic(s, th, nz, centred=True)
    assert np.max(np.abs(no_zero - good_c)) > 1e-3
    assert np.max(np.abs(centred - good_z)) > 1e-4
