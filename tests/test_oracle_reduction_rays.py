"""Pins the oracle's reduction on independent ray and plane grids
(orc_radon_reduction_rays, no reference counterpart) before it checks the GPU:

- on identical full-turn grids it is the reference-form reduction, bit for bit (that form
  is pinned by the reference's own functions and closed forms in test_oracle.py);
- half-turn rays (angles in [0, pi)) give, at azimuths in [pi, 2 pi), what the full-turn
  rays give: P(z, s, phi + pi) = P(z, -s, phi) with the row read at -s;
- the phi stencil is exact on cubics in phi, on both sides of the half turn;
- an off-centre Gaussian, whose line and plane integrals are closed forms, reduces from
  180 half-turn angles to planes at arbitrary azimuths within the discretisation error;
- a mistake planted in the stencil (no s -> -s, a one-angle index shift, linear instead of
  cubic weights) fails that closed-form check.
"""
import math

import numpy as np
import pytest

import orc
from golden_inputs import gauss_planes, gauss_rays, volume_formula


def _small_rays(nz=17, ns=17, nphi=8):
    vol = volume_formula(nz)
    z, s, phi = orc.zslice_grid(nz), orc.shift_grid(ns), orc.phi_grid(nphi)
    return orc.ray_grid(vol, z, s, phi), z, s, phi


def test_identical_grids_equal_reference_form_bitwise():
    rays, z, s, phi = _small_rays()
    th, _ = orc.theta_grid(4)
    got = orc.radon_reduction_rays(rays, s, False, s, phi, th)
    want = orc.radon_reduction(rays, s, phi, th)
    assert np.array_equal(got, want)


def test_half_turn_rays_reflect_onto_the_full_turn():
    # full-turn rays built from half-turn ones by P(z, s, phi + pi) = P(z, -s, phi) (the
    # reference's trilinear ray integrator itself breaks that symmetry on axis-aligned lines
    # that run along lattice planes, so its rays are not used for this check)
    m = 6
    rays, z, s, phi = _small_rays(nphi=2 * m)
    th, _ = orc.theta_grid(4)
    half = np.ascontiguousarray(rays[:, :, :m])
    full = np.ascontiguousarray(np.concatenate([half, half[:, ::-1, :]], axis=2))
    got = orc.radon_reduction_rays(half, s, True, s, phi, th)
    want = orc.radon_reduction(full, s, phi, th)
    # azimuths below pi read the same rows
    assert np.array_equal(got[:, :m], want[:, :m])
    # above pi: the half-turn row's spline at -s against the reflected row's at s
    scale = np.max(np.abs(want))
    assert np.max(np.abs(got[:, m:] - want[:, m:])) <= 1e-13 * scale
    # and the reflection is what makes them agree: the unreflected rows do not
    wrong = orc.radon_reduction(np.ascontiguousarray(np.concatenate([half, half], axis=2)),
                                s, phi, th)
    assert np.max(np.abs(got[:, m:] - wrong[:, m:])) > 1e-3 * scale


def _cubic(ph):
    return 1.0 + 0.7 * ph - 0.3 * ph ** 2 + 0.05 * ph ** 3


@pytest.mark.parametrize("half_turn", [False, True])
def test_stencil_exact_on_cubics_in_phi(half_turn):
    nz, ns, nr = 17, 17, 12
    z, s = orc.zslice_grid(nz), orc.shift_grid(ns)
    h = gauss_rays(z, s, [0.0], c=(0.0, 0.0, 0.1), sigma=0.35)  # even in s
    period = math.pi if half_turn else 2 * math.pi
    dphi = period / nr
    ang = dphi * np.arange(nr)
    rays = np.ascontiguousarray(h * _cubic(ang)[None, None, :])
    th, _ = orc.theta_grid(3)
    rng = np.random.default_rng(3)
    # azimuths whose stencils stay inside one period of the unrolled grid
    base = rng.uniform(1.5 * dphi, period - 2.5 * dphi, 6)
    phis = np.concatenate([base, base + math.pi]) if half_turn else base
    got = orc.radon_reduction_rays(rays, s, half_turn, s, phis, th)
    flat = np.ascontiguousarray(np.broadcast_to(h, (nz, ns, len(phis))))
    r_h = orc.radon_reduction(flat, s, phis, th)
    loc = np.mod(phis, period)
    want = r_h * _cubic(loc)[None, :, None]
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


def _gauss_case():
    nz, ns, nr = 64, 129, 90
    z, rs = orc.zslice_grid(nz), orc.shift_grid(ns)
    ang = (math.pi / nr) * np.arange(nr)
    rays = np.ascontiguousarray(gauss_rays(z, rs, ang))
    p = orc.shift_grid(33)
    rng = np.random.default_rng(11)
    phis = np.sort(rng.uniform(0, 2 * math.pi, 12))
    th, _ = orc.theta_grid(6)
    # the closed-form rays on each azimuth itself, reduced by the reference form: what the
    # phi stencil approximates
    exact = _reduce_rows(gauss_rays(z, rs, phis), rs, p, th)
    return rays, rs, p, phis, th, gauss_planes(p, phis, th), exact


def _reduce_rows(rows, rs, p, th):
    # rows[z][s][j] already on azimuth j: each is reduced on its own, as a 4-angle ray grid
    # whose azimuth 0 takes row 0 alone
    out = np.empty((len(p), rows.shape[2], len(th)))
    for j in range(rows.shape[2]):
        one = np.ascontiguousarray(np.repeat(rows[:, :, j:j + 1], 4, axis=2))
        out[:, j, :] = orc.radon_reduction_rays(one, rs, False, p, [0.0], th)[:, 0, :]
    return out


def test_gaussian_closed_form_from_half_turn_rays():
    rays, rs, p, phis, th, want, exact = _gauss_case()
    got = orc.radon_reduction_rays(rays, rs, True, p, phis, th)
    # z / shift discretisation of the reference form at 64 slices and 129 shifts
    assert np.max(np.abs(exact - want)) / np.max(want) < 1e-3
    assert np.max(np.abs(got - want)) / np.max(want) < 1e-3
    # the phi stencil on 90 angles in [0, pi): cubic interpolation error only
    err = np.max(np.abs(got - exact)) / np.max(want)
    assert err < 4e-6, err


def _planted_rows(rays, phis, planted):
    # the stencil restated with a planted mistake, applied to the half-turn rays on the CPU
    nr = rays.shape[2]
    dphi = math.pi / nr
    node = [-1.0, 0.0, 1.0, 2.0]
    rows = np.empty(rays.shape[:2] + (len(phis),))
    for j, ph in enumerate(phis):
        a = math.fmod(ph, 2 * math.pi)
        flip = a >= math.pi
        if flip:
            a -= math.pi
        u = a / dphi
        k = int(math.floor(u))
        fr = u - k
        if planted == "index_shift":
            k += 1
        if planted == "linear":
            ks, ws = [k, k + 1], [1 - fr, fr]
        else:
            ks = [k - 1, k, k + 1, k + 2]
            ws = [np.prod([(fr - node[i]) / (node[m] - node[i]) for i in range(4) if i != m])
                  for m in range(4)]
        acc = np.zeros(rays.shape[:2])
        for kk, w in zip(ks, ws):
            q, idx = divmod(kk, nr)
            f = (flip ^ (q % 2 == 1)) and planted != "no_reflection"
            r = rays[:, :, idx]
            acc = acc + w * (r[:, ::-1] if f else r)
        rows[:, :, j] = acc
    return rows


@pytest.mark.parametrize("planted", ["none", "no_reflection", "index_shift", "linear"])
def test_phi_check_catches_planted_stencil_mistakes(planted):
    # without a mistake the CPU restatement meets the same bound as the oracle; each planted
    # mistake fails it
    rays, rs, p, phis, th, want, exact = _gauss_case()
    got = _reduce_rows(_planted_rows(rays, phis, planted), rs, p, th)
    err = np.max(np.abs(got - exact)) / np.max(want)
    if planted == "none":
        assert err < 4e-6, err
    else:
        assert err > 4e-6, err
