"""GPU parity of the plain FP64 forms (WRT_F64) against the oracle and the golden vectors,
through the C-ABI (lib/libwrt.so).  Bar: bit-exact where the kernel sums in the reference's
order (ray, reduction, adjoint, GL); relative L2 and max-relative error <= 1e-10 for the
plane integrals, whose 64-lane tree sum only reorders the additions."""
import math
import os

import numpy as np
import pytest

import orc
import wrt
from golden_inputs import volume_formula

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-10


def _npz(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def test_gl_nodes_gpu_bitwise(ctx):
    f = _npz("gl.npz")
    for k in [k for k in f.files if k.startswith("x")]:
        n = int(k[1:])
        x, w = ctx.gl_nodes(n)
        assert np.array_equal(x, f[k]), n
        assert np.array_equal(w, f["w%d" % n]), n


def test_radon_grid_init(ctx):
    phi, th, w, s = ctx.radon_grid_init(64, 32, 129)
    t2, w2 = orc.theta_grid(32)
    assert np.array_equal(phi, orc.phi_grid(64))
    assert np.array_equal(th, t2) and np.array_equal(w, w2)
    assert np.array_equal(s, orc.shift_grid(129))


def _uniq(a):
    """distinct values in order of first appearance"""
    _, idx = np.unique(a, return_index=True)
    return a[np.sort(idx)]


def _unique_grid(phi, theta, shift):
    up, ut, us = _uniq(phi), _uniq(theta), _uniq(shift)
    if len(up) * len(ut) * len(us) != len(phi):
        return None
    return up, ut, us


def test_radon_f64_golden(ctx):
    f = _npz("ref_radon_grid.npz")
    for k in [k for k in f.files if k.startswith("radon_") and not k.endswith(("_in", "_n"))]:
        n = int(f[k + "_n"])
        phi, theta, shift = f[k + "_in"]
        vol = volume_formula(n)
        g = _unique_grid(phi, theta, shift)
        if g is not None:  # golden order is (phi, theta, shift); product order [s][phi][theta]
            up, ut, us = g
            out, _ = ctx.radon_grid(vol, up, ut, us)
            got = out.transpose(1, 2, 0).ravel()
        else:
            got = np.array([ctx.radon_grid(vol, [p], [t], [s])[0].item()
                            for p, t, s in zip(phi, theta, shift)])
        want = f[k]
        assert orc.rel_l2(got, want) <= TOL and orc.max_rel(got, want) <= TOL, k


def test_radon_f64_c0_parity(ctx):
    # 64^3 cell-centred phantom of 10 ellipsoids, 64 phi x 32 GL theta x 129 shifts (full
    # grid on the GPU), oracle on a seeded subset of planes
    n = 64
    vol = wrt.phantom(n, True, 10, 0, 0.0)
    phi, th, w, s = ctx.radon_grid_init(64, 32, 129)
    out, st = ctx.radon_grid(vol, phi, th, s)
    assert out.shape == (129, 64, 32)
    rng = np.random.default_rng(0)
    idx = rng.choice(out.size, 384, replace=False)
    iS, iP, iT = np.unravel_index(idx, out.shape)
    want = orc.radon_planes(vol, phi[iP], th[iT], s[iS])
    got = out.ravel()[idx]
    assert orc.rel_l2(got, want) <= TOL and orc.max_rel(got, want) <= TOL
    # structured planes (phi = 0, pi/2, pi, 3pi/2) where index decisions hinge on rounding
    for ip in (0, 16, 32, 48):
        want = orc.radon_planes(vol, np.full(32, phi[ip]), th, np.full(32, s[40]))
        assert orc.rel_l2(out[40, ip, :], want) <= TOL
    # full-size property: sum_s R ds equals the volume's mass for every direction
    mass = vol.sum() * (2.0 / n) ** 3
    ms = out.sum(axis=0) * (2.0 / 128)
    assert np.max(np.abs(ms - mass)) < 0.02 * abs(mass)
    assert st["logical_samples"] == 16.0 * n * n * out.size


def test_ray_f64_golden_bitwise(ctx):
    f = _npz("ref_ray_grid.npz")
    for k in [k for k in f.files if k.startswith("ray_") and not k.endswith(("_in", "_n"))]:
        n = int(f[k + "_n"])
        z, s, phi = f[k + "_in"]
        vol = volume_formula(n)
        uz, us, up = _uniq(z), _uniq(s), _uniq(phi)
        if len(uz) * len(us) * len(up) == len(z):
            out, _ = ctx.ray_grid(vol, uz, us, up)
            got = out.ravel()  # golden meshgrid order (z, s, phi) == product order
        else:
            got = np.array([ctx.ray_grid(vol, [a], [b], [c])[0].item() for a, b, c in zip(z, s, phi)])
        assert np.array_equal(got, f[k]), k


def test_ray_f64_c1_parity(ctx):
    # 128^3 node-centred phantom, 128 z x 180 phi in [0, pi) x 257 shifts
    n = 128
    vol = wrt.phantom(n, False, 10, 1, 0.0)
    z = wrt.zslice_grid(128)
    s, _ = wrt.ray_grid_init(257, 1)
    phi = (math.pi / 180) * np.arange(180)
    out, st = ctx.ray_grid(vol, z, s, phi)
    rng = np.random.default_rng(1)
    idx = rng.choice(out.size, 1500, replace=False)
    iz, i_s, ip = np.unravel_index(idx, out.shape)
    want = orc.rays(vol, z[iz], s[i_s], phi[ip])
    assert np.array_equal(out.ravel()[idx], want)
    # the phi = 0 rays carry the truncation/fmod index quirk
    want = orc.rays(vol, z[::8], np.full(16, s[100]), np.zeros(16))
    assert np.array_equal(out[::8, 100, 0], want)


def test_reduction_f64_bitwise(ctx):
    nz, ns, nphi, nth = 33, 33, 16, 8
    vol = wrt.phantom(nz, False, 10, 2, 0.0)
    z = wrt.zslice_grid(nz)
    s, phi = wrt.ray_grid_init(ns, nphi)
    rays = orc.ray_grid(vol, z, s, phi)
    th, _ = orc.theta_grid(nth)
    out, st = ctx.radon_reduction(rays, s, phi, th)
    want = orc.radon_reduction(rays, s, phi, th)
    assert np.array_equal(out, want)


def test_reduction_f64_c1_chain(ctx):
    # C1 parity chain: 128 z x 64 phi in [0, 2pi) x 129 s rays -> 64 phi x 32 theta x 129 s
    nz, ns, nphi, nth = 128, 129, 64, 32
    vol = wrt.phantom(nz, False, 10, 3, 0.0)
    z = wrt.zslice_grid(nz)
    s, phi = wrt.ray_grid_init(ns, nphi)
    rays, _ = ctx.ray_grid(vol, z, s, phi)
    _, th, _, _ = ctx.radon_grid_init(nphi, nth, ns)
    out, st = ctx.radon_reduction(rays, s, phi, th)
    rng = np.random.default_rng(5)
    for _ in range(24):
        i, j, k = rng.integers(ns), rng.integers(nphi), rng.integers(nth)
        want = orc.L().orc_radon_reduction_one(orc._p(np.ascontiguousarray(rays)), nz, orc._p(s), ns,
                                               orc._p(phi), nphi, orc._p(th), nth, int(i), int(j), int(k))
        assert out[i, j, k] == want


def test_adjoint_f64_bitwise(ctx):
    n, nphi, nth, ns = 12, 16, 8, 33
    phi, th, w, s = ctx.radon_grid_init(nphi, nth, ns)
    vol = wrt.phantom(16, True, 10, 4, 0.0)
    radon = orc.radon_grid(vol, phi, th, s)
    out, st = ctx.radon_adjoint(radon, phi, th, w, s, n)
    want = orc.radon_adjoint(radon, phi, th, w, s, n)
    assert np.array_equal(out, want)
    c = np.full_like(radon, 0.5)
    out, _ = ctx.radon_adjoint(c, phi, th, w, s, n)
    xs = -1.0 + 1.0 / n + np.arange(n) * (2.0 / n)
    Z, X, Y = np.meshgrid(xs, xs, xs, indexing="ij")
    inside = X ** 2 + Y ** 2 + Z ** 2 <= 1.0
    assert np.max(np.abs(out[inside] - 2 * math.pi)) < 1e-12


def test_transposes_f64_dot_and_oracle(ctx):
    rng = np.random.default_rng(8)
    n = 16
    phi, th, w, s = ctx.radon_grid_init(12, 6, 17)
    f = rng.standard_normal((n, n, n))
    g = rng.standard_normal((17, 12, 6))
    Af, _ = ctx.radon_grid(f, phi, th, s)
    ATg, _ = ctx.radon_grid_T(g, n, phi, th, s)
    assert abs(np.sum(Af * g) - np.sum(f * ATg)) <= 1e-12 * np.linalg.norm(Af) * np.linalg.norm(g)
    assert orc.rel_l2(ATg, orc.radon_grid_T(g, n, phi, th, s)) <= 1e-13
    n = 17
    z = wrt.zslice_grid(9)
    s, phi = wrt.ray_grid_init(17, 12)
    f = rng.standard_normal((n, n, n))
    g = rng.standard_normal((9, 17, 12))
    Af, _ = ctx.ray_grid(f, z, s, phi)
    ATg, _ = ctx.ray_grid_T(g, n, z, s, phi)
    assert abs(np.sum(Af * g) - np.sum(f * ATg)) <= 1e-12 * np.linalg.norm(Af) * np.linalg.norm(g)
    assert orc.rel_l2(ATg, orc.ray_grid_T(g, n, z, s, phi)) <= 1e-13


def test_incube_count_matches_oracle(ctx):
    for n, (nphi, nth, ns) in ((16, (8, 4, 9)), (33, (6, 3, 7)), (64, (4, 4, 5))):
        phi, th, w, s = ctx.radon_grid_init(nphi, nth, ns)
        want = sum(orc.plane_incube(n, p, t, v) for v in s for p in phi for t in th)
        assert ctx.radon_count_incube(n, phi, th, s) == want


def test_invalid_arguments(ctx):
    with pytest.raises(wrt.WrtError) as e:
        ctx.radon_grid(np.zeros((4, 4, 4)), [], [1.0], [0.0])
    assert e.value.code == wrt.WRT_E_ARG
    with pytest.raises(wrt.WrtError) as e:
        ctx.radon_grid(np.zeros((4, 4, 4)), [float("nan")], [1.0], [0.0])
    assert e.value.code == wrt.WRT_E_ARG
    with pytest.raises(wrt.WrtError) as e:
        ctx.radon_reduction(np.zeros((2, 5, 3)), np.linspace(-1, 1, 5), np.zeros(3), np.ones(2))
    assert e.value.code == wrt.WRT_E_ARG
