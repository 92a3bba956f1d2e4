"""CPU tests of the C-ABI library: it loads, exports every symbol include/wrt.h declares,
and its host-only parts (grids, synthetic inputs, argument checks) behave without a GPU."""
import ctypes as C
import math
import os
import re

import numpy as np

import orc
import wrt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "wrt.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(wrt_[A-Za-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = wrt.lib()
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(set(wrt.ABI_SYMBOLS)) == names


def test_host_grids_match_oracle_bits():
    s, p = wrt.ray_grid_init(257, 180)
    assert np.array_equal(s, orc.shift_grid(257))
    assert np.array_equal(p, orc.phi_grid(180))
    assert np.array_equal(wrt.zslice_grid(128), orc.zslice_grid(128))


def test_phantom_properties():
    n = 32
    v = wrt.phantom(n, cell_centred=True, nell=10, seed=0, sigma=0.0)
    v2 = wrt.phantom(n, cell_centred=True, nell=10, seed=0, sigma=0.0)
    assert np.array_equal(v, v2)
    xs = -1.0 + (2.0 * np.arange(n) + 1.0) / n
    X, Y, Z = np.meshgrid(xs, xs, xs, indexing="ij")
    outside = X ** 2 + Y ** 2 + Z ** 2 > 1.0
    assert np.all(v[outside] == 0.0)
    assert np.count_nonzero(v) > 0
    P = wrt.phantom_params(10, seed=0)
    assert np.all(np.linalg.norm(P[:, 0:3], axis=1) < 0.5)
    assert np.all((P[:, 3:6] >= 0.05) & (P[:, 3:6] <= 0.4))
    assert np.all((P[:, 15] >= -0.5) & (P[:, 15] <= 1.0))
    for e in range(10):
        R = P[e, 6:15].reshape(3, 3)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-12)
        assert abs(np.linalg.det(R) - 1.0) < 1e-12
    t = wrt.phantom(n, cell_centred=False, nell=10, seed=0, sigma=0.01)
    assert not np.array_equal(t, wrt.phantom(n, cell_centred=False, nell=10, seed=0))


def test_gaussian_moments():
    g = wrt.gaussian((200000,), seed=3)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1.0) < 0.01
    assert np.array_equal(g[:100], wrt.gaussian((100,), seed=3))


def test_job_sizes():
    phi = np.zeros(4)
    th = np.zeros(3)
    s = np.linspace(-1, 1, 5)
    assert wrt.job_bytes("radon", "f64", n=8, phi=phi, theta=th, shift=s) > 8 ** 3 * 8
    assert wrt.job_bytes("radon", "f64", n=0, phi=phi, theta=th, shift=s) == 0
    assert wrt.job_bytes("reduction", "f32", phi=phi, theta=th, shift=s, nz=8) == 0
    # the FP32 adjoint needs a uniform shift grid (it locates intervals arithmetically)
    w = np.ones(3)
    assert wrt.job_bytes("adjoint", "f32", n=8, phi=phi, theta=th, theta_w=w, shift=s) > 0
    s2 = s.copy()
    s2[2] = 0.1
    assert wrt.job_bytes("adjoint", "f32", n=8, phi=phi, theta=th, theta_w=w, shift=s2) == 0


def test_shift_order_contract():
    # plane and line integrals take shifts in any order (plane_integral and ray_transform
    # take any s); the reduction and the backprojection spline over the shift grid, which
    # GSL's gsl_interp_init rejects unless strictly increasing
    phi, th, w = np.zeros(4), np.zeros(3), np.ones(3)
    s = np.linspace(-1, 1, 9)
    rng = np.random.default_rng(0)
    for bad in (s[::-1], rng.permutation(s), np.r_[s[:4], s[3:]]):
        for prec in ("f64", "f32", "bf16"):
            assert wrt.job_bytes("radon", prec, n=8, phi=phi, theta=th, shift=bad) > 0
        assert wrt.job_bytes("radon_T", "f32", n=8, phi=phi, theta=th, shift=bad) > 0
        assert wrt.job_bytes("ray", "f32", n=8, phi=phi, z=s, shift=bad) > 0
        for prec in ("f64", "f32"):
            assert wrt.job_bytes("adjoint", prec, n=8, phi=phi, theta=th, theta_w=w,
                                 shift=bad) == 0
        assert wrt.job_bytes("reduction", "f64", phi=phi, theta=th, shift=bad, nz=8) == 0
    assert wrt.job_bytes("adjoint", "f64", n=8, phi=phi, theta=th, theta_w=w, shift=s) > 0
    assert wrt.job_bytes("reduction", "f64", phi=phi, theta=th, shift=s, nz=8) > 0


def test_init_without_device_reports_nodev():
    h = C.c_void_p()
    rc = wrt.lib().wrt_init(C.byref(h), 0)
    if rc == wrt.WRT_OK:  # a GPU is present: this is a CPU-only check
        wrt.lib().wrt_free(h)
        return
    assert rc == wrt.WRT_E_NODEV
    assert "device" in wrt.strerror(rc)


def test_fn_examples_are_the_reference_example_functions():
    from golden_inputs import EXAMPLES, example_terms

    for name in EXAMPLES:
        got = wrt.fn_example(name)
        want = example_terms(name)
        assert len(got) == len(want)
        for (k, c, p), (k2, c2, p2) in zip(got, want):
            assert k == k2 and tuple(c) == tuple(c2)
            assert tuple(p[:len(p2)]) == tuple(p2) and not any(p[len(p2):])
    try:
        wrt.fn_example("no_such_function")
        assert False
    except wrt.WrtError as e:
        assert e.code == wrt.WRT_E_ARG


def test_fn_phantom_terms_are_the_phantom_ellipsoids():
    P = wrt.phantom_params(10, seed=4)
    T = wrt.fn_phantom(10, seed=4)
    for e in range(10):
        kind, c, p = T[e]
        assert kind == wrt.FN["ellipsoid"]
        assert np.array_equal(c, P[e, 0:3]) and p[0] == P[e, 15]
        assert np.array_equal(p[1:4], P[e, 3:6]) and np.array_equal(p[4:13], P[e, 6:15])
    # the same evaluation as the voxelised phantom at its cell centres
    n = 24
    vol = wrt.phantom(n, True, 10, 4, 0.0)
    xs = -1.0 + (2.0 * np.arange(n) + 1.0) / n
    rng = np.random.default_rng(0)
    for _ in range(300):
        i, j, k = rng.integers(n, size=3)
        want = vol[i, j, k]
        got = orc.fn_eval(T, xs[i], xs[j], xs[k]) if xs[i] ** 2 + xs[j] ** 2 + xs[k] ** 2 <= 1 else 0.0
        assert got == want


def test_analytic_zslices_match_oracle():
    for g in (6, 17, 129):
        assert np.array_equal(wrt.analytic_zslice_grid(g), orc.analytic_zslices(g))
