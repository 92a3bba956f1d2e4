"""Writes the golden fixtures under tests/golden/.  Run in the build container only:

    make -C oracle ref && python tests/golden/make_golden.py

Sources of truth (the reference has no tests of its own):
  ref_radon_grid.npz  plane_integral / zero_interp from the reference's own
                      radon_grid/src/{integral,interp}.c, compiled by oracle/Makefile into
                      oracle/_ref (never copied into this repository)
  ref_ray_grid.npz    ray_transform / cube_trilinear_interp from ray_grid/src/{integral,interp}.c
  ref_analytic.npz    plane_integral / ray_integral / test_function of the analytic programs,
                      from radon_analytic/src/integral.c and ray_analytic/integral.c compiled
                      with each example test function (examples-test-functions/*.c)
  gl.npz              Gauss-Legendre nodes/weights, 50-digit mpmath, correctly rounded
  spline.npz          scipy CubicSpline(bc_type='natural') values
Volumes are integer-formula volumes, rebuilt exactly by volume_formula() in the tests, so
the fixtures hold only grids and outputs.
"""
import ctypes as C
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
from golden_inputs import volume_formula, radon_cases, ray_cases, point_cases  # noqa: E402

D = C.c_double
P = C.POINTER(D)


def jagged(v):
    n = v.shape[0]
    top = (C.POINTER(P) * n)()
    keep = []
    for i in range(n):
        mid = (P * n)()
        for j in range(n):
            mid[j] = v[i, j].ctypes.data_as(P)
        keep.append(mid)
        top[i] = C.cast(mid, C.POINTER(P))
    return top, keep


def ref_libs():
    rr = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_radon_grid.so"))
    ry = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_ray_grid.so"))
    for f in (rr.plane_integral, ry.ray_transform, rr.zero_interp, ry.cube_trilinear_interp):
        f.restype = D
        f.argtypes = [C.c_void_p, C.c_int, D, D, D]
    return rr, ry


def make_ref():
    rr, ry = ref_libs()
    out = {}
    for key, (n, phi, theta, shift) in radon_cases().items():
        v = volume_formula(n)
        j, keep = jagged(v)
        out["radon_" + key] = np.array(
            [rr.plane_integral(j, n, p, t, s) for p, t, s in zip(phi, theta, shift)])
        out["radon_" + key + "_in"] = np.stack([phi, theta, shift])
        out["radon_" + key + "_n"] = np.array(n)
    for key, (n, xyz) in point_cases("cell").items():
        v = volume_formula(n)
        j, keep = jagged(v)
        out["zero_" + key] = np.array([rr.zero_interp(j, n, *p) for p in xyz])
        out["zero_" + key + "_in"] = np.array(xyz)
        out["zero_" + key + "_n"] = np.array(n)
    np.savez_compressed(os.path.join(HERE, "ref_radon_grid.npz"), **out)
    out = {}
    for key, (n, z, s, phi) in ray_cases().items():
        v = volume_formula(n)
        j, keep = jagged(v)
        out["ray_" + key] = np.array([ry.ray_transform(j, n, a, b, c) for a, b, c in zip(z, s, phi)])
        out["ray_" + key + "_in"] = np.stack([z, s, phi])
        out["ray_" + key + "_n"] = np.array(n)
    for key, (n, xyz) in point_cases("node").items():
        v = volume_formula(n)
        j, keep = jagged(v)
        out["tri_" + key] = np.array([ry.cube_trilinear_interp(j, n, *p) for p in xyz])
        out["tri_" + key + "_in"] = np.array(xyz)
        out["tri_" + key + "_n"] = np.array(n)
    np.savez_compressed(os.path.join(HERE, "ref_ray_grid.npz"), **out)


def make_analytic():
    """plane_integral / ray_integral / test_function of the analytic programs, from the
    reference's radon_analytic/src/integral.c and ray_analytic/integral.c compiled with each
    of its example test functions (oracle/_ref/libref_{radon,ray}_analytic_<example>.so)."""
    from golden_inputs import EXAMPLES, analytic_plane_cases, analytic_points, analytic_ray_cases

    out = {}
    for name in EXAMPLES:
        rl = C.CDLL(os.path.join(ROOT, "oracle", "_ref", f"libref_radon_analytic_{name}.so"))
        yl = C.CDLL(os.path.join(ROOT, "oracle", "_ref", f"libref_ray_analytic_{name}.so"))
        rl.plane_integral.restype = D
        rl.plane_integral.argtypes = [C.c_void_p, C.c_int, D, D, D]
        yl.ray_integral.restype = D
        yl.ray_integral.argtypes = [C.c_void_p, D, D, D, C.c_int]
        rl.test_function.restype = D
        rl.test_function.argtypes = [D, D, D]
        fr = C.cast(rl.test_function, C.c_void_p)
        fy = C.cast(yl.test_function, C.c_void_p)
        for key, (g, phi, theta, shift) in analytic_plane_cases().items():
            out[f"{name}_plane_{key}"] = np.array(
                [rl.plane_integral(fr, g, p, t, s) for p, t, s in zip(phi, theta, shift)])
            out[f"{name}_plane_{key}_in"] = np.stack([phi, theta, shift])
            out[f"{name}_plane_{key}_g"] = np.array(g)
        for key, (g, z, s, phi) in analytic_ray_cases().items():
            out[f"{name}_ray_{key}"] = np.array(
                [yl.ray_integral(fy, a, b, c, g) for a, b, c in zip(z, s, phi)])
            out[f"{name}_ray_{key}_in"] = np.stack([z, s, phi])
            out[f"{name}_ray_{key}_g"] = np.array(g)
        pts = analytic_points(name)
        out[f"{name}_fn"] = np.array([rl.test_function(*q) for q in pts])
        out[f"{name}_fn_in"] = np.array(pts)
    np.savez_compressed(os.path.join(HERE, "ref_analytic.npz"), **out)


def make_gl():
    import mpmath as mp

    mp.mp.dps = 60
    out = {}
    for n in (1, 2, 3, 5, 8, 12, 16, 24, 31, 32, 40, 48, 64, 96, 128):
        xs, ws = [], []
        start, _ = np.polynomial.legendre.leggauss(n)
        for k in range(n):
            # Newton polish of numpy's node in 60-digit arithmetic
            x = mp.mpf(float(start[k]))
            for _ in range(8):
                pn, pm = mp.legendre(n, x), mp.legendre(n - 1, x)
                if x * x == 1:
                    break
                x = x - pn / (n * (x * pn - pm) / (x * x - 1))
            pm = mp.legendre(n - 1, x)
            w = 2 * (1 - x * x) / (n * pm) ** 2
            xs.append(float(x))
            ws.append(float(w))
        if n % 2 == 1:
            xs[n // 2] = 0.0
        out[f"x{n}"] = np.array(xs)
        out[f"w{n}"] = np.array(ws)
    np.savez_compressed(os.path.join(HERE, "gl.npz"), **out)


def make_spline():
    from scipy.interpolate import CubicSpline

    rng = np.random.default_rng(7)
    out = {}
    for ns in (3, 5, 9, 33, 129):
        x = -1.0 + np.arange(ns) * (2.0 / (ns - 1))
        y = rng.standard_normal(ns)
        q = np.concatenate([x, rng.uniform(-1, 1, 200), [1.0, -1.0]])
        cs = CubicSpline(x, y, bc_type="natural")
        out[f"x{ns}"] = x
        out[f"y{ns}"] = y
        out[f"q{ns}"] = q
        out[f"v{ns}"] = cs(q)
    np.savez_compressed(os.path.join(HERE, "spline.npz"), **out)


if __name__ == "__main__":
    make_ref()
    make_analytic()
    make_gl()
    make_spline()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
