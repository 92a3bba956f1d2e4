"""Inputs of the golden fixtures (shared by tests/golden/make_golden.py and the tests).

Volumes are integer formulas with exact binary-fraction values, so they are rebuilt
bit-identically anywhere; the angle/shift arrays are stored in the fixtures next to the
outputs (tests read them back rather than recomputing them).
"""
import math

import numpy as np


def volume_formula(n):
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    v = ((i * 7 + j * 13 + k * 29 + i * j * 3 + j * k * 5) % 97) - 48
    return np.ascontiguousarray(v.astype(np.float64) / 64.0)


def _gl_theta(nt):
    x, _ = np.polynomial.legendre.leggauss(nt)
    x = np.where(np.abs(x) < 1e-15, 0.0, x)
    return np.arccos(x)


def _grid(nphi, ntheta, nshift):
    phi = (2 * math.pi / nphi) * np.arange(nphi)
    theta = _gl_theta(ntheta)
    shift = -1.0 + np.arange(nshift) * (2.0 / (nshift - 1))
    P, T, S = np.meshgrid(phi, theta, shift, indexing="ij")
    return P.ravel(), T.ravel(), S.ravel()


def radon_cases():
    """name -> (n, phi[], theta[], shift[]) for plane_integral"""
    rng = np.random.default_rng(11)
    cases = {}
    cases["n4"] = (4,) + _grid(8, 4, 9)
    cases["n7"] = (7,) + _grid(6, 3, 5)
    p, t, s = _grid(4, 2, 5)
    m = 64
    cases["n16"] = (16, np.concatenate([p, rng.uniform(0, 2 * math.pi, m)]),
                    np.concatenate([t, np.arccos(rng.uniform(-1, 1, m))]),
                    np.concatenate([s, rng.uniform(-1, 1, m)]))
    p, t, s = _grid(16, 8, 33)
    idx = rng.choice(len(p), 48, replace=False)
    cases["n32"] = (32, p[idx], t[idx], s[idx])
    return cases


def ray_cases():
    """name -> (n, z[], shift[], phi[]) for ray_transform"""
    rng = np.random.default_rng(12)
    cases = {}

    def grid(n, nz, ns, nphi, span=2 * math.pi):
        z = -1.0 + np.arange(nz) * (2.0 / (nz - 1))
        s = -1.0 + np.arange(ns) * (2.0 / (ns - 1))
        ph = (span / nphi) * np.arange(nphi)
        Z, S, PH = np.meshgrid(z, s, ph, indexing="ij")
        return Z.ravel(), S.ravel(), PH.ravel()

    cases["n9"] = (9,) + grid(9, 9, 9, 8)
    cases["n17"] = (17,) + grid(5, 5, 17, 16)
    z, s, ph = grid(128, 128, 257, 180, math.pi)
    idx = np.concatenate([np.flatnonzero(ph == 0.0)[::97], rng.choice(len(z), 40, replace=False)])
    cases["n128"] = (128, z[idx], s[idx], ph[idx])
    m = 64
    cases["n64r"] = (64, rng.uniform(-1, 1, m), rng.uniform(-1, 1, m), rng.uniform(0, 2 * math.pi, m))
    return cases


def point_cases(kind):
    """name -> (n, [(x, y, z), ...]) for zero_interp ('cell') or cube_trilinear_interp ('node')"""
    rng = np.random.default_rng(13 if kind == "cell" else 14)
    cases = {}
    for n in ((8, 5) if kind == "cell" else (9, 10)):
        h = 2.0 / n if kind == "cell" else 2.0 / (n - 1)
        edges = [-1.0 + k * h for k in range(n + 1)] + [1.0, -1.0, 1.0 + 1e-12, -1.0 - 1e-12,
                                                       np.nextafter(1.0, 2.0), np.nextafter(-1.0, -2.0)]
        pts = []
        for _ in range(200):
            pts.append(tuple(rng.choice(edges, 3)))
        for _ in range(200):
            pts.append(tuple(rng.uniform(-1.05, 1.05, 3)))
        # quotients that round up to an integer (the trilinear index/fraction quirk)
        for k in range(1, n):
            a = k * h
            pts.append((np.nextafter(a - 1.0, -2.0), 0.1, -0.3))
            pts.append((0.2, np.nextafter(a - 1.0, -2.0), np.nextafter(a - 1.0, 2.0)))
        cases[f"n{n}"] = (n, [tuple(float(c) for c in p) for p in pts])
    return cases
