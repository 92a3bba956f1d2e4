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


# ------------------------------------------------------------------ analytic programs ---
def example_terms(name):
    """The reference's example test functions as term tables (kind, centre, parameters):
    radon_analytic/examples-test-functions/exponential_bump.c:5-16 (kind 0: r2max, eps) and
    sinusoidal_bump.c:22-35 (kind 1: amplitude, rin, rout, three bumps summed in order)."""
    if name == "exponential_bump":
        return [(0, (0.2, 0.0, 0.0), (0.5, 1e-2))]
    if name == "sinusoidal_bump":
        return [(1, (0.0, 0.2, 0.0), (1.0, 0.4, 0.6)),
                (1, (-0.5, 0.0, 0.0), (0.6, 0.3, 0.45)),
                (1, (0.0, 0.2, 0.5), (0.3, 0.4, 0.6))]
    raise ValueError(name)


EXAMPLES = ("exponential_bump", "sinusoidal_bump")


def analytic_plane_cases():
    """name -> (ngrid, phi[], theta[], shift[]) for the analytic plane_integral"""
    rng = np.random.default_rng(21)
    cases = {"g17": (17,) + _grid(8, 4, 9)}
    p, t, s = _grid(16, 8, 33)
    idx = rng.choice(len(p), 40, replace=False)
    m = 24
    cases["g33"] = (33, np.concatenate([p[idx], rng.uniform(0, 2 * math.pi, m)]),
                    np.concatenate([t[idx], np.arccos(rng.uniform(-1, 1, m))]),
                    np.concatenate([s[idx], rng.uniform(-1, 1, m)]))
    return cases


def analytic_zslices(ngrid):
    """z_i = (-1.0) + i*2.0/(ngrid-1), as ray_analytic/ray_analytic.c:255 forms it"""
    return np.array([(-1.0) + i * 2.0 / (ngrid - 1) for i in range(ngrid)])


def analytic_ray_cases():
    """name -> (ngrid, z[], shift[], phi[]) for the analytic ray_integral"""
    rng = np.random.default_rng(22)
    z = analytic_zslices(17)
    s = -1.0 + np.arange(9) * (2.0 / 8)
    ph = (2 * math.pi / 8) * np.arange(8)
    Z, S, PH = np.meshgrid(z, s, ph, indexing="ij")
    cases = {"g17": (17, Z.ravel(), S.ravel(), PH.ravel())}
    m = 64
    cases["g64r"] = (64, rng.uniform(-1, 1, m), rng.uniform(-1, 1, m), rng.uniform(0, 2 * math.pi, m))
    return cases


def analytic_points(name):
    """points for test_function: random, near the centres and on the terms' radii"""
    rng = np.random.default_rng(23)
    pts = [tuple(rng.uniform(-1.2, 1.2, 3)) for _ in range(300)]
    for kind, c, p in example_terms(name):
        radii = [math.sqrt(p[0] - p[1]), math.sqrt(p[0])] if kind == 0 else [p[1], p[2]]
        for r in radii + [0.0, 0.1]:
            for _ in range(6):
                d = rng.standard_normal(3)
                d /= np.linalg.norm(d)
                pts.append(tuple(float(c[k] + r * d[k]) for k in range(3)))
        pts.append(tuple(float(v) for v in c))
    return [tuple(float(v) for v in q) for q in pts]


# the off-centre Gaussian of the inversion tests: below 1e-5 of its peak outside the unit ball,
# so that the plane integrals on s in [-1, 1] hold all of it
GAUSS_C = (0.2, -0.12, 0.15)
GAUSS_SIGMA = 0.2


def fourier_case(nphi, ntheta, nshift, c=GAUSS_C, sigma=GAUSS_SIGMA):
    """the grids of rtft3d.m (uniform phi on [0, 2 pi), theta = arccos of the Gauss-Legendre
    nodes with their weights, uniform shifts on [-1, 1]) and radon[s][phi][theta] of
    f(x) = exp(-|x - c|^2 / sigma^2), whose plane integrals are pi sigma^2 exp(-(s - c.u)^2 /
    sigma^2)"""
    phi = (2 * math.pi / nphi) * np.arange(nphi)
    x, w = np.polynomial.legendre.leggauss(ntheta)
    theta = np.arccos(x)
    s = -1.0 + np.arange(nshift) * (2.0 / (nshift - 1))
    P, T = np.meshgrid(phi, theta, indexing="ij")
    cu = (np.sin(T) * np.cos(P) * c[0] + np.sin(T) * np.sin(P) * c[1] + np.cos(T) * c[2])
    radon = math.pi * sigma ** 2 * np.exp(-((s[:, None, None] - cu[None]) ** 2) / sigma ** 2)
    return radon, phi, theta, w


def gauss_on_grid(ngrid, c=GAUSS_C, sigma=GAUSS_SIGMA):
    """f at ((ix, iy, iz) - ngrid/2) * 2/ngrid, array [ix][iy][iz]"""
    x = (np.arange(ngrid) - ngrid // 2) * (2.0 / ngrid)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    return np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / sigma ** 2)


def fourier_case2d(nphi, nshift, c=GAUSS_C[:2], sigma=GAUSS_SIGMA):
    """rtft2d.m's grids (uniform phi on [0, 2 pi), uniform shifts on [-1, 1]) and radon[s][phi]
    of f(x) = exp(-|x - c|^2 / sigma^2) in 2D: sqrt(pi) sigma exp(-(s - c.u)^2 / sigma^2)"""
    phi = (2 * math.pi / nphi) * np.arange(nphi)
    s = -1.0 + np.arange(nshift) * (2.0 / (nshift - 1))
    cu = np.cos(phi) * c[0] + np.sin(phi) * c[1]
    radon = math.sqrt(math.pi) * sigma * np.exp(-((s[:, None] - cu[None]) ** 2) / sigma ** 2)
    return radon, phi


def gauss2d_on_grid(ngrid, c=GAUSS_C[:2], sigma=GAUSS_SIGMA):
    x = (np.arange(ngrid) - ngrid // 2) * (2.0 / ngrid)
    X, Y = np.meshgrid(x, x, indexing="ij")
    return np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2) / sigma ** 2)


def gauss_rays(z, s, phi, c=GAUSS_C, sigma=GAUSS_SIGMA):
    """rays[z][s][phi] of f(x) = exp(-|x - c|^2 / sigma^2) along the lines
    (s cos phi - t sin phi, s sin phi + t cos phi, z): sigma sqrt(pi) exp(-d^2 / sigma^2), d the
    distance from c to the line"""
    z, s, phi = (np.asarray(a, dtype=np.float64) for a in (z, s, phi))
    cs = c[0] * np.cos(phi) + c[1] * np.sin(phi)
    d2 = (z[:, None, None] - c[2]) ** 2 + (s[None, :, None] - cs[None, None, :]) ** 2
    return sigma * math.sqrt(math.pi) * np.exp(-d2 / sigma ** 2)


def gauss_planes(p, phi, theta, c=GAUSS_C, sigma=GAUSS_SIGMA):
    """R[p][phi][theta] of the same f: pi sigma^2 exp(-(p - c.u)^2 / sigma^2),
    u = (sin theta cos phi, sin theta sin phi, cos theta)"""
    p, phi, theta = (np.asarray(a, dtype=np.float64) for a in (p, phi, theta))
    cu = (c[0] * np.sin(theta)[None, :] * np.cos(phi)[:, None]
          + c[1] * np.sin(theta)[None, :] * np.sin(phi)[:, None] + c[2] * np.cos(theta)[None, :])
    return math.pi * sigma ** 2 * np.exp(-(p[:, None, None] - cu[None]) ** 2 / sigma ** 2)
