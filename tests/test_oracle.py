"""Pins the CPU oracle (oracle/wrt_oracle.c) before it is trusted as the checker.

- bit-exact against the reference's own GSL-free functions (golden fixtures made from
  radon_grid/src/{integral,interp}.c and ray_grid/src/{integral,interp}.c);
- Gauss-Legendre rule against 60-digit mpmath (correctly rounded) and the README values;
- natural spline against scipy; the quadratic against exact polynomials;
- closed forms: ellipsoid plane/line integrals, mass conservation, constant and linear
  data through the adjoint, a Gaussian through ray -> Radon reduction;
- the dot-product test for the transposes.
"""
import math
import os

import numpy as np
import pytest

import orc
from golden_inputs import volume_formula

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _npz(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


# --------------------------------------------------------------- reference goldens ---
def test_plane_integral_matches_reference_bits():
    f = _npz("ref_radon_grid.npz")
    keys = [k for k in f.files if k.startswith("radon_") and not k.endswith(("_in", "_n"))]
    assert len(keys) >= 4
    for k in keys:
        n = int(f[k + "_n"])
        phi, theta, shift = f[k + "_in"]
        vol = volume_formula(n)
        got = np.array([orc.plane_integral(vol, p, t, s) for p, t, s in zip(phi, theta, shift)])
        assert np.array_equal(got, f[k]), k


def test_zero_interp_matches_reference_bits():
    f = _npz("ref_radon_grid.npz")
    for k in [k for k in f.files if k.startswith("zero_") and not k.endswith(("_in", "_n"))]:
        n = int(f[k + "_n"])
        vol = volume_formula(n)
        got = np.array([orc.zero_interp(vol, *p) for p in f[k + "_in"]])
        assert np.array_equal(got, f[k]), k


def test_ray_transform_matches_reference_bits():
    f = _npz("ref_ray_grid.npz")
    keys = [k for k in f.files if k.startswith("ray_") and not k.endswith(("_in", "_n"))]
    assert "ray_n128" in keys
    for k in keys:
        n = int(f[k + "_n"])
        z, s, phi = f[k + "_in"]
        vol = volume_formula(n)
        got = orc.rays(vol, z, s, phi)
        assert np.array_equal(got, f[k]), k


def test_trilinear_matches_reference_bits():
    f = _npz("ref_ray_grid.npz")
    for k in [k for k in f.files if k.startswith("tri_") and not k.endswith(("_in", "_n"))]:
        n = int(f[k + "_n"])
        vol = volume_formula(n)
        got = np.array([orc.trilinear(vol, *p) for p in f[k + "_in"]])
        assert np.array_equal(got, f[k]), k


def test_trilinear_index_quirk_is_reproduced():
    # a quotient that rounds up to an integer k reads cell k with fraction ~1, i.e. ~node k+1
    n = 9
    vol = volume_formula(n)
    h = 2.0 / (n - 1)
    hits = 0
    for k in range(1, n - 1):
        x = np.nextafter(k * h - 1.0, -2.0)
        a = x + 1.0
        if int(a / h) == k and math.fmod(a, h) / h > 0.5:
            hits += 1
            v = orc.trilinear(vol, x, -1.0, -1.0)
            assert abs(v - vol[k + 1, 0, 0]) <= 1e-12 * max(1.0, abs(vol[k + 1, 0, 0]))
    assert hits >= 0


# ---------------------------------------------------------------------- grids / GL ---
def test_gl_nodes_correctly_rounded():
    f = _npz("gl.npz")
    for k in [k for k in f.files if k.startswith("x")]:
        n = int(k[1:])
        x, w = orc.gl_nodes(n)
        assert np.array_equal(x, f[k]), n
        assert np.array_equal(w, f["w%d" % n]), n
        assert np.array_equal(x, -x[::-1])
        assert abs(w.sum() - 2.0) < 1e-14
        assert np.all(np.diff(x) > 0)


def test_gl_exactness_on_polynomials():
    x, w = orc.gl_nodes(12)
    for k in range(0, 24):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(w * x ** k) - exact) < 1e-14


def test_readme_grid_values():
    # radon_analytic/README.md:76-78 (ntheta = 128) and ray_grid/README.md:91-93 (nphi = 256)
    th, _ = orc.theta_grid(128)
    assert ["%f" % t for t in th[:3]] == ["3.122878", "3.098635", "3.074249"]
    phi = orc.phi_grid(256)
    assert ["%f" % p for p in phi[227:230]] == ["5.571418", "5.595962", "5.620506"]
    s = orc.shift_grid(129)
    assert s[0] == -1.0 and s[-1] == 1.0 and s[64] == 0.0


# ------------------------------------------------------------ spline / polynomial ---
def test_natural_spline_vs_scipy():
    f = _npz("spline.npz")
    for k in [k for k in f.files if k.startswith("x")]:
        ns = k[1:]
        x, y, q, v = f["x" + ns], f["y" + ns], f["q" + ns], f["v" + ns]
        c = orc.cspline(x, y)
        assert c[0] == 0.0 and c[-1] == 0.0
        got = np.array([orc.cspline_eval(x, y, c, t) for t in q])
        assert np.max(np.abs(got - v)) <= 1e-13 * max(1.0, np.max(np.abs(v))), ns


def test_spline_reproduces_linear_data_and_nodes():
    x = -1.0 + np.arange(17) * (2.0 / 16)
    y = 0.25 - 1.5 * x
    c = orc.cspline(x, y)
    assert np.max(np.abs(c)) < 1e-14
    for t in np.linspace(-1, 1, 51):
        assert abs(orc.cspline_eval(x, y, c, t) - (0.25 - 1.5 * t)) < 1e-14
    y = np.cos(3 * x)
    c = orc.cspline(x, y)
    for i, t in enumerate(x):
        assert abs(orc.cspline_eval(x, y, c, t) - y[i]) < 1e-15


def test_poly3_exact_on_quadratics():
    rng = np.random.default_rng(3)
    for _ in range(50):
        a, b, cc = rng.standard_normal(3)
        xa = np.sort(rng.uniform(-1, 1, 3))
        ya = a + b * xa + cc * xa ** 2
        for t in rng.uniform(-1, 1, 5):
            assert abs(orc.poly3(xa, ya, t) - (a + b * t + cc * t * t)) < 1e-11


# -------------------------------------------------------------------- closed forms ---
def _one_ellipsoid(n, cell):
    import wrt

    p = wrt.phantom_params(1, seed=5)[0]
    vol = wrt.phantom(n, cell_centred=cell, nell=1, seed=5)
    return p, vol


def test_radon_ellipsoid_closed_form_and_mass():
    n = 48
    p, vol = _one_ellipsoid(n, True)
    c, ax, R, rho = p[0:3], p[3:6], p[6:15], p[15]
    rng = np.random.default_rng(1)
    for _ in range(12):
        phi, th, s = rng.uniform(0, 2 * math.pi), math.acos(rng.uniform(-1, 1)), rng.uniform(-0.6, 0.6)
        nr, _, _ = orc.plane_basis(phi, th)
        got = orc.plane_integral(vol, phi, th, s)
        want = orc.ellipsoid_plane(c, ax, R, rho, nr, s)
        assert abs(got - want) <= 0.06 * abs(rho) * math.pi * max(ax) ** 2 + 1e-12
    # mass: sum_s R(s) ds = sum f h^3 (exact in the continuum, discretisation here)
    ns = 97
    shift = orc.shift_grid(ns)
    R1 = orc.radon_grid(vol, np.array([0.7]), np.array([1.1]), shift)[:, 0, 0]
    mass = R1.sum() * (2.0 / (ns - 1))
    assert abs(mass - vol.sum() * (2.0 / n) ** 3) < 0.02 * abs(vol.sum() * (2.0 / n) ** 3)


def test_ray_ellipsoid_closed_form():
    n = 49
    p, vol = _one_ellipsoid(n, False)
    c, ax, R, rho = p[0:3], p[3:6], p[6:15], p[15]
    rng = np.random.default_rng(2)
    for _ in range(16):
        z, s, phi = rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(0, 2 * math.pi)
        got = orc.ray_transform(vol, z, s, phi)
        p0 = np.array([s * math.cos(phi), s * math.sin(phi), z])
        d = np.array([math.cos(phi + math.pi / 2), math.sin(phi + math.pi / 2), 0.0])
        want = orc.ellipsoid_line(c, ax, R, rho, p0, d)
        assert abs(got - want) <= 0.08 * abs(rho) * 2 * max(ax) + 1e-12


def test_adjoint_constant_and_linear_data():
    nphi, nth, ns, n = 12, 6, 17, 6
    phi = orc.phi_grid(nphi)
    th, w = orc.theta_grid(nth)
    s = orc.shift_grid(ns)
    const = np.full((ns, nphi, nth), 0.75)
    adj = orc.radon_adjoint(const, phi, th, w, s, n)
    xs = -1.0 + 1.0 / n + np.arange(n) * (2.0 / n)
    Z, X, Y = np.meshgrid(xs, xs, xs, indexing="ij")
    inside = (X ** 2 + Y ** 2 + Z ** 2) <= 1.0
    assert np.max(np.abs(adj[inside] - 4 * math.pi * 0.75)) < 1e-12
    lin = np.broadcast_to(s[:, None, None], (ns, nphi, nth)).copy()
    adj = orc.radon_adjoint(lin, phi, th, w, s, n)
    assert np.max(np.abs(adj[inside])) < 1e-12


def test_reduction_gaussian_closed_form():
    # f = exp(-|x|^2/sg^2): P(z,s) = sg sqrt(pi) exp(-(z^2+s^2)/sg^2), R(p) = pi sg^2 exp(-p^2/sg^2)
    sg = 0.3
    nz, ns, nphi, nth = 65, 65, 4, 6
    z = orc.zslice_grid(nz)
    s = orc.shift_grid(ns)
    phi = orc.phi_grid(nphi)
    th, _ = orc.theta_grid(nth)
    rays = sg * math.sqrt(math.pi) * np.exp(-(z[:, None, None] ** 2 + s[None, :, None] ** 2) / sg ** 2)
    rays = np.broadcast_to(rays, (nz, ns, nphi)).copy()
    R = orc.radon_reduction(rays, s, phi, th)
    want = math.pi * sg ** 2 * np.exp(-s ** 2 / sg ** 2)
    err = np.max(np.abs(R - want[:, None, None]))
    assert err < 2e-3 * want.max(), err


def test_ray_symmetry_phi_plus_pi():
    n = 17
    vol = volume_formula(n)
    rng = np.random.default_rng(4)
    for _ in range(20):
        z, s, phi = rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0, math.pi)
        a = orc.ray_transform(vol, z, s, phi + math.pi)
        b = orc.ray_transform(vol, z, -s, phi)
        assert abs(a - b) <= 1e-10 * max(1.0, abs(b))


# ---------------------------------------------------------------------- transposes ---
def test_transposes_dot_product():
    rng = np.random.default_rng(9)
    n = 6
    phi, s = orc.phi_grid(6), orc.shift_grid(7)
    th, _ = orc.theta_grid(3)
    f = rng.standard_normal((n, n, n))
    g = rng.standard_normal((len(s), len(phi), len(th)))
    Af = orc.radon_grid(f, phi, th, s)
    ATg = orc.radon_grid_T(g, n, phi, th, s)
    lhs, rhs = np.sum(Af * g), np.sum(f * ATg)
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(Af) * np.linalg.norm(g)
    n = 7
    z = orc.zslice_grid(5)
    s = orc.shift_grid(9)
    phi = orc.phi_grid(8)
    f = rng.standard_normal((n, n, n))
    g = rng.standard_normal((len(z), len(s), len(phi)))
    Af = orc.ray_grid(f, z, s, phi)
    ATg = orc.ray_grid_T(g, n, z, s, phi)
    lhs, rhs = np.sum(Af * g), np.sum(f * ATg)
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(Af) * np.linalg.norm(g)


@pytest.mark.parametrize("planted", ["sign", "transpose", "delta"])
def test_checks_catch_planted_mistakes(planted):
    # a mistake planted in the oracle's answer must fail the closed-form check on some plane
    n = 48
    p, vol = _one_ellipsoid(n, True)
    c, ax, R, rho = p[0:3], p[3:6], p[6:15], p[15]
    tol = 0.06 * abs(rho) * math.pi * max(ax) ** 2
    volT = np.ascontiguousarray(vol.transpose(2, 1, 0))
    rng = np.random.default_rng(1)
    worst = 0.0
    for _ in range(12):
        phi, th = rng.uniform(0, 2 * math.pi), math.acos(rng.uniform(-1, 1))
        nr, _, _ = orc.plane_basis(phi, th)
        s = float(np.dot(c, nr)) + rng.uniform(-0.05, 0.05)
        if planted == "sign":
            got = -orc.plane_integral(vol, phi, th, s)
        elif planted == "transpose":
            got = orc.plane_integral(volT, phi, th, s)
        else:  # one factor delta dropped from the term (value*delta)*delta
            got = orc.plane_integral(vol, phi, th, s) * n
        want = orc.ellipsoid_plane(c, ax, R, rho, nr, s)
        worst = max(worst, abs(got - want))
    assert worst > tol


def test_corrected_reciprocal_division_is_ieee():
    # the fast plane kernel replaces (x+1)/(2/n) by a corrected reciprocal when n is not a
    # power of two; it must decide every cell index exactly as the IEEE quotient does
    import ctypes as C
    f = orc.L().orc_div_rn_mismatches
    f.restype = C.c_longlong
    f.argtypes = [C.c_int, C.c_longlong, C.c_ulonglong]
    for n in list(range(3, 1025)):
        if n & (n - 1):
            assert f(n, 2000, n) == 0, n


# ----------------------------------------------------------------- analytic programs ---
from golden_inputs import EXAMPLES, analytic_zslices, example_terms  # noqa: E402


def test_analytic_test_functions_match_reference_bits():
    f = _npz("ref_analytic.npz")
    for name in EXAMPLES:
        terms = example_terms(name)
        pts = f[name + "_fn_in"]
        got = np.array([orc.fn_eval(terms, *p) for p in pts])
        assert np.array_equal(got, f[name + "_fn"]), name
        assert np.count_nonzero(got) > 20


def test_analytic_plane_integral_matches_reference_bits():
    f = _npz("ref_analytic.npz")
    for name in EXAMPLES:
        terms = example_terms(name)
        keys = [k for k in f.files if k.startswith(name + "_plane_") and not k.endswith(("_in", "_g"))]
        assert len(keys) == 2
        for k in keys:
            g = int(f[k + "_g"])
            phi, theta, shift = f[k + "_in"]
            got = np.array([orc.analytic_plane(terms, g, p, t, s) for p, t, s in zip(phi, theta, shift)])
            assert np.array_equal(got, f[k]), k
            assert np.count_nonzero(got) > len(got) // 4


def test_analytic_ray_integral_matches_reference_bits():
    f = _npz("ref_analytic.npz")
    for name in EXAMPLES:
        terms = example_terms(name)
        keys = [k for k in f.files if k.startswith(name + "_ray_") and not k.endswith(("_in", "_g"))]
        assert len(keys) == 2
        for k in keys:
            g = int(f[k + "_g"])
            z, s, phi = f[k + "_in"]
            got = np.array([orc.analytic_ray(terms, a, b, c, g) for a, b, c in zip(z, s, phi)])
            assert np.array_equal(got, f[k]), k


def test_analytic_zslices_follow_their_own_rounding():
    # ray_analytic forms z = -1 + (i*2.0)/(ngrid-1), ray_grid z = -1 + i*(2.0/(n-1)): the two
    # grids differ in the last bit for some i, and each program keeps its own
    for g in (6, 17, 33, 100, 129, 257):
        assert np.array_equal(orc.analytic_zslices(g), analytic_zslices(g))
    diffs = sum(np.count_nonzero(orc.analytic_zslices(g) != orc.zslice_grid(g)) for g in (6, 7, 11, 100))
    assert diffs > 0


def _exp_bump_plane(r2max, eps, d):
    """pi * int_eps^A exp(-1/t) dt, A = r2max - d^2: the bump's plane integral at distance d
    (t e^(-1/t) - E1(1/t) is an antiderivative of e^(-1/t))"""
    from scipy.special import exp1

    a = r2max - d * d
    if a <= eps:
        return 0.0
    F = lambda t: t * math.exp(-1.0 / t) - exp1(1.0 / t)  # noqa: E731
    return math.pi * (F(a) - F(eps))


def _exp_bump_line(r2max, eps, d):
    from scipy.integrate import quad

    a = r2max - d * d
    if a <= eps:
        return 0.0
    L = math.sqrt(a - eps)
    return quad(lambda u: math.exp(-1.0 / (a - u * u)), -L, L, epsabs=1e-16, epsrel=1e-13,
                limit=200)[0]


def test_analytic_exponential_bump_closed_forms():
    # The bump is smooth and compactly supported, so the lattice sums converge faster than
    # any power of delta: at ngrid = 65 they equal the exact integrals to ~1e-11 (plane,
    # via the exponential integral E1) and ~1e-11 (line, via quadrature).
    terms = example_terms("exponential_bump")
    c = np.array(terms[0][1])
    r2max, eps = terms[0][2]
    rng = np.random.default_rng(5)
    peak_p, peak_l = _exp_bump_plane(r2max, eps, 0.0), _exp_bump_line(r2max, eps, 0.0)
    for _ in range(8):
        phi, th, s = rng.uniform(0, 2 * math.pi), math.acos(rng.uniform(-1, 1)), rng.uniform(-0.6, 0.6)
        n = np.array([math.sin(th) * math.cos(phi), math.sin(th) * math.sin(phi), math.cos(th)])
        got = orc.analytic_plane(terms, 65, phi, th, s)
        assert abs(got - _exp_bump_plane(r2max, eps, abs(c @ n - s))) <= 1e-10 * peak_p
        z, s, phi = rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), rng.uniform(0, 2 * math.pi)
        p0 = np.array([s * math.cos(phi), s * math.sin(phi), z])
        dv = np.array([math.cos(phi + math.pi / 2), math.sin(phi + math.pi / 2), 0.0])
        w = c - p0
        d = math.sqrt(max(w @ w - (w @ dv) ** 2, 0.0))
        got = orc.analytic_ray(terms, z, s, phi, 65)
        assert abs(got - _exp_bump_line(r2max, eps, d)) <= 1e-10 * peak_l


def test_analytic_ellipsoid_term_closed_form():
    # the indicator has a jump, so the lattice sum converges like delta: 1% at ngrid = 129
    rot = [0.36, -0.48, 0.8, 0.8, 0.6, 0.0, -0.48, 0.64, 0.6]
    terms = [(2, (0.1, 0.05, -0.1), [0.7, 0.5, 0.3, 0.2] + rot)]
    rng = np.random.default_rng(6)
    scale = 0.7 * math.pi * 0.5 * 0.3
    for _ in range(6):
        phi, th, s = rng.uniform(0, 2 * math.pi), math.acos(rng.uniform(-1, 1)), rng.uniform(-0.3, 0.3)
        n = np.array([math.sin(th) * math.cos(phi), math.sin(th) * math.sin(phi), math.cos(th)])
        want = orc.ellipsoid_plane(np.array(terms[0][1]), np.array([0.5, 0.3, 0.2]),
                                   np.array(rot), 0.7, n, s)
        assert abs(orc.analytic_plane(terms, 129, phi, th, s) - want) <= 1e-2 * scale


# ------------------------------------------------------------- spectral Laplacians ---
def _spectral():
    import importlib.util

    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                        "spectral.py")
    spec = importlib.util.spec_from_file_location("spectral", path)
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


@pytest.mark.parametrize("n", [8, 9, 16])
def test_spectral_laplacian_on_trigonometric_polynomials(n):
    # cos(2 pi (a.i)/n) is an eigenvector: Laplacian = -(2 pi / period)^2 |a|^2 times it, for
    # every frequency up to and including the Nyquist one (kept as +n/2 by laplace3dfft.m)
    sp = _spectral()
    period = 2.5
    i = np.arange(n)
    I, J, K = np.meshgrid(i, i, i, indexing="ij")
    for a in [(1, 0, 0), (0, 2, 1), (3, 1, 2), (n // 2, 0, 1), (n // 2, n // 2, n // 2)]:
        u = np.cos(2 * np.pi * (a[0] * I + a[1] * J + a[2] * K) / n)
        want = -(2 * np.pi / period) ** 2 * (a[0] ** 2 + a[1] ** 2 + a[2] ** 2) * u
        assert np.max(np.abs(sp.laplace_fft(u, period) - want)) <= 1e-11 * max(1.0, np.max(np.abs(want)))
    I2, J2 = np.meshgrid(i, i, indexing="ij")
    u = np.sin(2 * np.pi * (2 * I2 + 3 * J2) / n)
    want = -(2 * np.pi / period) ** 2 * 13 * u
    assert np.max(np.abs(sp.laplace_fft(u, period) - want)) <= 1e-11 * np.max(np.abs(want))
    f = sp.frequencies(n)
    assert f[0] == 0 and f[n // 2] == n // 2 and (n % 2 == 0 or f[n // 2 + 1] == -(n // 2))


def test_spectral_laplacian_is_axis_symmetric_and_smooth_limit():
    sp = _spectral()
    rng = np.random.default_rng(3)
    u = rng.standard_normal((12, 12, 12))
    assert np.allclose(sp.laplace_fft(u.transpose(2, 0, 1), 2.0),
                       sp.laplace_fft(u, 2.0).transpose(2, 0, 1), atol=1e-12)
    # a narrow Gaussian on a wide period (e^-36 at the edge): the spectral Laplacian is the
    # analytic one, (4 r^2 - 6) e^(-r^2)
    n, period = 128, 12.0
    x = (np.arange(n) - n // 2) * (period / n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    r2 = X ** 2 + Y ** 2 + Z ** 2
    u = np.exp(-r2)
    want = (4 * r2 - 6) * u
    assert np.max(np.abs(sp.laplace_fft(u, period) - want)) <= 1e-10


# ------------------------------------------------------ Fourier-slice inversion ---
def test_rtft_is_the_padded_dft_with_shift_correction_and_weights():
    # brute force on a tiny case: coef[d][m] = jw_d rho_m^2 exp(2 pi i rsupp k dfreq)
    # sum_j radon[j][d] exp(-2 pi i j k / ntotal), k = (m + ntotal/2) mod ntotal
    fo = orc.fourier_module()
    rng = np.random.default_rng(5)
    nshift, nphi, ntheta, rsupp, padding = 9, 3, 2, 1.0, 2.0
    radon = rng.standard_normal((nshift, nphi, ntheta))
    phi = np.array([0.1, 2.0, 4.0])
    theta = np.array([0.7, 2.1])
    tw = np.array([0.4, 1.6])
    rho, dirs, coef = fo.rtft3d(radon, phi, theta, tw, rsupp, padding)
    dshift = 2 * rsupp / (nshift - 1)
    n = int(padding * math.floor(1 / dshift))
    dfreq = 1 / (n * dshift)
    assert len(rho) == n == 8 and rho[0] == -4 * dfreq and rho[n // 2] == 0
    for ip in range(nphi):
        for it in range(ntheta):
            d = ip * ntheta + it
            u = [math.sin(theta[it]) * math.cos(phi[ip]), math.sin(theta[it]) * math.sin(phi[ip]),
                 math.cos(theta[it])]
            assert np.allclose(dirs[d], u, atol=1e-15)
            for m in range(n):
                k = (m + n // 2) % n
                x = sum(radon[j, ip, it] * np.exp(-2j * math.pi * j * k / n)
                        for j in range(min(nshift, n)))
                want = (0.5 * dshift * dfreq * (2 * math.pi / nphi) * tw[it] * rho[m] ** 2
                        * np.exp(2j * math.pi * rsupp * k * dfreq) * x)
                assert abs(coef[d, m] - want) <= 1e-13 * (1 + abs(want))


def test_fourier_adjoint_numpy_and_c_agree():
    fo = orc.fourier_module()
    rng = np.random.default_rng(6)
    rho = (np.arange(10) - 5) * 0.37
    dirs = rng.standard_normal((7, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    coef = rng.standard_normal((7, 10)) + 1j * rng.standard_normal((7, 10))
    a = fo.adjoint_sum(rho, dirs, coef, 6)
    b = orc.fourier_adjoint(rho, dirs, coef, 6)
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a))
    # a single node at rho = 0.37 along x with coefficient 1: f(p) = cos(2 pi 0.37 p_x)
    c1 = np.zeros((1, 10), complex)
    c1[0, 6] = 1.0
    x = (np.arange(6) - 3) * (2.0 / 6)
    want = np.broadcast_to(np.cos(2 * math.pi * 0.37 * x)[:, None, None], (6, 6, 6))
    assert np.allclose(orc.fourier_adjoint(rho, np.array([[1.0, 0, 0]]), c1, 6), want, atol=1e-14)


def test_chang3d_reconstructs_an_off_centre_gaussian():
    # closed-form plane integrals of exp(-|x - c|^2 / sigma^2) in, the function itself out (up
    # to the discretisation error); c off every axis pins the axis order [x][y][z], the sign of
    # the shift correction and the direction convention; w0 divides
    from golden_inputs import fourier_case, gauss_on_grid

    # (max error 3e-4 at this size, 1e-5 at 64 phi x 32 theta x 129 s -> 32^3)
    radon, phi, theta, w = fourier_case(48, 24, 65)
    got = orc.chang3d(radon, phi, theta, w, 24, w0=2.0)
    want = gauss_on_grid(24)
    assert np.max(np.abs(2.0 * got - want)) <= 5e-4
    assert orc.rel_l2(2.0 * got, want) <= 5e-4
    assert np.unravel_index(np.argmax(got), got.shape) == (12 + 2, 12 - 1, 12 + 2)
    # coarser sampling, larger error
    radon, phi, theta, w = fourier_case(32, 16, 65)
    got = orc.chang3d(radon, phi, theta, w, 16)
    assert 1e-3 <= np.max(np.abs(got - gauss_on_grid(16))) <= 0.01


def test_rtft2d_is_the_padded_dft_with_shift_correction_and_weights():
    fo = orc.fourier_module()
    rng = np.random.default_rng(7)
    nshift, nphi, rsupp, padding = 9, 3, 1.0, 2.0
    radon = rng.standard_normal((nshift, nphi))
    phi = np.array([0.3, 2.5, 5.0])
    rho, dirs, coef = fo.rtft2d(radon, phi, rsupp, padding)
    dshift = 2 * rsupp / (nshift - 1)
    n = int(padding * math.floor(1 / dshift))
    dfreq = 1 / (n * dshift)
    for ip in range(nphi):
        assert np.allclose(dirs[ip], [math.cos(phi[ip]), math.sin(phi[ip])], atol=1e-15)
        for m in range(n):
            k = (m + n // 2) % n
            x = sum(radon[j, ip] * np.exp(-2j * math.pi * j * k / n) for j in range(min(nshift, n)))
            want = (0.5 * dfreq * (2 * math.pi / nphi) * abs(rho[m]) * dshift
                    * np.exp(2j * math.pi * rsupp * k * dfreq) * x)
            assert abs(coef[ip, m] - want) <= 1e-13 * (1 + abs(want))


def test_chang2d_reconstructs_an_off_centre_gaussian():
    # the 2D weight |rho| has a kink at 0, so the rule over rho is second order in dfreq =
    # 1/(ntotal dshift), which the padding sets: the error falls about 4x per doubling of the
    # padding (a wrong weight or sign would not converge), and is 2.6e-4 at padding 16
    from golden_inputs import fourier_case2d, gauss2d_on_grid

    radon, phi = fourier_case2d(64, 129)
    want = gauss2d_on_grid(32)
    errs = [np.max(np.abs(orc.chang2d(radon, phi, 32, padding=p) - want)) for p in (4.0, 8.0, 16.0)]
    assert errs[2] <= 4e-4 and orc.rel_l2(orc.chang2d(radon, phi, 32, padding=16.0), want) <= 3e-3
    assert 3.0 <= errs[0] / errs[1] <= 6.0 and 3.0 <= errs[1] / errs[2] <= 6.0
    got = orc.chang2d(radon, phi, 32, w0=2.0, padding=16.0)
    assert np.unravel_index(np.argmax(got), got.shape) == (16 + 3, 16 - 2)
