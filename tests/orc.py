"""ctypes wrapper of oracle/liboracle.so -- the CPU FP64 checker.

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
use it.  Layouts match the product's (see include/wrt.h).
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# ORACLE_LIB selects another build of the checker (the sanitizer build oracle/_san)
ORACLE = os.environ.get("ORACLE_LIB") or os.path.join(ROOT, "oracle", "liboracle.so")
REF_RADON = os.path.join(ROOT, "oracle", "_ref", "libref_radon_grid.so")
REF_RAY = os.path.join(ROOT, "oracle", "_ref", "libref_ray_grid.so")

_L = None
D = C.c_double
I = C.c_int
P = C.c_void_p


def L():
    global _L
    if _L is None:
        l = C.CDLL(ORACLE)
        l.orc_zero_interp.restype = D
        l.orc_zero_interp.argtypes = [P, I, D, D, D]
        l.orc_plane_integral.restype = D
        l.orc_plane_integral.argtypes = [P, I, D, D, D]
        l.orc_plane_incube.restype = C.c_longlong
        l.orc_plane_incube.argtypes = [I, D, D, D]
        l.orc_trilinear.restype = D
        l.orc_trilinear.argtypes = [P, I, D, D, D]
        l.orc_ray_transform.restype = D
        l.orc_ray_transform.argtypes = [P, I, D, D, D]
        l.orc_cspline_eval.restype = D
        l.orc_cspline_eval.argtypes = [P, P, P, I, D]
        l.orc_poly3_eval.restype = D
        l.orc_poly3_eval.argtypes = [P, P, D]
        l.orc_radon_reduction_one.restype = D
        l.orc_radon_reduction_one.argtypes = [P, I, P, I, P, I, P, I, I, I, I]
        l.orc_radon_reduction_rays_one.restype = D
        l.orc_radon_reduction_rays_one.argtypes = [P, I, P, I, I, I, D, D, D]
        l.orc_radon_adjoint_point.restype = D
        l.orc_radon_adjoint_point.argtypes = [P, P, I, P, P, I, P, I, D, D, D]
        l.orc_ellipsoid_plane.restype = D
        l.orc_ellipsoid_plane.argtypes = [P, P, P, D, P, D]
        l.orc_ellipsoid_line.restype = D
        l.orc_ellipsoid_line.argtypes = [P, P, P, D, P, P]
        _L = l
    return _L


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    return C.c_void_p(a.ctypes.data)


def phi_grid(nphi):
    o = np.empty(nphi)
    L().orc_phi_grid(nphi, _p(o))
    return o


def shift_grid(ns):
    o = np.empty(ns)
    L().orc_shift_grid(ns, _p(o))
    return o


def zslice_grid(nz):
    o = np.empty(nz)
    L().orc_zslice_grid(nz, _p(o))
    return o


def gl_nodes(n):
    x, w = np.empty(n), np.empty(n)
    L().orc_gl_nodes(n, _p(x), _p(w))
    return x, w


def theta_grid(n):
    t, w = np.empty(n), np.empty(n)
    L().orc_theta_grid(n, _p(t), _p(w))
    return t, w


def plane_basis(phi, theta):
    nr, v1, v2 = np.empty(3), np.empty(3), np.empty(3)
    L().orc_plane_basis(D(phi), D(theta), _p(nr), _p(v1), _p(v2))
    return nr, v1, v2


def plane_integral(vol, phi, theta, shift):
    vol = _f(vol)
    return L().orc_plane_integral(_p(vol), vol.shape[0], phi, theta, shift)


def plane_incube(n, phi, theta, shift):
    return L().orc_plane_incube(n, phi, theta, shift)


def zero_interp(vol, x, y, z):
    vol = _f(vol)
    return L().orc_zero_interp(_p(vol), vol.shape[0], x, y, z)


def trilinear(vol, x, y, z):
    vol = _f(vol)
    return L().orc_trilinear(_p(vol), vol.shape[0], x, y, z)


def ray_transform(vol, z, s, phi):
    vol = _f(vol)
    return L().orc_ray_transform(_p(vol), vol.shape[0], z, s, phi)


def radon_grid(vol, phi, theta, shift):
    vol, phi, theta, shift = map(_f, (vol, phi, theta, shift))
    out = np.empty((len(shift), len(phi), len(theta)))
    L().orc_radon_grid(_p(vol), vol.shape[0], _p(phi), len(phi), _p(theta), len(theta),
                       _p(shift), len(shift), _p(out))
    return out


def radon_planes(vol, phi, theta, shift):
    vol, phi, theta, shift = map(_f, (vol, phi, theta, shift))
    out = np.empty(len(phi))
    L().orc_radon_planes(_p(vol), vol.shape[0], len(phi), _p(phi), _p(theta), _p(shift), _p(out))
    return out


def ray_grid(vol, z, shift, phi):
    vol, z, shift, phi = map(_f, (vol, z, shift, phi))
    out = np.empty((len(z), len(shift), len(phi)))
    L().orc_ray_grid(_p(vol), vol.shape[0], _p(z), len(z), _p(shift), len(shift), _p(phi),
                     len(phi), _p(out))
    return out


def rays(vol, z, shift, phi):
    vol, z, shift, phi = map(_f, (vol, z, shift, phi))
    out = np.empty(len(z))
    L().orc_rays(_p(vol), vol.shape[0], len(z), _p(z), _p(shift), _p(phi), _p(out))
    return out


def cspline(x, y):
    x, y = _f(x), _f(y)
    c = np.empty(len(x))
    L().orc_cspline_init(_p(x), _p(y), len(x), _p(c))
    return c


def cspline_eval(x, y, c, v):
    x, y, c = map(_f, (x, y, c))
    return L().orc_cspline_eval(_p(x), _p(y), _p(c), len(x), v)


def poly3(xa, ya, x):
    xa, ya = _f(xa), _f(ya)
    return L().orc_poly3_eval(_p(xa), _p(ya), x)


def radon_reduction(rays_, shift, phi, theta):
    rays_, shift, phi, theta = map(_f, (rays_, shift, phi, theta))
    out = np.empty((len(shift), len(phi), len(theta)))
    L().orc_radon_reduction_grid(_p(rays_), rays_.shape[0], _p(shift), len(shift), _p(phi),
                                 len(phi), _p(theta), len(theta), _p(out))
    return out


def radon_reduction_rays(rays_, rshift, half_turn, shift, phi, theta):
    """rays_[z][rs][k] on nrphi = rays_.shape[2] angles k*pi/nrphi (half_turn) or k*2pi/nrphi
    -> out[s][phi][theta]"""
    rays_, rshift, shift, phi, theta = map(_f, (rays_, rshift, shift, phi, theta))
    out = np.empty((len(shift), len(phi), len(theta)))
    L().orc_radon_reduction_rays(_p(rays_), rays_.shape[0], _p(rshift), len(rshift),
                                 rays_.shape[2], int(bool(half_turn)), _p(shift), len(shift),
                                 _p(phi), len(phi), _p(theta), len(theta), _p(out))
    return out


def radon_reduction_rays_one(rays_, rshift, half_turn, p, phi, theta):
    rays_, rshift = _f(rays_), _f(rshift)
    return L().orc_radon_reduction_rays_one(_p(rays_), rays_.shape[0], _p(rshift), len(rshift),
                                            rays_.shape[2], int(bool(half_turn)), float(p),
                                            float(phi), float(theta))


def radon_adjoint(radon, phi, theta, theta_w, shift, n):
    radon, phi, theta, theta_w, shift = map(_f, (radon, phi, theta, theta_w, shift))
    out = np.empty((n, n, n))
    L().orc_radon_adjoint_grid(_p(radon), _p(phi), len(phi), _p(theta), _p(theta_w), len(theta),
                               _p(shift), len(shift), n, _p(out))
    return out


def radon_adjoint_point(radon, phi, theta, theta_w, shift, x, y, z):
    radon, phi, theta, theta_w, shift = map(_f, (radon, phi, theta, theta_w, shift))
    return L().orc_radon_adjoint_point(_p(radon), _p(phi), len(phi), _p(theta), _p(theta_w),
                                       len(theta), _p(shift), len(shift), x, y, z)


def radon_grid_T(g, n, phi, theta, shift):
    g, phi, theta, shift = map(_f, (g, phi, theta, shift))
    out = np.empty((n, n, n))
    L().orc_radon_grid_T(_p(g), n, _p(phi), len(phi), _p(theta), len(theta), _p(shift),
                         len(shift), _p(out))
    return out


def ray_grid_T(g, n, z, shift, phi):
    g, z, shift, phi = map(_f, (g, z, shift, phi))
    out = np.empty((n, n, n))
    L().orc_ray_grid_T(_p(g), n, _p(z), len(z), _p(shift), len(shift), _p(phi), len(phi),
                       _p(out))
    return out


def ellipsoid_plane(c, axes, rot, rho, nrm, p):
    c, axes, rot, nrm = map(_f, (c, axes, rot, nrm))
    return L().orc_ellipsoid_plane(_p(c), _p(axes), _p(rot), rho, _p(nrm), p)


def ellipsoid_line(c, axes, rot, rho, p0, d):
    c, axes, rot, p0, d = map(_f, (c, axes, rot, p0, d))
    return L().orc_ellipsoid_line(_p(c), _p(axes), _p(rot), rho, _p(p0), _p(d))


# ------------------------------------------------------------------ analytic programs ---
class FnTerm(C.Structure):
    """orc_fn_term (same layout as the product's wrt_fn_term)"""
    _fields_ = [("kind", C.c_int), ("c", C.c_double * 3), ("p", C.c_double * 13)]


def fn_terms(terms):
    """[(kind, centre[3], params[...]), ...] -> ctypes array of FnTerm"""
    arr = (FnTerm * max(1, len(terms)))()
    for k, (kind, c, p) in enumerate(terms):
        arr[k].kind = kind
        for i in range(3):
            arr[k].c[i] = c[i]
        for i, v in enumerate(p):
            arr[k].p[i] = v
    return arr


def _fn_lib():
    l = L()
    if not getattr(l, "_fn_ready", False):
        l.orc_fn_eval.restype = D
        l.orc_fn_eval.argtypes = [P, I, D, D, D]
        l.orc_analytic_plane.restype = D
        l.orc_analytic_plane.argtypes = [P, I, I, D, D, D]
        l.orc_analytic_ray.restype = D
        l.orc_analytic_ray.argtypes = [P, I, D, D, D, I]
        l._fn_ready = True
    return l


def fn_eval(terms, x, y, z):
    t = fn_terms(terms)
    return _fn_lib().orc_fn_eval(C.cast(t, P), len(terms), x, y, z)


def analytic_plane(terms, ngrid, phi, theta, shift):
    t = fn_terms(terms)
    return _fn_lib().orc_analytic_plane(C.cast(t, P), len(terms), ngrid, phi, theta, shift)


def analytic_ray(terms, zslice, shift, phi, ngrid):
    t = fn_terms(terms)
    return _fn_lib().orc_analytic_ray(C.cast(t, P), len(terms), zslice, shift, phi, ngrid)


def analytic_radon_grid(terms, ngrid, phi, theta, shift):
    phi, theta, shift = map(_f, (phi, theta, shift))
    t = fn_terms(terms)
    out = np.empty((len(shift), len(phi), len(theta)))
    _fn_lib().orc_analytic_radon_grid(C.cast(t, P), len(terms), ngrid, _p(phi), len(phi),
                                      _p(theta), len(theta), _p(shift), len(shift), _p(out))
    return out


def analytic_ray_grid(terms, ngrid, z, shift, phi):
    z, shift, phi = map(_f, (z, shift, phi))
    t = fn_terms(terms)
    out = np.empty((len(z), len(shift), len(phi)))
    _fn_lib().orc_analytic_ray_grid(C.cast(t, P), len(terms), ngrid, _p(z), len(z), _p(shift),
                                    len(shift), _p(phi), len(phi), _p(out))
    return out


def analytic_zslices(ngrid):
    o = np.empty(ngrid)
    L().orc_analytic_zslices(ngrid, _p(o))
    return o


def fourier_adjoint(rho, dirs, coef, ngrid):
    """exact adjoint sum of the Fourier-slice inversion (C, OpenMP); coef complex [ndir][n]"""
    rho, dirs = _f(rho), _f(dirs)
    c = np.ascontiguousarray(np.stack([coef.real, coef.imag], axis=-1), dtype=np.float64)
    out = np.empty((ngrid, ngrid, ngrid))
    L().orc_fourier_adjoint(_p(c), _p(rho), len(rho), _p(dirs), dirs.shape[0], ngrid, _p(out))
    return out


def fourier_module():
    """oracle/fourier.py (numpy restatement of rtft3d.m / chang3d.m)"""
    import importlib.util

    spec = importlib.util.spec_from_file_location("fourier", os.path.join(ROOT, "oracle", "fourier.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def chang3d(radon, phi, theta, theta_w, ngrid, w0=1.0, rsupp=1.0, padding=4.0):
    """the oracle's Fourier-slice inversion: numpy rtft3d, C exact adjoint sum"""
    rho, dirs, coef = fourier_module().rtft3d(radon, phi, theta, theta_w, rsupp, padding)
    return fourier_adjoint(rho, dirs, coef, ngrid) / w0


def chang2d(radon, phi, ngrid, w0=1.0, rsupp=1.0, padding=4.0):
    """the oracle's 2D Fourier-slice inversion (numpy rtft2d and exact adjoint sum)"""
    return fourier_module().chang2d(radon, phi, ngrid, w0, rsupp, padding)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def max_rel(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    m = np.max(np.abs(b)) if b.size else 0.0
    return float(np.max(np.abs(a - b)) / (m if m > 0 else 1.0)) if a.size else 0.0
