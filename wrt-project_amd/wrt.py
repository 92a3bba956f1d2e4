"""ctypes binding of lib/libwrt.so (the C-ABI declared in include/wrt.h).

Plumbing for tests and bench.py: numpy arrays in, numpy arrays out.  The binding loads the
in-tree library and raises WrtError if it is missing -- there is no fallback path.

Array layouts follow the reference's files: volumes vol[x][y][z]; radon [s][phi][theta];
rays [z][s][phi]; adjoint output [z][x][y].
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# WRT_LIB selects another build of the library (the CPU sanitizer build lib/san/libwrt.so)
LIB_PATH = os.environ.get("WRT_LIB") or os.path.join(HERE, "lib", "libwrt.so")

(WRT_OK, WRT_E_ARG, WRT_E_NODEV, WRT_E_HIP, WRT_E_NOMEM, WRT_E_ADMIT, WRT_E_IO, WRT_E_STATE,
 WRT_E_SCRATCH) = (0, -1, -2, -3, -4, -5, -6, -7, -8)
PREC = {"f64": 0, "f32": 1, "bf16": 2}
OP = {"radon": 0, "ray": 1, "reduction": 2, "adjoint": 3, "radon_T": 4, "ray_T": 5}


class WrtError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        super().__init__(f"{where}: {strerror(code)} ({code})")


class Stats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("kernel_seconds", C.c_double),
                ("logical_samples", C.c_double), ("alg_bytes", C.c_double),
                ("outputs", C.c_longlong)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class JobDesc(C.Structure):
    _fields_ = [("op", C.c_int), ("prec", C.c_int), ("n", C.c_int),
                ("phi", C.c_void_p), ("nphi", C.c_int),
                ("theta", C.c_void_p), ("theta_w", C.c_void_p), ("ntheta", C.c_int),
                ("shift", C.c_void_p), ("nshift", C.c_int),
                ("z", C.c_void_p), ("nz", C.c_int)]


class ServeCfg(C.Structure):
    _fields_ = [("qps", C.c_double), ("nreq", C.c_int), ("seed", C.c_ulonglong),
                ("mix", C.c_double * 3), ("edges", C.POINTER(C.c_int)), ("nedges", C.c_int),
                ("ndatasets", C.c_int), ("hbm_budget", C.c_size_t), ("max_inflight", C.c_int),
                ("plain", C.c_int)]


class ServeReq(C.Structure):
    _fields_ = [("op", C.c_int), ("prec", C.c_int), ("edge", C.c_int), ("dataset", C.c_int),
                ("data_seed", C.c_ulonglong), ("arrival_s", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServeReport(C.Structure):
    _fields_ = [("completed", C.c_int), ("wall_s", C.c_double), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double),
                ("samples", C.c_double), ("samples_per_s", C.c_double), ("loads", C.c_int),
                ("evictions", C.c_int), ("peak_inflight", C.c_int), ("peak_bytes", C.c_size_t),
                ("batched", C.c_int), ("rounds", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None

# every symbol declared in include/wrt.h
ABI_SYMBOLS = [
    "wrt_init", "wrt_free", "wrt_strerror", "wrt_mem_info", "wrt_gl_nodes",
    "wrt_radon_grid_init", "wrt_ray_grid_init", "wrt_zslice_grid", "wrt_radon_grid",
    "wrt_ray_grid", "wrt_radon_reduction", "wrt_radon_reduction_rays", "wrt_radon_adjoint", "wrt_radon_grid_T",
    "wrt_ray_grid_T", "wrt_job_bytes", "wrt_job_input_count", "wrt_job_output_count",
    "wrt_job_create", "wrt_job_load", "wrt_job_load_synthetic", "wrt_job_run",
    "wrt_job_fetch", "wrt_job_destroy", "wrt_job_tuning", "wrt_batch_run", "wrt_phantom",
    "wrt_phantom_params", "wrt_gaussian", "wrt_radon_count_incube", "wrt_job_launch",
    "wrt_job_poll", "wrt_job_wait", "wrt_serve_describe", "wrt_serve_grids", "wrt_serve_desc",
    "wrt_serve_input", "wrt_serve_run", "wrt_batch_dgroup",
    "wrt_fn_example", "wrt_fn_phantom", "wrt_analytic_zslice_grid", "wrt_radon_analytic",
    "wrt_ray_analytic", "wrt_laplace3d", "wrt_laplace2d", "wrt_read_text",
    "wrt_write_text_grid", "wrt_write_text_values", "wrt_write_f64", "wrt_read_f64",
    "wrt_fourier_inversion", "wrt_fourier_inversion2d", "wrt_radon_grid_noalloc",
]

FN = {"exp_bump": 0, "sine_bump": 1, "ellipsoid": 2}
FN_MAX_TERMS = 256


class FnTerm(C.Structure):
    """wrt_fn_term: kind, centre, 13 parameters"""
    _fields_ = [("kind", C.c_int), ("c", C.c_double * 3), ("p", C.c_double * 13)]


def fn_terms(terms):
    """[(kind, centre[3], params[...]), ...] (kind an int or a FN key) -> ctypes array"""
    arr = (FnTerm * max(1, len(terms)))()
    for k, (kind, c, p) in enumerate(terms):
        arr[k].kind = FN[kind] if isinstance(kind, str) else int(kind)
        for i in range(3):
            arr[k].c[i] = c[i]
        for i, v in enumerate(p):
            arr[k].p[i] = v
    return arr


def _terms_list(arr, count):
    return [(arr[k].kind, tuple(arr[k].c), tuple(arr[k].p)) for k in range(count)]


def fn_example(name):
    """the reference's example test function as terms"""
    arr = (FnTerm * 8)()
    rc = lib().wrt_fn_example(name.encode(), arr, 8)
    if rc < 0:
        raise WrtError(rc, "wrt_fn_example")
    return _terms_list(arr, rc)


def fn_phantom(nell, seed=0):
    arr = (FnTerm * max(1, nell))()
    _chk(lib().wrt_fn_phantom(nell, seed, arr), "wrt_fn_phantom")
    return _terms_list(arr, nell)


def read_text(path, count, cols):
    out = np.empty(count)
    _chk(lib().wrt_read_text(str(path).encode(), count, cols, _p(out)), "wrt_read_text")
    return out


def write_text_grid(path, a, b, c, values):
    a, b, c, values = map(_f64, (a, b, c, values))
    assert values.size == a.size * b.size * c.size
    _chk(lib().wrt_write_text_grid(str(path).encode(), _p(a), a.size, _p(b), b.size, _p(c), c.size,
                                   _p(values)), "wrt_write_text_grid")


def write_text_values(path, values):
    values = _f64(values).ravel()
    _chk(lib().wrt_write_text_values(str(path).encode(), _p(values), values.size),
         "wrt_write_text_values")


def write_f64(path, values):
    values = _f64(values).ravel()
    _chk(lib().wrt_write_f64(str(path).encode(), _p(values), values.size), "wrt_write_f64")


def read_f64(path, count):
    out = np.empty(count)
    _chk(lib().wrt_read_f64(str(path).encode(), count, _p(out)), "wrt_read_f64")
    return out


def analytic_zslice_grid(ngrid):
    o = np.empty(ngrid)
    _chk(lib().wrt_analytic_zslice_grid(ngrid, _p(o)), "wrt_analytic_zslice_grid")
    return o


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise WrtError(WRT_E_NODEV, f"libwrt.so not built at {LIB_PATH}")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        I = C.c_int
        D = C.c_double
        L.wrt_strerror.restype = C.c_char_p
        L.wrt_strerror.argtypes = [I]
        L.wrt_init.argtypes = [C.POINTER(P), I]
        L.wrt_free.argtypes = [P]
        L.wrt_free.restype = None
        L.wrt_mem_info.argtypes = [P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        L.wrt_gl_nodes.argtypes = [P, I, P, P]
        L.wrt_radon_grid_init.argtypes = [P, I, I, I, P, P, P, P]
        L.wrt_ray_grid_init.argtypes = [I, I, P, P]
        L.wrt_zslice_grid.argtypes = [I, P]
        L.wrt_radon_grid.argtypes = [P, P, I, P, I, P, I, P, I, I, P, P]
        L.wrt_radon_grid_noalloc.argtypes = [P, P, I, P, I, P, I, P, I, P, P]
        L.wrt_ray_grid.argtypes = [P, P, I, P, I, P, I, P, I, I, P, P]
        L.wrt_radon_reduction.argtypes = [P, P, I, P, I, P, I, P, I, P, P]
        L.wrt_radon_reduction_rays.argtypes = [P, P, I, P, I, I, I, P, I, P, I, P, I, P, P]
        L.wrt_radon_adjoint.argtypes = [P, P, P, I, P, P, I, P, I, I, I, P, P]
        L.wrt_radon_grid_T.argtypes = [P, P, I, P, I, P, I, P, I, I, P, P]
        L.wrt_ray_grid_T.argtypes = [P, P, I, P, I, P, I, P, I, I, P, P]
        L.wrt_job_bytes.argtypes = [P]
        L.wrt_job_bytes.restype = C.c_size_t
        L.wrt_job_input_count.argtypes = [P]
        L.wrt_job_input_count.restype = C.c_size_t
        L.wrt_job_output_count.argtypes = [P]
        L.wrt_job_output_count.restype = C.c_size_t
        L.wrt_job_create.argtypes = [P, P, C.POINTER(P)]
        L.wrt_job_load.argtypes = [P, P]
        L.wrt_job_load_synthetic.argtypes = [P, C.c_ulonglong, I, D]
        L.wrt_job_run.argtypes = [P, P]
        L.wrt_job_launch.argtypes = [P]
        L.wrt_job_poll.argtypes = [P, P]
        L.wrt_job_wait.argtypes = [P, P]
        L.wrt_job_fetch.argtypes = [P, P]
        L.wrt_job_tuning.argtypes = [P, C.POINTER(I), C.POINTER(I)]
        L.wrt_job_destroy.argtypes = [P]
        L.wrt_job_destroy.restype = None
        L.wrt_batch_run.argtypes = [P, P, I, P]
        L.wrt_phantom.argtypes = [I, I, I, C.c_ulonglong, D, P]
        L.wrt_phantom_params.argtypes = [I, C.c_ulonglong, P]
        L.wrt_gaussian.argtypes = [C.c_size_t, C.c_ulonglong, P]
        L.wrt_radon_count_incube.argtypes = [P, I, P, I, P, I, P, I, P]
        L.wrt_batch_dgroup.argtypes = [P]
        L.wrt_serve_describe.argtypes = [P, I, P]
        L.wrt_serve_grids.argtypes = [P, I, P, P, P, P, P]
        L.wrt_serve_desc.argtypes = [P, P, P, P, P, P, P]
        L.wrt_serve_input.argtypes = [P, P]
        L.wrt_serve_run.argtypes = [P, P, P, P, P]
        L.wrt_read_text.argtypes = [C.c_char_p, C.c_size_t, I, P]
        L.wrt_write_text_grid.argtypes = [C.c_char_p, P, I, P, I, P, I, P]
        L.wrt_write_text_values.argtypes = [C.c_char_p, P, C.c_size_t]
        L.wrt_write_f64.argtypes = [C.c_char_p, P, C.c_size_t]
        L.wrt_read_f64.argtypes = [C.c_char_p, C.c_size_t, P]
        L.wrt_fourier_inversion.argtypes = [P, P, P, I, P, P, I, I, D, D, D, I, P, P]
        L.wrt_fourier_inversion2d.argtypes = [P, P, P, I, I, D, D, D, I, P, P]
        L.wrt_laplace3d.argtypes = [P, P, I, D, P, P]
        L.wrt_laplace2d.argtypes = [P, P, I, D, P, P]
        L.wrt_fn_example.argtypes = [C.c_char_p, P, I]
        L.wrt_fn_phantom.argtypes = [I, C.c_ulonglong, P]
        L.wrt_analytic_zslice_grid.argtypes = [I, P]
        L.wrt_radon_analytic.argtypes = [P, P, I, I, P, I, P, I, P, I, P, P]
        L.wrt_ray_analytic.argtypes = [P, P, I, I, P, I, P, I, P, I, P, P]
        _lib = L
    return _lib


def strerror(code):
    try:
        return lib().wrt_strerror(code).decode()
    except Exception:  # library missing
        return f"error {code}"


def _chk(rc, where):
    if rc != WRT_OK:
        raise WrtError(rc, where)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    return None if a is None else a.ctypes.data


# ------------------------------------------------------------------ host-only helpers ---
def phantom(n, cell_centred=True, nell=10, seed=0, sigma=0.0):
    out = np.empty((n, n, n))
    _chk(lib().wrt_phantom(n, int(bool(cell_centred)), nell, seed, sigma, _p(out)), "wrt_phantom")
    return out


def phantom_params(nell, seed=0):
    out = np.empty((nell, 16))
    _chk(lib().wrt_phantom_params(nell, seed, _p(out)), "wrt_phantom_params")
    return out


def gaussian(shape, seed=0):
    out = np.empty(shape)
    _chk(lib().wrt_gaussian(out.size, seed, _p(out)), "wrt_gaussian")
    return out


def ray_grid_init(nshift, nphi):
    s, p = np.empty(nshift), np.empty(nphi)
    _chk(lib().wrt_ray_grid_init(nshift, nphi, _p(s), _p(p)), "wrt_ray_grid_init")
    return s, p


def zslice_grid(nz):
    z = np.empty(nz)
    _chk(lib().wrt_zslice_grid(nz, _p(z)), "wrt_zslice_grid")
    return z


# ---------------------------------------------------------------------------- context ---
class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        _chk(lib().wrt_init(C.byref(h), device), "wrt_init")
        self.h = h

    def close(self):
        if self.h:
            lib().wrt_free(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def mem_info(self):
        f, t = C.c_size_t(), C.c_size_t()
        _chk(lib().wrt_mem_info(self.h, C.byref(f), C.byref(t)), "wrt_mem_info")
        return f.value, t.value

    def gl_nodes(self, n):
        x, w = np.empty(n), np.empty(n)
        _chk(lib().wrt_gl_nodes(self.h, n, _p(x), _p(w)), "wrt_gl_nodes")
        return x, w

    def radon_grid_init(self, nphi, ntheta, nshift):
        phi, th, w, s = np.empty(nphi), np.empty(ntheta), np.empty(ntheta), np.empty(nshift)
        _chk(lib().wrt_radon_grid_init(self.h, nphi, ntheta, nshift, _p(phi), _p(th), _p(w),
                                       _p(s)), "wrt_radon_grid_init")
        return phi, th, w, s

    def radon_grid(self, vol, phi, theta, shift, prec="f64"):
        vol, phi, theta, shift = map(_f64, (vol, phi, theta, shift))
        n = vol.shape[0]
        out = np.empty((len(shift), len(phi), len(theta)))
        st = Stats()
        _chk(lib().wrt_radon_grid(self.h, _p(vol), n, _p(phi), len(phi), _p(theta), len(theta),
                                  _p(shift), len(shift), PREC[prec], _p(out), C.byref(st)),
             "wrt_radon_grid")
        return out, st.as_dict()

    def radon_grid_noalloc(self, vol, phi, theta, shift):
        """wrt_radon_grid's FP32 form through the entry point that allocates nothing"""
        vol, phi, theta, shift = map(_f64, (vol, phi, theta, shift))
        n = vol.shape[0]
        out = np.empty((len(shift), len(phi), len(theta)))
        st = Stats()
        _chk(lib().wrt_radon_grid_noalloc(self.h, _p(vol), n, _p(phi), len(phi), _p(theta),
                                          len(theta), _p(shift), len(shift), _p(out),
                                          C.byref(st)), "wrt_radon_grid_noalloc")
        return out, st.as_dict()

    def ray_grid(self, vol, z, shift, phi, prec="f64"):
        vol, z, shift, phi = map(_f64, (vol, z, shift, phi))
        n = vol.shape[0]
        out = np.empty((len(z), len(shift), len(phi)))
        st = Stats()
        _chk(lib().wrt_ray_grid(self.h, _p(vol), n, _p(z), len(z), _p(shift), len(shift),
                                _p(phi), len(phi), PREC[prec], _p(out), C.byref(st)),
             "wrt_ray_grid")
        return out, st.as_dict()

    def radon_reduction(self, rays, shift, phi, theta):
        rays, shift, phi, theta = map(_f64, (rays, shift, phi, theta))
        out = np.empty((len(shift), len(phi), len(theta)))
        st = Stats()
        _chk(lib().wrt_radon_reduction(self.h, _p(rays), rays.shape[0], _p(shift), len(shift),
                                       _p(phi), len(phi), _p(theta), len(theta), _p(out),
                                       C.byref(st)), "wrt_radon_reduction")
        return out, st.as_dict()

    def radon_reduction_rays(self, rays, rshift, half_turn, shift, phi, theta):
        """rays[z][s][k] on their own shift grid and nrphi = rays.shape[2] angles k*pi/nrphi
        (half_turn) or k*2pi/nrphi -> out[s][phi][theta] at the planes (shift, phi, theta)."""
        rays, rshift, shift, phi, theta = map(_f64, (rays, rshift, shift, phi, theta))
        if rays.ndim != 3 or rays.shape[1] != len(rshift):
            raise ValueError("rays must be [z][s][phi] on the ray shift grid")
        out = np.empty((len(shift), len(phi), len(theta)))
        st = Stats()
        _chk(lib().wrt_radon_reduction_rays(self.h, _p(rays), rays.shape[0], _p(rshift),
                                            len(rshift), rays.shape[2], int(bool(half_turn)),
                                            _p(shift), len(shift), _p(phi), len(phi), _p(theta),
                                            len(theta), _p(out), C.byref(st)),
             "wrt_radon_reduction_rays")
        return out, st.as_dict()

    def radon_adjoint(self, radon, phi, theta, theta_w, shift, n, prec="f64"):
        radon, phi, theta, theta_w, shift = map(_f64, (radon, phi, theta, theta_w, shift))
        out = np.empty((n, n, n))
        st = Stats()
        _chk(lib().wrt_radon_adjoint(self.h, _p(radon), _p(phi), len(phi), _p(theta),
                                     _p(theta_w), len(theta), _p(shift), len(shift), n,
                                     PREC[prec], _p(out), C.byref(st)), "wrt_radon_adjoint")
        return out, st.as_dict()

    def radon_grid_T(self, g, n, phi, theta, shift, prec="f64"):
        g, phi, theta, shift = map(_f64, (g, phi, theta, shift))
        out = np.empty((n, n, n))
        st = Stats()
        _chk(lib().wrt_radon_grid_T(self.h, _p(g), n, _p(phi), len(phi), _p(theta), len(theta),
                                    _p(shift), len(shift), PREC[prec], _p(out), C.byref(st)),
             "wrt_radon_grid_T")
        return out, st.as_dict()

    def ray_grid_T(self, g, n, z, shift, phi, prec="f64"):
        g, z, shift, phi = map(_f64, (g, z, shift, phi))
        out = np.empty((n, n, n))
        st = Stats()
        _chk(lib().wrt_ray_grid_T(self.h, _p(g), n, _p(z), len(z), _p(shift), len(shift),
                                  _p(phi), len(phi), PREC[prec], _p(out), C.byref(st)),
             "wrt_ray_grid_T")
        return out, st.as_dict()

    def radon_analytic(self, terms, ngrid, phi, theta, shift):
        """plane integrals of the test function `terms` (list of (kind, c, p)), FP64"""
        phi, theta, shift = map(_f64, (phi, theta, shift))
        t = fn_terms(terms)
        out = np.empty((len(shift), len(phi), len(theta)))
        st = Stats()
        _chk(lib().wrt_radon_analytic(self.h, t, len(terms), ngrid, _p(phi), len(phi), _p(theta),
                                      len(theta), _p(shift), len(shift), _p(out), C.byref(st)),
             "wrt_radon_analytic")
        return out, st.as_dict()

    def ray_analytic(self, terms, ngrid, z, shift, phi):
        z, shift, phi = map(_f64, (z, shift, phi))
        t = fn_terms(terms)
        out = np.empty((len(z), len(shift), len(phi)))
        st = Stats()
        _chk(lib().wrt_ray_analytic(self.h, t, len(terms), ngrid, _p(z), len(z), _p(shift),
                                    len(shift), _p(phi), len(phi), _p(out), C.byref(st)),
             "wrt_ray_analytic")
        return out, st.as_dict()

    def fourier_inversion(self, radon, phi, theta, theta_w, ngrid, w0=1.0, rsupp=1.0,
                          padding=4.0):
        """Chang-type Fourier-slice inversion of radon[s][phi][theta] to ngrid^3, FP64"""
        radon, phi, theta, theta_w = map(_f64, (radon, phi, theta, theta_w))
        out = np.empty((ngrid, ngrid, ngrid))
        st = Stats()
        _chk(lib().wrt_fourier_inversion(self.h, _p(radon), _p(phi), len(phi), _p(theta),
                                         _p(theta_w), len(theta), radon.shape[0], rsupp, padding,
                                         w0, ngrid, _p(out), C.byref(st)), "wrt_fourier_inversion")
        return out, st.as_dict()

    def fourier_inversion2d(self, radon, phi, ngrid, w0=1.0, rsupp=1.0, padding=4.0):
        """Chang-type Fourier-slice inversion of radon[s][phi] to ngrid^2 [x][y], FP64"""
        radon, phi = _f64(radon), _f64(phi)
        out = np.empty((ngrid, ngrid))
        st = Stats()
        _chk(lib().wrt_fourier_inversion2d(self.h, _p(radon), _p(phi), len(phi), radon.shape[0],
                                           rsupp, padding, w0, ngrid, _p(out), C.byref(st)),
             "wrt_fourier_inversion2d")
        return out, st.as_dict()

    def chang2d(self, ngrid, w0, input_filename, nphi, nshift, rsupp=1.0, padd=4.0):
        """the reference's chang2d(ngrid, w0, input_filename, nphi, nshift, rsupp, padd)
        (radon_inversion/chang2d.m:3-26): reads the last column of the "s, phi, value" file
        (s -> phi), phi = k 2 pi / nphi, and returns the ngrid^2 reconstruction [x][y] divided
        by w0"""
        radon = read_text(input_filename, nshift * nphi, 3)
        phi = np.arange(nphi) * (2 * np.pi / nphi)
        out, _ = self.fourier_inversion2d(radon.reshape(nshift, nphi), phi, ngrid, w0=w0,
                                          rsupp=rsupp, padding=padd)
        return out

    def chang3d(self, ngrid, w0, input_filename, nphi, ntheta, nshift, rsupp=1.0, padd=4.0):
        """the reference's chang3d(ngrid, w0, input_filename, nphi, ntheta, nshift, rsupp,
        padd) (radon_inversion/chang3d.m:3-30): reads the last column of the Radon data file
        (s -> phi -> theta), with phi = k 2 pi / nphi and theta = acos of the Gauss-Legendre
        nodes in descending order (lgwt's order, rtft3d.m:32-33), and returns the
        ngrid^3 reconstruction [x][y][z] divided by w0"""
        radon = read_text(input_filename, nshift * nphi * ntheta, 4)
        _, theta, w, _ = self.radon_grid_init(nphi, ntheta, nshift)
        theta, w = theta[::-1].copy(), w[::-1].copy()  # nodes cos(theta) descending
        phi = np.arange(nphi) * (2 * np.pi / nphi)
        out, _ = self.fourier_inversion(radon.reshape(nshift, nphi, ntheta), phi, theta, w,
                                        ngrid, w0=w0, rsupp=rsupp, padding=padd)
        return out

    def laplace(self, u, period):
        """spectral Laplacian of a 2-D or 3-D FP64 array with equal sides"""
        u = _f64(u)
        n = u.shape[0]
        out = np.empty_like(u)
        st = Stats()
        fn = lib().wrt_laplace3d if u.ndim == 3 else lib().wrt_laplace2d
        _chk(fn(self.h, _p(u), n, float(period), _p(out), C.byref(st)), "wrt_laplace")
        return out, st.as_dict()

    def radon_count_incube(self, n, phi, theta, shift):
        phi, theta, shift = map(_f64, (phi, theta, shift))
        out = C.c_double()
        _chk(lib().wrt_radon_count_incube(self.h, n, _p(phi), len(phi), _p(theta), len(theta),
                                          _p(shift), len(shift), C.byref(out)),
             "wrt_radon_count_incube")
        return out.value

    def serve_grids(self, edge):
        """phi, theta, theta_w, shift, z of a serve request of this cube edge"""
        phi, th, w = np.empty(edge // 2), np.empty(edge // 4), np.empty(edge // 4)
        s, z = np.empty(edge + 1), np.empty(edge)
        _chk(lib().wrt_serve_grids(self.h, edge, _p(phi), _p(th), _p(w), _p(s), _p(z)),
             "wrt_serve_grids")
        return phi, th, w, s, z

    def serve_job(self, req):
        """A job for a serve request, on its grids, loaded with its dataset's input."""
        phi, th, w, s, z = self.serve_grids(req.edge)
        name = {v: k for k, v in OP.items()}[req.op]
        prec = {v: k for k, v in PREC.items()}[req.prec]
        kw = dict(phi=phi, theta=th, shift=s)
        if name == "reduction":
            kw["z"] = z
        else:
            kw["n"] = req.edge
        if name == "adjoint":
            kw["theta_w"] = w
        j = self.job(name, prec, **kw)
        j.load(serve_input(req))
        return j

    def serve(self, cfg, keep=()):
        """Run the stream; returns (report dict, latencies in ms, {request: output})."""
        L = lib()
        lat = np.empty(cfg.nreq)
        outs = {}
        ptrs = (C.c_void_p * cfg.nreq)()
        for i in keep:
            r = serve_describe(cfg, i)
            e = r.edge
            cnt = {OP["radon"]: (e + 1) * (e // 2) * (e // 4), OP["adjoint"]: e ** 3,
                   OP["reduction"]: (e + 1) * (e // 2) * (e // 4)}[r.op]
            outs[i] = np.empty(cnt)
            ptrs[i] = outs[i].ctypes.data
        rep = ServeReport()
        _chk(L.wrt_serve_run(self.h, C.byref(cfg), C.byref(rep), _p(lat), ptrs), "wrt_serve_run")
        return rep.as_dict(), lat, outs

    def job(self, op, prec="f64", n=0, phi=None, theta=None, theta_w=None, shift=None, z=None,
            nz=None):
        return Job(self, op, prec, n, phi, theta, theta_w, shift, z, nz)

    def batch_run(self, jobs):
        arr = (C.c_void_p * len(jobs))(*[j.h for j in jobs])
        st = Stats()
        _chk(lib().wrt_batch_run(self.h, arr, len(jobs), C.byref(st)), "wrt_batch_run")
        return st.as_dict()

    def batch_dgroup(self):
        return lib().wrt_batch_dgroup(self.h)


def make_desc(op, prec="f64", n=0, phi=None, theta=None, theta_w=None, shift=None, z=None,
              nz=None):
    keep = [None if a is None else _f64(a) for a in (phi, theta, theta_w, shift, z)]
    phi, theta, theta_w, shift, z = keep
    d = JobDesc(OP[op], PREC[prec], n, _p(phi), 0 if phi is None else len(phi), _p(theta),
                _p(theta_w), 0 if theta is None else len(theta), _p(shift),
                0 if shift is None else len(shift), _p(z),
                nz if nz is not None else (0 if z is None else len(z)))
    return d, keep


class Job:
    def __init__(self, ctx, op, prec, n, phi, theta, theta_w, shift, z, nz):
        self.ctx = ctx
        self.desc, self._keep = make_desc(op, prec, n, phi, theta, theta_w, shift, z, nz)
        h = C.c_void_p()
        _chk(lib().wrt_job_create(ctx.h, C.byref(self.desc), C.byref(h)), "wrt_job_create")
        self.h = h
        self.nin = lib().wrt_job_input_count(C.byref(self.desc))
        self.nout = lib().wrt_job_output_count(C.byref(self.desc))

    @property
    def bytes(self):
        return lib().wrt_job_bytes(C.byref(self.desc))

    def load(self, a):
        a = _f64(a)
        assert a.size == self.nin, (a.size, self.nin)
        _chk(lib().wrt_job_load(self.h, _p(a)), "wrt_job_load")

    def load_synthetic(self, seed, nell=10, sigma=0.0):
        _chk(lib().wrt_job_load_synthetic(self.h, seed, nell, sigma), "wrt_job_load_synthetic")

    def run(self):
        st = Stats()
        _chk(lib().wrt_job_run(self.h, C.byref(st)), "wrt_job_run")
        return st.as_dict()

    def launch(self):
        _chk(lib().wrt_job_launch(self.h), "wrt_job_launch")

    def poll(self):
        st = Stats()
        rc = lib().wrt_job_poll(self.h, C.byref(st))
        if rc == 1:
            return None
        _chk(rc, "wrt_job_poll")
        return st.as_dict()

    def wait(self):
        st = Stats()
        _chk(lib().wrt_job_wait(self.h, C.byref(st)), "wrt_job_wait")
        return st.as_dict()

    def fetch(self):
        out = np.empty(self.nout)
        _chk(lib().wrt_job_fetch(self.h, _p(out)), "wrt_job_fetch")
        return out

    def tuning(self):
        """(rows per band, directions per task) of a fast plane-integral job after its first
        launch."""
        br, dg = C.c_int(0), C.c_int(0)
        _chk(lib().wrt_job_tuning(self.h, C.byref(br), C.byref(dg)), "wrt_job_tuning")
        return br.value, dg.value

    def close(self):
        if self.h:
            lib().wrt_job_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def job_bytes(op, prec="f64", **kw):
    d, keep = make_desc(op, prec, **kw)
    return lib().wrt_job_bytes(C.byref(d))


# ------------------------------------------------------------------------ serve loop ---
def serve_cfg(qps=8.0, nreq=64, seed=1, mix=(0.5, 0.3, 0.2), edges=(64, 96, 128, 192, 256),
              ndatasets=2, hbm_budget=0, max_inflight=4, plain=False):
    """A wrt_serve_cfg and the array it points to (keep both alive while in use)."""
    e = (C.c_int * len(edges))(*edges)
    cfg = ServeCfg(qps, nreq, seed, (C.c_double * 3)(*mix), e, len(edges), ndatasets,
                   hbm_budget, max_inflight, int(bool(plain)))
    cfg._edges = e
    return cfg


def serve_describe(cfg, i):
    r = ServeReq()
    _chk(lib().wrt_serve_describe(C.byref(cfg), i, C.byref(r)), "wrt_serve_describe")
    return r


def serve_input(req):
    """The request's dataset input as the flat array wrt_job_load takes."""
    e = req.edge
    count = {OP["radon"]: e ** 3, OP["adjoint"]: (e + 1) * (e // 2) * (e // 4),
             OP["reduction"]: e * (e + 1) * (e // 2)}[req.op]
    out = np.empty(count)
    _chk(lib().wrt_serve_input(C.byref(req), _p(out)), "wrt_serve_input")
    return out
