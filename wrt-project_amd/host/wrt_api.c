/*
 * wrt_api.c -- C host layer of the wrt C-ABI: context, grids, per-direction tables, jobs
 * and the host-buffer entry points.  All device work is enqueued through the launchers of
 * csrc/wrt_dev.h on the context stream.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "wrt_internal.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define HIPCHK(x)                                \
  do {                                           \
    hipError_t e__ = (x);                        \
    if (e__ != hipSuccess) return WRT_E_HIP;     \
  } while (0)
#define RCHK(x)                 \
  do {                          \
    int r__ = (x);              \
    if (r__ != WRT_OK) return r__; \
  } while (0)

int wrt_hip(hipError_t e) { return e == hipSuccess ? WRT_OK : WRT_E_HIP; }

const char* wrt_strerror(int code) {
  switch (code) {
    case WRT_OK: return "ok";
    case WRT_E_ARG: return "invalid argument";
    case WRT_E_NODEV: return "no usable HIP device";
    case WRT_E_HIP: return "HIP runtime error";
    case WRT_E_NOMEM: return "out of memory";
    case WRT_E_ADMIT: return "job does not fit the free-HBM budget";
    case WRT_E_IO: return "file open or parse error";
    case WRT_E_STATE: return "call out of order";
    case WRT_E_SCRATCH: return "the call needs more than the fixed no-allocation scratch";
    default: return "unknown error";
  }
}

/* ----------------------------------------------------------------------- context --- */
int wrt_init(wrt_ctx** out, int device) {
  if (!out) return WRT_E_ARG;
  *out = NULL;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return WRT_E_NODEV;
  if (device < 0 || device >= count) return WRT_E_ARG;
  HIPCHK(hipSetDevice(device));
  wrt_ctx* c = (wrt_ctx*)calloc(1, sizeof(wrt_ctx));
  if (!c) return WRT_E_NOMEM;
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      hipEventCreate(&c->ev2) != hipSuccess || hipEventCreate(&c->ev3) != hipSuccess) {
    free(c);
    return WRT_E_HIP;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    cus = 256;
  c->cu_count = cus;
  /* the device address of this GPU's copy of the no-allocation scratch */
  if (wrtk_noalloc_setup(device) != WRT_OK) {
    hipEventDestroy(c->ev0);
    hipEventDestroy(c->ev1);
    hipEventDestroy(c->ev2);
    hipEventDestroy(c->ev3);
    hipStreamDestroy(c->stream);
    free(c);
    return WRT_E_HIP;
  }
  *out = c;
  return WRT_OK;
}

void wrt_batch_release(wrt_ctx* ctx); /* wrt_batch.c */

void wrt_free(wrt_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  wrt_batch_release(c);
  wrt_spectral_release(c);
  hipEventDestroy(c->ev0);
  hipEventDestroy(c->ev1);
  hipEventDestroy(c->ev2);
  hipEventDestroy(c->ev3);
  hipStreamDestroy(c->stream);
  free(c);
}

int wrt_mem_info(wrt_ctx* c, size_t* fr, size_t* tot) {
  if (!c) return WRT_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  size_t f = 0, t = 0;
  HIPCHK(hipMemGetInfo(&f, &t));
  if (fr) *fr = f;
  if (tot) *tot = t;
  return WRT_OK;
}

/* ------------------------------------------------------------------------- grids --- */
int wrt_gl_nodes(wrt_ctx* c, int n, double* x, double* w) {
  if (!c || n < 1 || !x || !w) return WRT_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  double* d = NULL;
  HIPCHK(hipMalloc((void**)&d, sizeof(double) * 2 * (size_t)n));
  int rc = wrtk_gl(n, d, d + n, c->stream);
  if (rc == WRT_OK) rc = wrt_hip(hipMemcpyAsync(x, d, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipMemcpyAsync(w, d + n, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipStreamSynchronize(c->stream));
  hipFree(d);
  return rc;
}

int wrt_radon_grid_init(wrt_ctx* c, int nphi, int ntheta, int nshift, double* phi,
                        double* theta, double* theta_w, double* shift) {
  if (nphi < 1 || ntheta < 1 || nshift < 2 || !phi || !theta || !shift) return WRT_E_ARG;
  const double dphi = (2 * M_PI) / nphi;
  for (int k = 0; k < nphi; ++k) phi[k] = dphi * k;
  const double ds = 2.0 / (nshift - 1);
  for (int j = 0; j < nshift; ++j) shift[j] = (-1.0) + j * ds;
  double* x = (double*)malloc(sizeof(double) * 2 * (size_t)ntheta);
  if (!x) return WRT_E_NOMEM;
  int rc = wrt_gl_nodes(c, ntheta, x, x + ntheta);
  if (rc == WRT_OK)
    for (int i = 0; i < ntheta; ++i) {
      theta[i] = acos(x[i]);
      if (theta_w) theta_w[i] = x[ntheta + i];
    }
  free(x);
  return rc;
}

int wrt_ray_grid_init(int nshift, int nphi, double* shift, double* phi) {
  if (nshift < 2 || nphi < 1 || !shift || !phi) return WRT_E_ARG;
  const double ds = 2.0 / (nshift - 1);
  for (int j = 0; j < nshift; ++j) shift[j] = (-1.0) + j * ds;
  const double dphi = (2 * M_PI) / nphi;
  for (int k = 0; k < nphi; ++k) phi[k] = dphi * k;
  return WRT_OK;
}

int wrt_zslice_grid(int nslices, double* z) {
  if (nslices < 2 || !z) return WRT_E_ARG;
  const double dz = 2.0 / (nslices - 1);
  for (int l = 0; l < nslices; ++l) z[l] = (-1.0) + l * dz;
  return WRT_OK;
}

/* normal, v1 = direction(phi, theta + pi/2), v2 = normal x v1 (with the reference's sign
 * convention), then a = delta*v1, b = delta*v2: radon_grid/src/integral.c:10-36,47-49 */
void wrt_host_radon_dirs(int n, const double* phi, int nphi, const double* theta, int ntheta,
                         double* dirs) {
  wrt_host_plane_dirs((1.0 / (n)), phi, nphi, theta, ntheta, dirs);
}

/* the same basis for a lattice step delta (radon_analytic/src/integral.c:9-22,36-45 forms
 * it identically with delta = (2/(ngrid-1))/2) */
void wrt_host_plane_dirs(double delta, const double* phi, int nphi, const double* theta,
                         int ntheta, double* dirs) {
  for (int ip = 0; ip < nphi; ++ip)
    for (int it = 0; it < ntheta; ++it) {
      const double p = phi[ip], t = theta[it];
      double nr[3], v1[3], v2[3];
      nr[0] = sin(t) * cos(p);
      nr[1] = sin(t) * sin(p);
      nr[2] = cos(t);
      const double t2 = t + M_PI / 2;
      v1[0] = sin(t2) * cos(p);
      v1[1] = sin(t2) * sin(p);
      v1[2] = cos(t2);
      v2[0] = nr[1] * v1[2] - nr[2] * v1[1];
      v2[1] = -(nr[0] * v1[2] - nr[2] * v1[0]);
      v2[2] = nr[0] * v1[1] - nr[1] * v1[0];
      double* D = dirs + 9 * ((size_t)ip * ntheta + it);
      D[0] = nr[0];
      D[1] = nr[1];
      D[2] = nr[2];
      for (int k = 0; k < 3; ++k) {
        D[3 + k] = delta * v1[k];
        D[6 + k] = delta * v2[k];
      }
    }
}

/* normal (cos phi, sin phi), direction (cos(phi+pi/2), sin(phi+pi/2)):
 * ray_grid/src/integral.c:16-24 */
void wrt_host_ray_trig(const double* phi, int nphi, double* trig) {
  for (int k = 0; k < nphi; ++k) {
    trig[4 * k + 0] = cos(phi[k]);
    trig[4 * k + 1] = sin(phi[k]);
    trig[4 * k + 2] = cos(phi[k] + M_PI / 2);
    trig[4 * k + 3] = sin(phi[k] + M_PI / 2);
  }
}

int wrt_check_grid(const double* g, int count, int min_count) {
  if (count < min_count || !g) return WRT_E_ARG;
  for (int i = 0; i < count; ++i)
    if (!isfinite(g[i])) return WRT_E_ARG;
  return WRT_OK;
}

/* --------------------------------------------------- host-buffer entry points --- */
static int run_once(wrt_ctx* c, const wrt_job_desc* d, const double* in, double* out,
                    wrt_stats* st) {
  if (!c || !in || !out) return WRT_E_ARG;
  wrt_job* j = NULL;
  RCHK(wrt_job_create(c, d, &j));
  int rc = wrt_job_load(j, in);
  if (rc == WRT_OK) rc = wrt_job_run(j, st);
  if (rc == WRT_OK) rc = wrt_job_fetch(j, out);
  wrt_job_destroy(j);
  return rc;
}

int wrt_radon_grid(wrt_ctx* c, const double* vol, int n, const double* phi, int nphi,
                   const double* theta, int ntheta, const double* shift, int nshift,
                   wrt_prec prec, double* out, wrt_stats* st) {
  wrt_job_desc d = {WRT_OP_RADON, prec, n, phi, nphi, theta, NULL, ntheta, shift, nshift, NULL, 0};
  return run_once(c, &d, vol, out, st);
}

int wrt_ray_grid(wrt_ctx* c, const double* vol, int n, const double* z, int nz,
                 const double* shift, int nshift, const double* phi, int nphi, wrt_prec prec,
                 double* out, wrt_stats* st) {
  wrt_job_desc d = {WRT_OP_RAY, prec, n, phi, nphi, NULL, NULL, 0, shift, nshift, z, nz};
  return run_once(c, &d, vol, out, st);
}

int wrt_radon_reduction(wrt_ctx* c, const double* rays, int nz, const double* shift,
                        int nshift, const double* phi, int nphi, const double* theta,
                        int ntheta, double* out, wrt_stats* st) {
  wrt_job_desc d = {WRT_OP_REDUCTION, WRT_F64, 0, phi, nphi, theta, NULL, ntheta, shift, nshift, NULL, nz};
  return run_once(c, &d, rays, out, st);
}

int wrt_radon_adjoint(wrt_ctx* c, const double* radon, const double* phi, int nphi,
                      const double* theta, const double* theta_w, int ntheta,
                      const double* shift, int nshift, int n, wrt_prec prec, double* out,
                      wrt_stats* st) {
  wrt_job_desc d = {WRT_OP_ADJOINT, prec, n, phi, nphi, theta, theta_w, ntheta, shift, nshift, NULL, 0};
  return run_once(c, &d, radon, out, st);
}

int wrt_radon_grid_T(wrt_ctx* c, const double* g, int n, const double* phi, int nphi,
                     const double* theta, int ntheta, const double* shift, int nshift,
                     wrt_prec prec, double* out, wrt_stats* st) {
  wrt_job_desc d = {WRT_OP_RADON_T, prec, n, phi, nphi, theta, NULL, ntheta, shift, nshift, NULL, 0};
  return run_once(c, &d, g, out, st);
}

int wrt_ray_grid_T(wrt_ctx* c, const double* g, int n, const double* z, int nz,
                   const double* shift, int nshift, const double* phi, int nphi, wrt_prec prec,
                   double* out, wrt_stats* st) {
  wrt_job_desc d = {WRT_OP_RAY_T, prec, n, phi, nphi, NULL, NULL, 0, shift, nshift, z, nz};
  return run_once(c, &d, g, out, st);
}

int wrt_radon_count_incube(wrt_ctx* c, int n, const double* phi, int nphi, const double* theta,
                           int ntheta, const double* shift, int nshift, double* count) {
  if (!c || !count || n < 1 || wrt_check_grid(phi, nphi, 1) || wrt_check_grid(theta, ntheta, 1) ||
      wrt_check_grid(shift, nshift, 1))
    return WRT_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  const size_t ndir = (size_t)nphi * ntheta;
  double* h = (double*)malloc(sizeof(double) * (9 * ndir + nshift));
  if (!h) return WRT_E_NOMEM;
  wrt_host_radon_dirs(n, phi, nphi, theta, ntheta, h);
  memcpy(h + 9 * ndir, shift, sizeof(double) * nshift);
  double* d = NULL;
  unsigned long long* tot = NULL;
  int rc = wrt_hip(hipMalloc((void**)&d, sizeof(double) * (9 * ndir + nshift)));
  if (rc == WRT_OK) rc = wrt_hip(hipMalloc((void**)&tot, sizeof(unsigned long long)));
  if (rc == WRT_OK) rc = wrt_hip(hipMemcpyAsync(d, h, sizeof(double) * (9 * ndir + nshift), hipMemcpyHostToDevice, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipMemsetAsync(tot, 0, sizeof(unsigned long long), c->stream));
  if (rc == WRT_OK) rc = wrtk_radon_incube(n, d, (int)ndir, d + 9 * ndir, nshift, tot, c->stream);
  unsigned long long v = 0;
  if (rc == WRT_OK) rc = wrt_hip(hipMemcpyAsync(&v, tot, sizeof(v), hipMemcpyDeviceToHost, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipStreamSynchronize(c->stream));
  *count = (double)v;
  hipFree(d);
  hipFree(tot);
  free(h);
  return rc;
}
