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
  *out = c;
  return WRT_OK;
}

void wrt_batch_release(wrt_ctx* ctx); /* wrt_batch.c */

void wrt_free(wrt_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  wrt_batch_release(c);
  hipStreamSynchronize(c->stream);
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
  const double delta = (1.0 / (n));
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

/* -------------------------------------------------------------------------- jobs --- */
static int desc_ok(const wrt_job_desc* d) {
  if (!d) return 0;
  if (d->prec != WRT_F64 && d->prec != WRT_F32 && d->prec != WRT_BF16) return 0;
  switch (d->op) {
    case WRT_OP_RADON:
    case WRT_OP_RADON_T:
      return d->n >= 1 && d->n <= 4096 && wrt_check_grid(d->phi, d->nphi, 1) == WRT_OK &&
             wrt_check_grid(d->theta, d->ntheta, 1) == WRT_OK &&
             wrt_check_grid(d->shift, d->nshift, 1) == WRT_OK;
    case WRT_OP_RAY:
    case WRT_OP_RAY_T:
      return d->n >= 2 && d->n <= 4096 && wrt_check_grid(d->phi, d->nphi, 1) == WRT_OK &&
             wrt_check_grid(d->z, d->nz, 1) == WRT_OK &&
             wrt_check_grid(d->shift, d->nshift, 1) == WRT_OK;
    case WRT_OP_REDUCTION:
      return d->prec == WRT_F64 && d->nz >= 3 && wrt_check_grid(d->phi, d->nphi, 1) == WRT_OK &&
             wrt_check_grid(d->theta, d->ntheta, 1) == WRT_OK &&
             wrt_check_grid(d->shift, d->nshift, 3) == WRT_OK;
    case WRT_OP_ADJOINT:
      return d->n >= 1 && d->n <= 4096 && wrt_check_grid(d->phi, d->nphi, 1) == WRT_OK &&
             wrt_check_grid(d->theta, d->ntheta, 1) == WRT_OK &&
             wrt_check_grid(d->theta_w, d->ntheta, 1) == WRT_OK &&
             wrt_check_grid(d->shift, d->nshift, 3) == WRT_OK;
    default: return 0;
  }
}

static size_t n3(const wrt_job_desc* d) { return (size_t)d->n * d->n * d->n; }
static size_t nradon(const wrt_job_desc* d) { return (size_t)d->nshift * d->nphi * d->ntheta; }
static size_t nrays(const wrt_job_desc* d) { return (size_t)d->nz * d->nshift * d->nphi; }

size_t wrt_job_input_count(const wrt_job_desc* d) {
  if (!d) return 0;
  switch (d->op) {
    case WRT_OP_RADON: case WRT_OP_RAY: return n3(d);
    case WRT_OP_REDUCTION: case WRT_OP_RAY_T: return nrays(d);
    case WRT_OP_ADJOINT: case WRT_OP_RADON_T: return nradon(d);
  }
  return 0;
}

size_t wrt_job_output_count(const wrt_job_desc* d) {
  if (!d) return 0;
  switch (d->op) {
    case WRT_OP_RADON: case WRT_OP_REDUCTION: return nradon(d);
    case WRT_OP_RAY: return nrays(d);
    case WRT_OP_ADJOINT: case WRT_OP_RADON_T: case WRT_OP_RAY_T: return n3(d);
  }
  return 0;
}

static size_t table_doubles(const wrt_job_desc* d) {
  size_t t = (size_t)d->nshift + 2 * (size_t)d->nshift; /* shift + spline factor */
  switch (d->op) {
    case WRT_OP_RADON: case WRT_OP_RADON_T: t += 9 * (size_t)d->nphi * d->ntheta; break;
    case WRT_OP_RAY: case WRT_OP_RAY_T: t += (size_t)d->nz + 4 * (size_t)d->nphi; break;
    case WRT_OP_REDUCTION:
      t += 2 * (size_t)d->ntheta + 2 * (size_t)d->nshift * d->nz * d->nphi;
      break;
    case WRT_OP_ADJOINT:
      t += 2 * (size_t)d->nphi + 3 * (size_t)d->ntheta + 2 * nradon(d);
      break;
  }
  return t;
}

/* fast-path hooks (csrc/k_fast.hip, host/wrt_fast.c) */
size_t wrt_fast_bytes(const wrt_job_desc* d);
int wrt_fast_create(wrt_job* j);
int wrt_fast_load(wrt_job* j, const double* in);
int wrt_fast_run(wrt_job* j, wrt_stats* st);
void wrt_fast_destroy(wrt_job* j);
int wrt_fast_supported(const wrt_job_desc* d);

static size_t elem_bytes(wrt_prec p) { return p == WRT_F64 ? 8 : 4; }

size_t wrt_job_bytes(const wrt_job_desc* d) {
  if (!desc_ok(d)) return 0;
  size_t b = table_doubles(d) * 8;
  if (d->prec == WRT_F64) {
    b += wrt_job_input_count(d) * 8 + wrt_job_output_count(d) * 8;
  } else {
    b += wrt_fast_bytes(d);
  }
  return b;
}

static double* dup_grid(const double* g, int n) {
  if (!g || n <= 0) return NULL;
  double* p = (double*)malloc(sizeof(double) * (size_t)n);
  if (p) memcpy(p, g, sizeof(double) * (size_t)n);
  return p;
}

void wrt_job_destroy(wrt_job* j) {
  if (!j) return;
  hipSetDevice(j->ctx->device);
  hipStreamSynchronize(j->ctx->stream);
  if (j->fast) wrt_fast_destroy(j);
  hipFree(j->d_in);
  hipFree(j->d_out);
  hipFree(j->d_tabs);
  hipFree(j->d_work);
  free(j->phi);
  free(j->theta);
  free(j->theta_w);
  free(j->shift);
  free(j->z);
  free(j);
}

int wrt_job_create(wrt_ctx* c, const wrt_job_desc* d, wrt_job** out) {
  if (!c || !out) return WRT_E_ARG;
  *out = NULL;
  if (!desc_ok(d)) return WRT_E_ARG;
  if (d->prec != WRT_F64 && !wrt_fast_supported(d)) return WRT_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  wrt_job* j = (wrt_job*)calloc(1, sizeof(wrt_job));
  if (!j) return WRT_E_NOMEM;
  j->ctx = c;
  j->d = *d;
  j->phi = dup_grid(d->phi, d->nphi);
  j->theta = dup_grid(d->theta, d->ntheta);
  j->theta_w = dup_grid(d->theta_w, d->ntheta);
  j->shift = dup_grid(d->shift, d->nshift);
  j->z = dup_grid(d->z, d->nz);
  j->d.phi = j->phi;
  j->d.theta = j->theta;
  j->d.theta_w = j->theta_w;
  j->d.shift = j->shift;
  j->d.z = j->z;
  j->in_count = wrt_job_input_count(d);
  j->out_count = wrt_job_output_count(d);

  /* tables, laid out in one allocation */
  const size_t nt = table_doubles(d);
  double* h = (double*)calloc(nt, sizeof(double));
  if (!h) { wrt_job_destroy(j); return WRT_E_NOMEM; }
  int rc = wrt_hip(hipMalloc((void**)&j->d_tabs, nt * sizeof(double)));
  if (rc != WRT_OK) { free(h); wrt_job_destroy(j); return WRT_E_NOMEM; }
  size_t off = 0;
  memcpy(h + off, j->shift, sizeof(double) * d->nshift);
  j->d_shift = j->d_tabs + off;
  off += d->nshift;
  j->d_fac = j->d_tabs + off;
  off += 2 * (size_t)d->nshift;
  switch (d->op) {
    case WRT_OP_RADON: case WRT_OP_RADON_T:
      wrt_host_radon_dirs(d->n, j->phi, d->nphi, j->theta, d->ntheta, h + off);
      j->d_dirs = j->d_tabs + off;
      off += 9 * (size_t)d->nphi * d->ntheta;
      break;
    case WRT_OP_RAY: case WRT_OP_RAY_T:
      memcpy(h + off, j->z, sizeof(double) * d->nz);
      j->d_z = j->d_tabs + off;
      off += d->nz;
      wrt_host_ray_trig(j->phi, d->nphi, h + off);
      j->d_dirs = j->d_tabs + off;
      off += 4 * (size_t)d->nphi;
      break;
    case WRT_OP_REDUCTION:
      for (int i = 0; i < d->ntheta; ++i) {
        h[off + 2 * i] = cos(j->theta[i]);
        h[off + 2 * i + 1] = sin(j->theta[i]);
      }
      j->d_ttrig = j->d_tabs + off;
      off += 2 * (size_t)d->ntheta;
      j->d_spl = j->d_tabs + off;
      off += 2 * (size_t)d->nshift * d->nz * d->nphi;
      break;
    case WRT_OP_ADJOINT: {
      for (int k = 0; k < d->nphi; ++k) {
        h[off + 2 * k] = cos(j->phi[k]);
        h[off + 2 * k + 1] = sin(j->phi[k]);
      }
      j->d_ptrig = j->d_tabs + off;
      off += 2 * (size_t)d->nphi;
      for (int i = 0; i < d->ntheta; ++i) {
        h[off + 2 * i] = cos(j->theta[i]);
        h[off + 2 * i + 1] = sin(j->theta[i]);
      }
      j->d_ttrig = j->d_tabs + off;
      off += 2 * (size_t)d->ntheta;
      const double dphi = 2 * M_PI / d->nphi; /* integral.c:11 */
      for (int i = 0; i < d->ntheta; ++i) h[off + i] = dphi * j->theta_w[i];
      j->d_wgt = j->d_tabs + off;
      off += d->ntheta;
      j->d_spl = j->d_tabs + off;
      off += 2 * nradon(d);
      break;
    }
  }
  rc = wrt_hip(hipMemcpyAsync(j->d_tabs, h, nt * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (rc == WRT_OK && (d->op == WRT_OP_REDUCTION || d->op == WRT_OP_ADJOINT))
    rc = wrtk_spline_factor(j->d_shift, d->nshift, j->d_fac, c->stream);
  if (rc == WRT_OK) rc = wrt_hip(hipStreamSynchronize(c->stream));
  free(h);
  if (rc != WRT_OK) { wrt_job_destroy(j); return rc; }

  if (d->prec == WRT_F64) {
    rc = wrt_hip(hipMalloc(&j->d_in, j->in_count * 8));
    if (rc == WRT_OK) rc = wrt_hip(hipMalloc(&j->d_out, j->out_count * 8));
    if (rc != WRT_OK) { wrt_job_destroy(j); return WRT_E_NOMEM; }
  } else {
    rc = wrt_fast_create(j);
    if (rc != WRT_OK) { wrt_job_destroy(j); return rc; }
  }
  j->bytes = wrt_job_bytes(d);
  *out = j;
  return WRT_OK;
}

int wrt_job_load(wrt_job* j, const double* in) {
  if (!j || !in) return WRT_E_ARG;
  HIPCHK(hipSetDevice(j->ctx->device));
  int rc;
  if (j->d.prec == WRT_F64) {
    rc = wrt_hip(hipMemcpyAsync(j->d_in, in, j->in_count * 8, hipMemcpyHostToDevice, j->ctx->stream));
    if (rc == WRT_OK) rc = wrt_hip(hipStreamSynchronize(j->ctx->stream));
  } else {
    rc = wrt_fast_load(j, in);
  }
  if (rc == WRT_OK) j->loaded = 1;
  return rc;
}

int wrt_job_load_synthetic(wrt_job* j, unsigned long long seed, int nell, double sigma) {
  if (!j) return WRT_E_ARG;
  double* h = (double*)malloc(j->in_count * sizeof(double));
  if (!h) return WRT_E_NOMEM;
  int rc;
  if (j->d.op == WRT_OP_RADON || j->d.op == WRT_OP_RAY)
    rc = wrt_phantom(j->d.n, j->d.op == WRT_OP_RADON, nell, seed, sigma, h);
  else
    rc = wrt_gaussian(j->in_count, seed, h);
  if (rc == WRT_OK) rc = wrt_job_load(j, h);
  free(h);
  return rc;
}

static double logical_samples(const wrt_job_desc* d) {
  switch (d->op) {
    case WRT_OP_RADON: case WRT_OP_RADON_T:
      return 16.0 * d->n * (double)d->n * (double)nradon(d);
    case WRT_OP_RAY: case WRT_OP_RAY_T:
      return (4.0 * d->n + 1.0) * (double)nrays(d);
    case WRT_OP_REDUCTION: return (4.0 * d->nz + 1.0) * (double)nradon(d);
    case WRT_OP_ADJOINT: return (double)d->nphi * d->ntheta * (double)n3(d);
  }
  return 0;
}

/* enqueue the job's device work between ev0 .. ev2 on the context stream (ev1 marks the
 * start of the dominant kernel) */
static int job_enqueue_f64(wrt_job* j) {
  wrt_ctx* c = j->ctx;
  const wrt_job_desc* d = &j->d;
  hipStream_t st = c->stream;
  int rc = WRT_OK;
  switch (d->op) {
    case WRT_OP_RADON:
      HIPCHK(hipEventRecord(c->ev1, st));
      rc = wrtk_radon_plain((const double*)j->d_in, d->n, j->d_dirs, d->nphi * d->ntheta,
                            j->d_shift, d->nshift, (double*)j->d_out, st);
      break;
    case WRT_OP_RAY:
      HIPCHK(hipEventRecord(c->ev1, st));
      rc = wrtk_ray_plain((const double*)j->d_in, d->n, j->d_z, d->nz, j->d_shift, d->nshift,
                          j->d_dirs, d->nphi, (double*)j->d_out, st);
      break;
    case WRT_OP_REDUCTION:
      rc = wrtk_spline_rows(j->d_shift, d->nshift, j->d_fac, (const double*)j->d_in, j->d_spl,
                            d->nz, d->nphi, (long long)d->nshift * d->nphi, d->nphi, st);
      if (rc != WRT_OK) return rc;
      HIPCHK(hipEventRecord(c->ev1, st));
      rc = wrtk_reduction(j->d_spl, d->nz, j->d_shift, d->nshift, d->nphi, j->d_ttrig,
                          d->ntheta, (double*)j->d_out, st);
      break;
    case WRT_OP_ADJOINT:
      rc = wrtk_spline_rows(j->d_shift, d->nshift, j->d_fac, (const double*)j->d_in, j->d_spl,
                            1, d->nphi * d->ntheta, 0, (long long)d->nphi * d->ntheta, st);
      if (rc != WRT_OK) return rc;
      HIPCHK(hipEventRecord(c->ev1, st));
      rc = wrtk_adjoint(j->d_spl, j->d_shift, d->nshift, j->d_ptrig, d->nphi, j->d_ttrig,
                        j->d_wgt, d->ntheta, d->n, (double*)j->d_out, st);
      break;
    case WRT_OP_RADON_T:
      rc = wrtk_fill_f64((double*)j->d_out, 0.0, j->out_count, st);
      if (rc != WRT_OK) return rc;
      HIPCHK(hipEventRecord(c->ev1, st));
      rc = wrtk_radon_T_plain((const double*)j->d_in, d->n, j->d_dirs, d->nphi * d->ntheta,
                              j->d_shift, d->nshift, (double*)j->d_out, st);
      break;
    case WRT_OP_RAY_T:
      rc = wrtk_fill_f64((double*)j->d_out, 0.0, j->out_count, st);
      if (rc != WRT_OK) return rc;
      HIPCHK(hipEventRecord(c->ev1, st));
      rc = wrtk_ray_T_plain((const double*)j->d_in, d->n, j->d_z, d->nz, j->d_shift,
                            d->nshift, j->d_dirs, d->nphi, (double*)j->d_out, st);
      break;
  }
  return rc;
}

int wrt_job_run(wrt_job* j, wrt_stats* st) {
  if (!j) return WRT_E_ARG;
  if (!j->loaded) return WRT_E_STATE;
  wrt_ctx* c = j->ctx;
  HIPCHK(hipSetDevice(c->device));
  if (j->d.prec != WRT_F64) return wrt_fast_run(j, st);
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  RCHK(job_enqueue_f64(j));
  HIPCHK(hipEventRecord(c->ev2, c->stream));
  HIPCHK(hipEventSynchronize(c->ev2));
  if (st) {
    float ms_all = 0, ms_k = 0;
    HIPCHK(hipEventElapsedTime(&ms_all, c->ev0, c->ev2));
    HIPCHK(hipEventElapsedTime(&ms_k, c->ev1, c->ev2));
    st->seconds = ms_all * 1e-3;
    st->kernel_seconds = ms_k * 1e-3;
    st->logical_samples = logical_samples(&j->d);
    st->alg_bytes = 0;
    st->outputs = (long long)j->out_count;
  }
  return WRT_OK;
}

int wrt_fast_fetch(wrt_job* j, double* out); /* wrt_fast.c */

int wrt_job_fetch(wrt_job* j, double* out) {
  if (!j || !out) return WRT_E_ARG;
  HIPCHK(hipSetDevice(j->ctx->device));
  if (j->d.prec != WRT_F64) return wrt_fast_fetch(j, out);
  HIPCHK(hipMemcpyAsync(out, j->d_out, j->out_count * 8, hipMemcpyDeviceToHost, j->ctx->stream));
  HIPCHK(hipStreamSynchronize(j->ctx->stream));
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
