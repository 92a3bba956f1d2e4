/*
 * wrt_job.c -- device-resident jobs: one operator on fixed grids, with its own input,
 * output, tables, HIP stream (default priority) and timing events.  A job is loaded once
 * and can be launched many times; launches of different jobs overlap on the device, which
 * is what the batch runner and the serve loop build on.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "wrt_internal.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define HIPCHK(x)                            \
  do {                                       \
    if ((x) != hipSuccess) return WRT_E_HIP; \
  } while (0)

/* fast-path hooks (host/wrt_fast.c) */
size_t wrt_fast_bytes(const wrt_job_desc* d);
int wrt_fast_supported(const wrt_job_desc* d);
int wrt_fast_create(wrt_job* j);
int wrt_fast_load(wrt_job* j, const double* in);
int wrt_fast_enqueue(wrt_job* j);
int wrt_fast_fetch(wrt_job* j, double* out);
void wrt_fast_destroy(wrt_job* j);

/* The reduction and the backprojection spline their data over the shift grid, which GSL's
 * gsl_interp_init (called through gsl_spline_init at radon_reduction/slicing.c:45-64 and
 * radon_inversion/backprojection/C/src/interp.c:17-25) rejects unless strictly increasing;
 * the plane and line integrals accept shifts in any order, as plane_integral and
 * ray_transform take any s. */
static int increasing(const double* s, int n) {
  for (int i = 1; i < n; ++i)
    if (!(s[i - 1] < s[i])) return 0;
  return 1;
}

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
             wrt_check_grid(d->shift, d->nshift, 3) == WRT_OK && increasing(d->shift, d->nshift);
    case WRT_OP_ADJOINT:
      return d->n >= 1 && d->n <= 4096 && wrt_check_grid(d->phi, d->nphi, 1) == WRT_OK &&
             wrt_check_grid(d->theta, d->ntheta, 1) == WRT_OK &&
             wrt_check_grid(d->theta_w, d->ntheta, 1) == WRT_OK &&
             wrt_check_grid(d->shift, d->nshift, 3) == WRT_OK && increasing(d->shift, d->nshift);
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

size_t wrt_job_bytes(const wrt_job_desc* d) {
  if (!desc_ok(d)) return 0;
  if (d->prec != WRT_F64 && !wrt_fast_supported(d)) return 0;
  size_t b = table_doubles(d) * 8;
  if (d->prec == WRT_F64)
    b += wrt_job_input_count(d) * 8 + wrt_job_output_count(d) * 8;
  else
    b += wrt_fast_bytes(d);
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
  if (j->st) hipStreamSynchronize(j->st);
  /* a captured batch graph that names this job's buffers must never be replayed */
  wrt_batch_forget(j->ctx, j);
  if (j->fast) wrt_fast_destroy(j);
  hipFree(j->d_in);
  hipFree(j->d_out);
  hipFree(j->d_tabs);
  hipFree(j->d_work);
  if (j->e0) hipEventDestroy(j->e0);
  if (j->e1) hipEventDestroy(j->e1);
  if (j->e2) hipEventDestroy(j->e2);
  if (j->e3) hipEventDestroy(j->e3);
  if (j->st) hipStreamDestroy(j->st);
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
  if (hipStreamCreateWithFlags(&j->st, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&j->e0) != hipSuccess || hipEventCreate(&j->e1) != hipSuccess ||
      hipEventCreate(&j->e2) != hipSuccess || hipEventCreate(&j->e3) != hipSuccess) {
    wrt_job_destroy(j);
    return WRT_E_HIP;
  }
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

  /* small tables, laid out in one allocation */
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
      const double dphi = 2 * M_PI / d->nphi; /* radon_inversion/.../integral.c:11 */
      for (int i = 0; i < d->ntheta; ++i) h[off + i] = dphi * j->theta_w[i];
      j->d_wgt = j->d_tabs + off;
      off += d->ntheta;
      j->d_spl = j->d_tabs + off;
      off += 2 * nradon(d);
      break;
    }
  }
  rc = wrt_hip(hipMemcpyAsync(j->d_tabs, h, nt * sizeof(double), hipMemcpyHostToDevice, j->st));
  if (rc == WRT_OK && (d->op == WRT_OP_REDUCTION || d->op == WRT_OP_ADJOINT))
    rc = wrtk_spline_factor(j->d_shift, d->nshift, j->d_fac, j->st);
  if (rc == WRT_OK) rc = wrt_hip(hipStreamSynchronize(j->st));
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
  if (j->pending) return WRT_E_STATE;
  HIPCHK(hipSetDevice(j->ctx->device));
  int rc;
  if (j->d.prec == WRT_F64) {
    rc = wrt_hip(hipMemcpyAsync(j->d_in, in, j->in_count * 8, hipMemcpyHostToDevice, j->st));
    if (rc == WRT_OK) rc = wrt_hip(hipStreamSynchronize(j->st));
  } else {
    rc = wrt_fast_load(j, in);
  }
  if (rc == WRT_OK) {
    j->loaded = 1;
    j->gen = ++j->ctx->load_seq;
  }
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

/* the plain FP64 forms: e1 / e3 bracket the dominant kernel */
static int job_enqueue_f64(wrt_job* j) {
  const wrt_job_desc* d = &j->d;
  hipStream_t st = j->st;
  int rc = WRT_OK;
  switch (d->op) {
    case WRT_OP_RADON:
      HIPCHK(hipEventRecord(j->e1, st));
      rc = wrtk_radon_plain((const double*)j->d_in, d->n, j->d_dirs, d->nphi * d->ntheta,
                            j->d_shift, d->nshift, (double*)j->d_out, st);
      break;
    case WRT_OP_RAY:
      HIPCHK(hipEventRecord(j->e1, st));
      rc = wrtk_ray_plain((const double*)j->d_in, d->n, j->d_z, d->nz, j->d_shift, d->nshift,
                          j->d_dirs, d->nphi, (double*)j->d_out, st);
      break;
    case WRT_OP_REDUCTION:
      rc = wrtk_spline_rows(j->d_shift, d->nshift, j->d_fac, (const double*)j->d_in, j->d_spl,
                            d->nz, d->nphi, (long long)d->nshift * d->nphi, d->nphi, st);
      if (rc != WRT_OK) return rc;
      HIPCHK(hipEventRecord(j->e1, st));
      rc = wrtk_reduction(j->d_spl, d->nz, j->d_shift, d->nshift, j->d_shift, d->nshift,
                          d->nphi, j->d_ttrig, d->ntheta, (double*)j->d_out, st);
      break;
    case WRT_OP_ADJOINT:
      rc = wrtk_spline_rows(j->d_shift, d->nshift, j->d_fac, (const double*)j->d_in, j->d_spl,
                            1, d->nphi * d->ntheta, 0, (long long)d->nphi * d->ntheta, st);
      if (rc != WRT_OK) return rc;
      HIPCHK(hipEventRecord(j->e1, st));
      rc = wrtk_adjoint(j->d_spl, j->d_shift, d->nshift, j->d_ptrig, d->nphi, j->d_ttrig,
                        j->d_wgt, d->ntheta, d->n, (double*)j->d_out, st);
      break;
    case WRT_OP_RADON_T:
      rc = wrtk_fill_f64((double*)j->d_out, 0.0, j->out_count, st);
      if (rc != WRT_OK) return rc;
      HIPCHK(hipEventRecord(j->e1, st));
      rc = wrtk_radon_T_plain((const double*)j->d_in, d->n, j->d_dirs, d->nphi * d->ntheta,
                              j->d_shift, d->nshift, (double*)j->d_out, st);
      break;
    case WRT_OP_RAY_T:
      rc = wrtk_fill_f64((double*)j->d_out, 0.0, j->out_count, st);
      if (rc != WRT_OK) return rc;
      HIPCHK(hipEventRecord(j->e1, st));
      rc = wrtk_ray_T_plain((const double*)j->d_in, d->n, j->d_z, d->nz, j->d_shift,
                            d->nshift, j->d_dirs, d->nphi, (double*)j->d_out, st);
      break;
  }
  if (rc != WRT_OK) return rc;
  HIPCHK(hipEventRecord(j->e3, st));
  return WRT_OK;
}

int wrt_job_launch(wrt_job* j) {
  if (!j) return WRT_E_ARG;
  if (!j->loaded || j->pending) return WRT_E_STATE;
  HIPCHK(hipSetDevice(j->ctx->device));
  HIPCHK(hipEventRecord(j->e0, j->st));
  const int rc = j->d.prec == WRT_F64 ? job_enqueue_f64(j) : wrt_fast_enqueue(j);
  if (rc != WRT_OK) return rc;
  HIPCHK(hipEventRecord(j->e2, j->st));
  j->pending = 1;
  return WRT_OK;
}

static int fill_stats(wrt_job* j, wrt_stats* st) {
  if (!st) return WRT_OK;
  float ms_all = 0, ms_k = 0;
  HIPCHK(hipEventElapsedTime(&ms_all, j->e0, j->e2));
  HIPCHK(hipEventElapsedTime(&ms_k, j->e1, j->e3));
  st->seconds = ms_all * 1e-3;
  st->kernel_seconds = ms_k * 1e-3;
  st->logical_samples = logical_samples(&j->d);
  st->alg_bytes = 0;
  st->outputs = (long long)j->out_count;
  return WRT_OK;
}

int wrt_job_poll(wrt_job* j, wrt_stats* st) {
  if (!j) return WRT_E_ARG;
  if (!j->pending) return WRT_E_STATE;
  const hipError_t e = hipEventQuery(j->e2);
  if (e == hipErrorNotReady) return 1;
  if (e != hipSuccess) return WRT_E_HIP;
  j->pending = 0;
  return fill_stats(j, st);
}

int wrt_job_wait(wrt_job* j, wrt_stats* st) {
  if (!j) return WRT_E_ARG;
  if (!j->pending) return WRT_E_STATE;
  HIPCHK(hipEventSynchronize(j->e2));
  j->pending = 0;
  return fill_stats(j, st);
}

int wrt_job_run(wrt_job* j, wrt_stats* st) {
  const int rc = wrt_job_launch(j);
  if (rc != WRT_OK) return rc;
  return wrt_job_wait(j, st);
}

int wrt_job_tuning(const wrt_job* j, int* band_rows, int* dgroup) {
  return wrt_fast_tuning(j, band_rows, dgroup);
}

int wrt_job_fetch(wrt_job* j, double* out) {
  if (!j || !out) return WRT_E_ARG;
  if (j->pending) return WRT_E_STATE;
  HIPCHK(hipSetDevice(j->ctx->device));
  if (j->d.prec != WRT_F64) return wrt_fast_fetch(j, out);
  HIPCHK(hipMemcpyAsync(out, j->d_out, j->out_count * 8, hipMemcpyDeviceToHost, j->st));
  HIPCHK(hipStreamSynchronize(j->st));
  return WRT_OK;
}
