/*
 * wrt_analytic.c -- host side of the analytic programs (radon_analytic, ray_analytic): test
 * functions as sums of terms, their support balls, the reference's example functions, and
 * the two FP64 grid entry points.
 */
#include <math.h>
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

int wrtk_radon_analytic(const wrt_fn_term* terms, const double* rb2, int nt, int ngrid,
                        const double* dirs, int ndir, const double* shift, int nshift,
                        double* out, hipStream_t st);
int wrtk_ray_analytic(const wrt_fn_term* terms, const double* rb2, int nt, int ngrid,
                      const double* z, int nz, const double* shift, int nshift,
                      const double* trig, int nphi, double* out, hipStream_t st);

static void set_term(wrt_fn_term* t, int kind, double cx, double cy, double cz) {
  memset(t, 0, sizeof(*t));
  t->kind = kind;
  t->c[0] = cx;
  t->c[1] = cy;
  t->c[2] = cz;
}

/* radon_analytic/examples-test-functions/exponential_bump.c:3-16 and sinusoidal_bump.c:22-35
 * (identical copies under ray_analytic/test-functions-examples/) */
int wrt_fn_example(const char* name, wrt_fn_term* terms, int max_terms) {
  if (!name || !terms) return WRT_E_ARG;
  if (!strcmp(name, "exponential_bump")) {
    if (max_terms < 1) return WRT_E_ARG;
    set_term(terms, WRT_FN_EXP_BUMP, 0.2, 0.0, 0.0);
    terms[0].p[0] = 0.5;  /* the support's squared radius */
    terms[0].p[1] = 1e-2; /* eps */
    return 1;
  }
  if (!strcmp(name, "sinusoidal_bump")) {
    /* three balls: centre, amplitude, inner and outer radius */
    static const double B[3][6] = {{0.0, 0.2, 0.0, 1.0, 0.4, 0.6},
                                   {-0.5, 0.0, 0.0, 0.6, 0.3, 0.45},
                                   {0.0, 0.2, 0.5, 0.3, 0.4, 0.6}};
    if (max_terms < 3) return WRT_E_ARG;
    for (int k = 0; k < 3; ++k) {
      set_term(terms + k, WRT_FN_SINE_BUMP, B[k][0], B[k][1], B[k][2]);
      terms[k].p[0] = B[k][3];
      terms[k].p[1] = B[k][4];
      terms[k].p[2] = B[k][5];
    }
    return 3;
  }
  return WRT_E_ARG;
}

int wrt_fn_phantom(int nell, unsigned long long seed, wrt_fn_term* terms) {
  if (nell < 0 || nell > WRT_FN_MAX_TERMS || (nell > 0 && !terms)) return WRT_E_ARG;
  double* P = (double*)malloc(sizeof(double) * 16 * (size_t)(nell ? nell : 1));
  if (!P) return WRT_E_NOMEM;
  const int rc = wrt_phantom_params(nell, seed, P);
  for (int e = 0; rc == WRT_OK && e < nell; ++e) {
    const double* p = P + 16 * e;
    set_term(terms + e, WRT_FN_ELLIPSOID, p[0], p[1], p[2]);
    terms[e].p[0] = p[15];
    for (int k = 0; k < 3; ++k) terms[e].p[1 + k] = p[3 + k];
    for (int k = 0; k < 9; ++k) terms[e].p[4 + k] = p[6 + k];
  }
  free(P);
  return rc;
}

int wrt_analytic_zslice_grid(int ngrid, double* z) {
  if (ngrid < 2 || !z) return WRT_E_ARG;
  for (int i = 0; i < ngrid; ++i) z[i] = (-1.0) + i * 2.0 / (ngrid - 1);
  return WRT_OK;
}

/* Squared radius of a ball outside which the term is exactly 0, with a small relative
 * margin; WRT_E_ARG for a malformed term. */
static int term_ball(const wrt_fn_term* t, double* rb2) {
  for (int k = 0; k < 3; ++k)
    if (!isfinite(t->c[k]) || fabs(t->c[k]) > 1e6) return WRT_E_ARG;
  for (int k = 0; k < 13; ++k)
    if (!isfinite(t->p[k])) return WRT_E_ARG;
  double r2;
  switch (t->kind) {
    case WRT_FN_EXP_BUMP:
      /* non-zero only where |x-c|^2 < r2max - eps */
      r2 = t->p[0] - t->p[1];
      if (r2 <= 0.0) r2 = 0.0;
      break;
    case WRT_FN_SINE_BUMP: {
      /* 0 for r > rout, and also at r = rout; amplitude for r < rin */
      const double r = fmax(0.0, fmax(t->p[1], t->p[2]));
      r2 = r * r;
      break;
    }
    case WRT_FN_ELLIPSOID: {
      /* |x - c| = |R^T (x - c)| <= max semi-axis for an orthogonal R */
      const double* R = t->p + 4;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          const double d = R[i] * R[j] + R[3 + i] * R[3 + j] + R[6 + i] * R[6 + j];
          if (fabs(d - (i == j ? 1.0 : 0.0)) > 1e-9) return WRT_E_ARG;
        }
      if (t->p[1] <= 0.0 || t->p[2] <= 0.0 || t->p[3] <= 0.0) return WRT_E_ARG;
      const double a = fmax(t->p[1], fmax(t->p[2], t->p[3]));
      r2 = a * a;
      break;
    }
    default: return WRT_E_ARG;
  }
  /* the margin covers the rounding of |x - c|^2 and an R orthogonal to 1e-9 */
  *rb2 = r2 * (1.0 + 1e-6) + 1e-12;
  return WRT_OK;
}

/* device copies of the terms and their balls, in one allocation */
static int upload_terms(const wrt_fn_term* terms, int nterms, void** dbuf, hipStream_t st) {
  const size_t tb = sizeof(wrt_fn_term) * (size_t)(nterms ? nterms : 1);
  char* h = (char*)calloc(1, tb + 8 * (size_t)(nterms ? nterms : 1));
  if (!h) return WRT_E_NOMEM;
  int rc = WRT_OK;
  for (int k = 0; k < nterms && rc == WRT_OK; ++k) rc = term_ball(terms + k, (double*)(h + tb) + k);
  if (rc == WRT_OK) {
    memcpy(h, terms, sizeof(wrt_fn_term) * (size_t)nterms);
    rc = wrt_hip(hipMalloc(dbuf, tb + 8 * (size_t)(nterms ? nterms : 1)));
    if (rc == WRT_OK) rc = wrt_hip(hipMemcpyAsync(*dbuf, h, tb + 8 * (size_t)(nterms ? nterms : 1),
                                                  hipMemcpyHostToDevice, st));
    if (rc == WRT_OK) rc = wrt_hip(hipStreamSynchronize(st));
  }
  free(h);
  return rc;
}

static int check_terms(const wrt_fn_term* terms, int nterms) {
  if (nterms < 0 || nterms > WRT_FN_MAX_TERMS || (nterms > 0 && !terms)) return WRT_E_ARG;
  return WRT_OK;
}

int wrt_radon_analytic(wrt_ctx* c, const wrt_fn_term* terms, int nterms, int ngrid,
                       const double* phi, int nphi, const double* theta, int ntheta,
                       const double* shift, int nshift, double* out, wrt_stats* st) {
  if (!c || !out || ngrid < 2 || ngrid > (1 << 20) || check_terms(terms, nterms) ||
      wrt_check_grid(phi, nphi, 1) || wrt_check_grid(theta, ntheta, 1) ||
      wrt_check_grid(shift, nshift, 1))
    return WRT_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  const size_t ndir = (size_t)nphi * ntheta, nplanes = ndir * nshift;
  double* h = (double*)malloc(sizeof(double) * (9 * ndir + nshift));
  if (!h) return WRT_E_NOMEM;
  wrt_host_plane_dirs((2.0 / (ngrid - 1)) / 2.0, phi, nphi, theta, ntheta, h);
  memcpy(h + 9 * ndir, shift, sizeof(double) * nshift);
  void* dterms = NULL;
  double *d = NULL, *dout = NULL;
  int rc = upload_terms(terms, nterms, &dterms, c->stream);
  if (rc == WRT_OK) rc = wrt_hip(hipMalloc((void**)&d, sizeof(double) * (9 * ndir + nshift)));
  if (rc == WRT_OK) rc = wrt_hip(hipMalloc((void**)&dout, sizeof(double) * nplanes));
  if (rc == WRT_OK)
    rc = wrt_hip(hipMemcpyAsync(d, h, sizeof(double) * (9 * ndir + nshift), hipMemcpyHostToDevice, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipEventRecord(c->ev0, c->stream));
  if (rc == WRT_OK)
    rc = wrtk_radon_analytic((const wrt_fn_term*)dterms,
                             (const double*)((char*)dterms + sizeof(wrt_fn_term) * (nterms ? nterms : 1)),
                             nterms, ngrid, d, (int)ndir, d + 9 * ndir, nshift, dout, c->stream);
  if (rc == WRT_OK) rc = wrt_hip(hipEventRecord(c->ev1, c->stream));
  if (rc == WRT_OK)
    rc = wrt_hip(hipMemcpyAsync(out, dout, sizeof(double) * nplanes, hipMemcpyDeviceToHost, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipStreamSynchronize(c->stream));
  if (rc == WRT_OK && st) {
    float ms = 0;
    rc = wrt_hip(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    st->seconds = st->kernel_seconds = ms * 1e-3;
    st->logical_samples = 16.0 * ngrid * (double)ngrid * (double)nplanes;
    st->alg_bytes = 0;
    st->outputs = (long long)nplanes;
  }
  hipFree(dterms);
  hipFree(d);
  hipFree(dout);
  free(h);
  return rc;
}

int wrt_ray_analytic(wrt_ctx* c, const wrt_fn_term* terms, int nterms, int ngrid,
                     const double* z, int nz, const double* shift, int nshift,
                     const double* phi, int nphi, double* out, wrt_stats* st) {
  if (!c || !out || ngrid < 2 || ngrid > (1 << 20) || check_terms(terms, nterms) ||
      wrt_check_grid(z, nz, 1) || wrt_check_grid(shift, nshift, 1) || wrt_check_grid(phi, nphi, 1))
    return WRT_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  const size_t nrays = (size_t)nz * nshift * nphi;
  const size_t nt = (size_t)nz + nshift + 4 * (size_t)nphi;
  double* h = (double*)malloc(sizeof(double) * nt);
  if (!h) return WRT_E_NOMEM;
  memcpy(h, z, sizeof(double) * nz);
  memcpy(h + nz, shift, sizeof(double) * nshift);
  wrt_host_ray_trig(phi, nphi, h + nz + nshift);
  void* dterms = NULL;
  double *d = NULL, *dout = NULL;
  int rc = upload_terms(terms, nterms, &dterms, c->stream);
  if (rc == WRT_OK) rc = wrt_hip(hipMalloc((void**)&d, sizeof(double) * nt));
  if (rc == WRT_OK) rc = wrt_hip(hipMalloc((void**)&dout, sizeof(double) * nrays));
  if (rc == WRT_OK) rc = wrt_hip(hipMemcpyAsync(d, h, sizeof(double) * nt, hipMemcpyHostToDevice, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipEventRecord(c->ev0, c->stream));
  if (rc == WRT_OK)
    rc = wrtk_ray_analytic((const wrt_fn_term*)dterms,
                           (const double*)((char*)dterms + sizeof(wrt_fn_term) * (nterms ? nterms : 1)),
                           nterms, ngrid, d, nz, d + nz, nshift, d + nz + nshift, nphi, dout, c->stream);
  if (rc == WRT_OK) rc = wrt_hip(hipEventRecord(c->ev1, c->stream));
  if (rc == WRT_OK)
    rc = wrt_hip(hipMemcpyAsync(out, dout, sizeof(double) * nrays, hipMemcpyDeviceToHost, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipStreamSynchronize(c->stream));
  if (rc == WRT_OK && st) {
    float ms = 0;
    rc = wrt_hip(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    st->seconds = st->kernel_seconds = ms * 1e-3;
    st->logical_samples = (4.0 * ngrid + 1.0) * (double)nrays;
    st->alg_bytes = 0;
    st->outputs = (long long)nrays;
  }
  hipFree(dterms);
  hipFree(d);
  hipFree(dout);
  free(h);
  return rc;
}
