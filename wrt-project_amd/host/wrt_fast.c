/*
 * wrt_fast.c -- host side of the fast (FP32 / BF16-brick) forms.
 *
 * RADON: the FP64 input is uploaded once, cut into 32^3 bricks (FP32 or BF16) on the
 * device, all-zero bricks are dropped, and the remaining bricks x direction groups form the
 * task list of the brick kernel (interior bricks and bricks touching the cube's faces are
 * separate launches).  The kernel accumulates each plane in FP64; the result is rounded to
 * FP32 on the device, which is the output of the form.
 */
#include <stdlib.h>
#include <math.h>
#include <string.h>

#include "../csrc/wrt_fast.h"
#include "wrt_internal.h"

#define HIPCHK(x)                            \
  do {                                       \
    if ((x) != hipSuccess) return WRT_E_HIP; \
  } while (0)

typedef struct {
  int nb;
  size_t nbricks;
  void* bricks;
  int* flags;       /* [nbricks] non-zero flags, then the support radius (8 bytes) */
  double rho2;      /* squared support radius of the loaded volume */
  int* tasks;       /* device, interior then boundary */
  int n_interior, n_boundary, dgroup;
  double* acc;      /* [nshift][ndir] */
  float* out;       /* [nshift][ndir] */
  double* staging;  /* n^3 FP64 input */
} radon_fast;

int wrt_fast_supported(const wrt_job_desc* d) {
  return d && d->op == WRT_OP_RADON && (d->prec == WRT_F32 || d->prec == WRT_BF16);
}

static int nb_of(int n) { return (n + WRT_BRICK - 1) / WRT_BRICK; }

size_t wrt_fast_bytes(const wrt_job_desc* d) {
  if (!wrt_fast_supported(d)) return 0;
  const size_t nb = (size_t)nb_of(d->n);
  const size_t nbr = nb * nb * nb;
  const size_t ndir = (size_t)d->nphi * d->ntheta;
  const size_t ngroups = (ndir + 7) / 8;
  const size_t eb = d->prec == WRT_BF16 ? 2 : 4;
  return nbr * WRT_BRICK * WRT_BRICK * WRT_BRICK * eb + nbr * 4 + 2 * nbr * ngroups * 4 +
         (size_t)d->nshift * ndir * 12 + (size_t)d->n * d->n * d->n * 8;
}

void wrt_fast_destroy(wrt_job* j) {
  radon_fast* f = (radon_fast*)j->fast;
  if (!f) return;
  hipFree(f->bricks);
  hipFree(f->flags);
  hipFree(f->tasks);
  hipFree(f->acc);
  hipFree(f->out);
  hipFree(f->staging);
  free(f);
  j->fast = NULL;
}

int wrt_fast_create(wrt_job* j) {
  const wrt_job_desc* d = &j->d;
  radon_fast* f = (radon_fast*)calloc(1, sizeof(radon_fast));
  if (!f) return WRT_E_NOMEM;
  j->fast = f;
  f->nb = nb_of(d->n);
  f->nbricks = (size_t)f->nb * f->nb * f->nb;
  const size_t eb = d->prec == WRT_BF16 ? 2 : 4;
  const size_t ndir = (size_t)d->nphi * d->ntheta;
  const size_t nout = (size_t)d->nshift * ndir;
  const size_t ngroups = (ndir + 7) / 8;
  if (hipMalloc(&f->bricks, f->nbricks * WRT_BRICK * WRT_BRICK * WRT_BRICK * eb) != hipSuccess ||
      hipMalloc((void**)&f->flags, f->nbricks * sizeof(int) + 16) != hipSuccess ||
      hipMalloc((void**)&f->tasks, 2 * f->nbricks * ngroups * sizeof(int)) != hipSuccess ||
      hipMalloc((void**)&f->acc, nout * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&f->out, nout * sizeof(float)) != hipSuccess ||
      hipMalloc((void**)&f->staging, (size_t)d->n * d->n * d->n * sizeof(double)) != hipSuccess) {
    wrt_fast_destroy(j);
    return WRT_E_NOMEM;
  }
  return WRT_OK;
}

int wrt_fast_load(wrt_job* j, const double* in) {
  radon_fast* f = (radon_fast*)j->fast;
  const wrt_job_desc* d = &j->d;
  hipStream_t st = j->ctx->stream;
  const size_t n3 = (size_t)d->n * d->n * d->n;
  HIPCHK(hipMemcpyAsync(f->staging, in, n3 * sizeof(double), hipMemcpyHostToDevice, st));
  /* flags, then (8-byte aligned) the squared support radius as ordered bits */
  const size_t r2off = (f->nbricks * sizeof(int) + 7) & ~(size_t)7;
  unsigned long long* d_r2 = (unsigned long long*)((char*)f->flags + r2off);
  HIPCHK(hipMemsetAsync(f->flags, 0, r2off + 8, st));
  int rc = wrtk_brickify(f->staging, d->n, f->nb, d->prec == WRT_BF16, f->bricks, f->flags, d_r2, st);
  if (rc != WRT_OK) return rc;
  int* flags = (int*)malloc(r2off + 8);
  if (!flags) return WRT_E_NOMEM;
  if (hipMemcpyAsync(flags, f->flags, r2off + 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    free(flags);
    return WRT_E_HIP;
  }
  {
    /* a sample farther than (farthest non-zero cell centre + half a cell diagonal) lies in a
     * zero cell; the small relative margin covers the rounding of the sample coordinates */
    double r2;
    memcpy(&r2, (char*)flags + r2off, sizeof(double));
    const double rho = sqrt(r2) + sqrt(3.0) / d->n;
    f->rho2 = rho * rho * (1.0 + 1e-12) + 1e-12;
  }
  const int ndir = d->nphi * d->ntheta;
  int nonempty = 0;
  for (size_t b = 0; b < f->nbricks; ++b) nonempty += flags[b] != 0;
  /* enough tasks to fill 256 CUs several times, at least 16 directions (one per wave) */
  long long want = (long long)ndir * (nonempty ? nonempty : 1) / 2048;
  int dg = (int)(want < 16 ? 16 : (want > 512 ? 512 : want));
  dg = (dg + 15) & ~15;
  const int ngroups = (ndir + dg - 1) / dg;
  f->dgroup = dg;
  int* t = (int*)malloc(2 * (f->nbricks * ngroups + 1) * sizeof(int));
  if (!t) { free(flags); return WRT_E_NOMEM; }
  int k = 0;
  for (int pass = 0; pass < 2; ++pass) {
    for (int g = 0; g < ngroups; ++g)
      for (size_t b = 0; b < f->nbricks; ++b) {
        if (!flags[b]) continue;
        const int bx = (int)(b / ((size_t)f->nb * f->nb)), by = (int)((b / f->nb) % f->nb), bz = (int)(b % f->nb);
        const int bnd = bx == 0 || by == 0 || bz == 0 || (bx + 1) * WRT_BRICK >= d->n ||
                        (by + 1) * WRT_BRICK >= d->n || (bz + 1) * WRT_BRICK >= d->n;
        if (bnd != pass) continue;
        t[2 * k] = (int)b;
        t[2 * k + 1] = g;
        ++k;
      }
    if (pass == 0) f->n_interior = k;
  }
  f->n_boundary = k - f->n_interior;
  free(flags);
  rc = WRT_OK;
  if (k > 0 && (hipMemcpyAsync(f->tasks, t, 2 * (size_t)k * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess))
    rc = WRT_E_HIP;
  free(t);
  return rc;
}

/* enqueue between ev0 and ev2 (ev1 = start of the brick kernels) */
int wrt_fast_enqueue(wrt_job* j) {
  radon_fast* f = (radon_fast*)j->fast;
  const wrt_job_desc* d = &j->d;
  wrt_ctx* c = j->ctx;
  hipStream_t st = c->stream;
  const size_t ndir = (size_t)d->nphi * d->ntheta;
  const size_t nout = (size_t)d->nshift * ndir;
  int rc = wrtk_fill_f64(f->acc, 0.0, nout, st);
  if (rc != WRT_OK) return rc;
  HIPCHK(hipEventRecord(c->ev1, st));
  wrt_radon_fast_args a;
  a.bricks = f->bricks;
  a.n = d->n;
  a.nb = f->nb;
  a.bf16 = d->prec == WRT_BF16;
  a.dirs = j->d_dirs;
  a.ndir = (int)ndir;
  a.dgroup = f->dgroup;
  a.shift = j->d_shift;
  a.nshift = d->nshift;
  a.acc = f->acc;
  a.rho2 = f->rho2;
  a.tasks = f->tasks;
  a.ntasks = f->n_interior;
  rc = wrtk_radon_fast(&a, 0, st);
  if (rc != WRT_OK) return rc;
  a.tasks = f->tasks + 2 * (size_t)f->n_interior;
  a.ntasks = f->n_boundary;
  rc = wrtk_radon_fast(&a, 1, st);
  if (rc != WRT_OK) return rc;
  HIPCHK(hipEventRecord(c->ev3, st));
  return wrtk_acc_to_f32(f->acc, f->out, nout, st);
}

int wrt_fast_run(wrt_job* j, wrt_stats* s) {
  wrt_ctx* c = j->ctx;
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  int rc = wrt_fast_enqueue(j);
  if (rc != WRT_OK) return rc;
  HIPCHK(hipEventRecord(c->ev2, c->stream));
  HIPCHK(hipEventSynchronize(c->ev2));
  if (s) {
    float ms_all = 0, ms_k = 0;
    HIPCHK(hipEventElapsedTime(&ms_all, c->ev0, c->ev2));
    HIPCHK(hipEventElapsedTime(&ms_k, c->ev1, c->ev3));
    const wrt_job_desc* d = &j->d;
    s->seconds = ms_all * 1e-3;
    s->kernel_seconds = ms_k * 1e-3;
    s->logical_samples = 16.0 * d->n * (double)d->n * (double)d->nshift * d->nphi * d->ntheta;
    s->alg_bytes = 0;
    s->outputs = (long long)d->nshift * d->nphi * d->ntheta;
  }
  return WRT_OK;
}

int wrt_fast_fetch(wrt_job* j, double* out) {
  radon_fast* f = (radon_fast*)j->fast;
  const wrt_job_desc* d = &j->d;
  const size_t nout = (size_t)d->nshift * d->nphi * d->ntheta;
  float* h = (float*)malloc(nout * sizeof(float));
  if (!h) return WRT_E_NOMEM;
  int rc = WRT_OK;
  if (hipMemcpyAsync(h, f->out, nout * sizeof(float), hipMemcpyDeviceToHost, j->ctx->stream) != hipSuccess ||
      hipStreamSynchronize(j->ctx->stream) != hipSuccess)
    rc = WRT_E_HIP;
  if (rc == WRT_OK)
    for (size_t i = 0; i < nout; ++i) out[i] = (double)h[i];
  free(h);
  return rc;
}
