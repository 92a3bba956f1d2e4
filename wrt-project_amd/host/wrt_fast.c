/*
 * wrt_fast.c -- host side of the fast (FP32 / BF16-brick) forms.
 *
 * RADON: the FP64 input is uploaded once, cut into 32^3 bricks (FP32 or BF16) on the
 * device, all-zero bricks are dropped, and the remaining bricks x direction groups form the
 * task list of the brick kernel (interior bricks and bricks touching the cube's faces are
 * separate launches).  The kernel accumulates each plane in FP64; the result is rounded to
 * FP32 on the device, which is the output of the form.
 *
 * RAY (FP32): the FP64 input is converted on the device to FP32 in [z][x][y] order; each
 * slice's z index and fraction are decided on the host as the reference decides them.
 *
 * ADJOINT (FP32): the Radon rows are uploaded in FP64 and splined in FP64 as in the plain
 * form; each interval's cubic is then rounded to FP32 (y, b, c, d), directions become float4
 * (unit normal, weight), and one thread per voxel sums the directions with Kahan compensation.
 * The interval of s is found arithmetically, so the shift grid must be uniform.
 *
 * RADON_T (FP32): the exact transpose of the plane integrals.  The plane values are rounded
 * to FP32; every (brick, direction group) task scatters its sections into an LDS brick
 * (index decisions of the forward) and stores it to the group's slot; the slots are summed
 * per voxel in FP64 and scaled by delta^2.  Every brick of the cube takes part.
 *
 * RAY_T (FP32): the exact transpose of the FP32 line integrals: per slice, bilinear weights
 * scattered into an LDS plane, added to an FP64 volume accumulator at the slice's two
 * z-nodes, rounded to FP32.
 */
#include <stdlib.h>
#include <float.h>
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
  int* flags;       /* [nbricks] non-zero voxel counts, then the support radius (8 bytes) */
  double rho2;      /* squared support radius of the loaded volume */
  int* tasks;       /* device, interior then boundary */
  int n_interior, n_boundary, dgroup;
  int band_rows;    /* rows per band of the walk (4, 8 or 16) */
  int tuned;        /* band_rows and dgroup settled for this load */
  long long* acc;   /* [nshift][ndir] plane sums, fixed point in units of qstep */
  double qstep;     /* wrt_plane_qstep of the loaded volume */
  float* out;       /* [nshift][ndir] */
  double* staging;  /* n^3 FP64 input */
  void* sorted;     /* ascending shifts, then their rows (sorted_shifts) */
  /* host copy of the non-empty bricks, heaviest first (batch planning) */
  int *ne_brick, *ne_count, n_ne;
  unsigned char* ne_bnd;
  hipStream_t side; /* the face-brick launch (side_init) */
  hipEvent_t fork, join;
} radon_fast;

/* ADJOINT in FP32: per-interval FP32 cubics of the FP64 spline, float4 direction table */
typedef struct {
  double* staging; /* FP64 Radon input [s][phi][theta] */
  void* cub;       /* [ndir][nshift-1] float4 (y, b, c, d) */
  void* dirs;      /* [ndir] float4 (n.x, n.y, n.z, (2pi/nphi) w_theta) */
  double* trig;    /* [ndir][4] sin t, cos p, sin p, cos t: exact cut decisions near |s| = 1 */
  float* out;      /* n^3 [z][x][y] */
  int* order;      /* the voxel tiles in launch order (wrt_adj_tile_order) */
  double s0, ds;
} adjoint_fast;

/* the FP32 adjoint evaluates the spline interval arithmetically: the shift grid must be
 * uniform (every grid the reference programs build is) */
static int uniform_grid(const double* s, int n, double* s0, double* ds) {
  if (n < 3) return 0;
  const double d = (s[n - 1] - s[0]) / (n - 1);
  for (int i = 0; i < n; ++i)
    if (fabs(s[i] - (s[0] + i * d)) > 1e-12 * (1.0 + fabs(s[i]))) return 0;
  *s0 = s[0];
  *ds = d;
  return d > 0;
}

/* RAY in FP32: FP32 volume in [z][x][y] order, slice decisions made on the host */
typedef struct {
  double* staging; /* n^3 FP64 input */
  float* vol32;    /* n^3 FP32 [z][x][y] */
  int* zi;         /* [nz] slice index (-1: |z| > 1) */
  double* zc;      /* [nz] slice fraction */
  float* out;      /* [nz][ns][nphi] */
} ray_fast;

static size_t ray_fast_bytes(const wrt_job_desc* d) {
  const size_t n3 = (size_t)d->n * d->n * d->n;
  return n3 * 12 + (size_t)d->nz * 12 + (size_t)d->nz * d->nshift * d->nphi * 4;
}

static void ray_fast_destroy(wrt_job* j) {
  ray_fast* f = (ray_fast*)j->fast;
  hipFree(f->staging);
  hipFree(f->vol32);
  hipFree(f->zi);
  hipFree(f->zc);
  hipFree(f->out);
  free(f);
  j->fast = NULL;
}

/* a slice's index and fraction exactly as cube_trilinear_interp decides them for the
 * constant z of every point of the slice (ray_grid/src/interp.c:25-47) */
static int slice_decisions(const wrt_job_desc* d, int* d_zi, double* d_zc, hipStream_t st) {
  int* zi = (int*)malloc((size_t)d->nz * sizeof(int));
  double* zc = (double*)malloc((size_t)d->nz * sizeof(double));
  if (!zi || !zc) {
    free(zi);
    free(zc);
    return WRT_E_NOMEM;
  }
  const double dg = 2.0 / (d->n - 1);
  for (int l = 0; l < d->nz; ++l) {
    const double z = d->z[l];
    if (fabs(z) > 1) {
      zi[l] = -1;
      zc[l] = 0.0;
      continue;
    }
    int iz = (int)((z + 1.0) / dg);
    double cz = fmod(z + 1.0, dg) / dg;
    if (iz == d->n - 1) {
      iz -= 1;
      cz = 1.0;
    }
    zi[l] = iz;
    zc[l] = cz;
  }
  int rc = WRT_OK;
  if (hipMemcpyAsync(d_zi, zi, (size_t)d->nz * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(d_zc, zc, (size_t)d->nz * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    rc = WRT_E_HIP;
  free(zi);
  free(zc);
  return rc;
}

static int ray_fast_create(wrt_job* j) {
  const wrt_job_desc* d = &j->d;
  ray_fast* f = (ray_fast*)calloc(1, sizeof(ray_fast));
  if (!f) return WRT_E_NOMEM;
  j->fast = f;
  const size_t n3 = (size_t)d->n * d->n * d->n;
  if (hipMalloc((void**)&f->staging, n3 * 8) != hipSuccess ||
      hipMalloc((void**)&f->vol32, n3 * 4) != hipSuccess ||
      hipMalloc((void**)&f->zi, (size_t)d->nz * sizeof(int)) != hipSuccess ||
      hipMalloc((void**)&f->zc, (size_t)d->nz * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&f->out, (size_t)d->nz * d->nshift * d->nphi * 4) != hipSuccess)
    return WRT_E_NOMEM;
  return slice_decisions(d, f->zi, f->zc, j->st);
}

static int ray_fast_load(wrt_job* j, const double* in) {
  ray_fast* f = (ray_fast*)j->fast;
  const size_t n3 = (size_t)j->d.n * j->d.n * j->d.n;
  HIPCHK(hipMemcpyAsync(f->staging, in, n3 * 8, hipMemcpyHostToDevice, j->st));
  const int rc = wrtk_vol_zxy_f32(f->staging, j->d.n, f->vol32, j->st);
  if (rc != WRT_OK) return rc;
  HIPCHK(hipStreamSynchronize(j->st));
  return WRT_OK;
}

static int ray_fast_enqueue(wrt_job* j) {
  ray_fast* f = (ray_fast*)j->fast;
  const wrt_job_desc* d = &j->d;
  HIPCHK(hipEventRecord(j->e1, j->st));
  const int rc = wrtk_ray_f32(f->vol32, d->n, f->zi, f->zc, d->nz, j->d_shift, d->nshift,
                              j->d_dirs, d->nphi, f->out, j->st);
  if (rc != WRT_OK) return rc;
  HIPCHK(hipEventRecord(j->e3, j->st));
  return WRT_OK;
}

int wrt_fast_supported(const wrt_job_desc* d) {
  if (!d) return 0;
  if (d->op == WRT_OP_RAY || d->op == WRT_OP_RAY_T) return d->prec == WRT_F32 && d->nz <= 65535;
  if (d->op == WRT_OP_RADON_T) return d->prec == WRT_F32 && d->n <= 4096;
  /* the brick kernel packs a row's first and last lattice column (|i2| <= 2n) in 16 bits */
  if (d->op == WRT_OP_RADON) return (d->prec == WRT_F32 || d->prec == WRT_BF16) && d->n <= 4096;
  if (d->op == WRT_OP_ADJOINT) {
    double s0, ds;
    return d->prec == WRT_F32 && d->shift && uniform_grid(d->shift, d->nshift, &s0, &ds);
  }
  return 0;
}

static size_t adjoint_fast_bytes(const wrt_job_desc* d) {
  const size_t ndir = (size_t)d->nphi * d->ntheta;
  return (size_t)d->nshift * ndir * 8 + ndir * (size_t)(d->nshift - 1) * 16 + ndir * 48 +
         (size_t)d->n * d->n * d->n * 4 + (size_t)wrt_adj_tiles(d->n, WRT_ADJ_VPT) * 4;
}

static void adjoint_fast_destroy(wrt_job* j) {
  adjoint_fast* f = (adjoint_fast*)j->fast;
  hipFree(f->staging);
  hipFree(f->cub);
  hipFree(f->dirs);
  hipFree(f->trig);
  hipFree(f->out);
  hipFree(f->order);
  free(f);
  j->fast = NULL;
}

static int adjoint_fast_create(wrt_job* j) {
  const wrt_job_desc* d = &j->d;
  adjoint_fast* f = (adjoint_fast*)calloc(1, sizeof(adjoint_fast));
  if (!f) return WRT_E_NOMEM;
  j->fast = f;
  uniform_grid(d->shift, d->nshift, &f->s0, &f->ds);
  const size_t ndir = (size_t)d->nphi * d->ntheta;
  if (hipMalloc((void**)&f->staging, (size_t)d->nshift * ndir * 8) != hipSuccess ||
      hipMalloc(&f->cub, ndir * (size_t)(d->nshift - 1) * 16) != hipSuccess ||
      hipMalloc(&f->dirs, ndir * 16) != hipSuccess ||
      hipMalloc((void**)&f->trig, ndir * 32) != hipSuccess ||
      hipMalloc((void**)&f->out, (size_t)d->n * d->n * d->n * 4) != hipSuccess)
    return WRT_E_NOMEM;
  const size_t ntiles = (size_t)wrt_adj_tiles(d->n, WRT_ADJ_VPT);
  if (hipMalloc((void**)&f->order, ntiles * 4) != hipSuccess) return WRT_E_NOMEM;
  /* unit normals (sin t cos p, sin t sin p, cos t) and weights (2pi/nphi) w_t, host libm */
  float* h = (float*)malloc(ndir * 16);
  double* tr = (double*)malloc(ndir * 32);
  int* ord = (int*)malloc(ntiles * 4);
  if (!h || !tr || !ord) {
    free(h);
    free(tr);
    free(ord);
    return WRT_E_NOMEM;
  }
  const double dphi = 2 * M_PI / d->nphi;
  for (int ip = 0; ip < d->nphi; ++ip)
    for (int it = 0; it < d->ntheta; ++it) {
      const size_t k = (size_t)ip * d->ntheta + it;
      double* t = tr + 4 * k;
      t[0] = sin(d->theta[it]);
      t[1] = cos(d->phi[ip]);
      t[2] = sin(d->phi[ip]);
      t[3] = cos(d->theta[it]);
      float* v = h + 4 * k;
      v[0] = (float)(t[0] * t[1]);
      v[1] = (float)(t[0] * t[2]);
      v[2] = (float)t[3];
      v[3] = (float)(dphi * d->theta_w[it]);
    }
  wrt_adj_tile_order(d->n, WRT_ADJ_VPT, ord);
  int rc = WRT_OK;
  if (hipMemcpyAsync(f->dirs, h, ndir * 16, hipMemcpyHostToDevice, j->st) != hipSuccess ||
      hipMemcpyAsync(f->trig, tr, ndir * 32, hipMemcpyHostToDevice, j->st) != hipSuccess ||
      hipMemcpyAsync(f->order, ord, ntiles * 4, hipMemcpyHostToDevice, j->st) != hipSuccess ||
      hipStreamSynchronize(j->st) != hipSuccess)
    rc = WRT_E_HIP;
  free(h);
  free(tr);
  free(ord);
  return rc;
}

static int adjoint_fast_load(wrt_job* j, const double* in) {
  adjoint_fast* f = (adjoint_fast*)j->fast;
  const wrt_job_desc* d = &j->d;
  const size_t ndir = (size_t)d->nphi * d->ntheta;
  HIPCHK(hipMemcpyAsync(f->staging, in, (size_t)d->nshift * ndir * 8, hipMemcpyHostToDevice, j->st));
  int rc = wrtk_spline_rows(j->d_shift, d->nshift, j->d_fac, f->staging, j->d_spl, 1, (int)ndir, 0,
                            (long long)ndir, j->st);
  if (rc == WRT_OK) rc = wrtk_adjoint_cubics(j->d_shift, d->nshift, j->d_spl, f->cub, (long long)ndir, j->st);
  if (rc == WRT_OK) HIPCHK(hipStreamSynchronize(j->st));
  return rc;
}

static int adjoint_fast_enqueue(wrt_job* j) {
  adjoint_fast* f = (adjoint_fast*)j->fast;
  const wrt_job_desc* d = &j->d;
  HIPCHK(hipEventRecord(j->e1, j->st));
  const int rc = wrtk_adjoint_f32(f->dirs, f->trig, d->nphi * d->ntheta, f->cub, d->nshift, f->s0, f->ds,
                                  d->n, f->order, f->out, j->st);
  if (rc != WRT_OK) return rc;
  HIPCHK(hipEventRecord(j->e3, j->st));
  return WRT_OK;
}

static int fetch_f32(wrt_job* j, const float* src, size_t count, double* out) {
  float* h = (float*)malloc(count * sizeof(float));
  if (!h) return WRT_E_NOMEM;
  int rc = WRT_OK;
  if (hipMemcpyAsync(h, src, count * sizeof(float), hipMemcpyDeviceToHost, j->st) != hipSuccess ||
      hipStreamSynchronize(j->st) != hipSuccess)
    rc = WRT_E_HIP;
  if (rc == WRT_OK)
    for (size_t i = 0; i < count; ++i) out[i] = (double)h[i];
  free(h);
  return rc;
}

static int nb_of(int n) { return (n + WRT_BRICK - 1) / WRT_BRICK; }

/* The brick kernel bisects the shift grid for the shifts that cut a brick, so it reads an
 * ascending copy together with each position's row in the caller's grid: plane_integral
 * (radon_grid/src/integral.c:26) takes any s, so a grid may come in any order and repeat
 * values.  dev receives nshift doubles (ascending) then nshift ints (rows); hs, when not
 * NULL, the ascending values. */
typedef struct {
  double v;
  int i;
} shift_key;

static int by_shift(const void* a, const void* b) {
  const shift_key *x = (const shift_key*)a, *y = (const shift_key*)b;
  if (x->v != y->v) return x->v < y->v ? -1 : 1;
  return (x->i > y->i) - (x->i < y->i);
}

static size_t sorted_shift_bytes(int nshift) { return (size_t)nshift * (sizeof(double) + sizeof(int)); }

static int sorted_shifts(const wrt_job_desc* d, void* dev, double* hs, hipStream_t st) {
  const int ns = d->nshift;
  shift_key* k = (shift_key*)malloc(sizeof(shift_key) * (size_t)ns);
  char* h = (char*)malloc(sorted_shift_bytes(ns));
  if (!k || !h) {
    free(k);
    free(h);
    return WRT_E_NOMEM;
  }
  for (int j = 0; j < ns; ++j) {
    k[j].v = d->shift[j];
    k[j].i = j;
  }
  qsort(k, (size_t)ns, sizeof(shift_key), by_shift);
  double* v = (double*)h;
  int* row = (int*)(h + (size_t)ns * sizeof(double));
  for (int j = 0; j < ns; ++j) {
    v[j] = k[j].v;
    row[j] = k[j].i;
    if (hs) hs[j] = k[j].v;
  }
  int rc = WRT_OK;
  if (hipMemcpyAsync(dev, h, sorted_shift_bytes(ns), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    rc = WRT_E_HIP;
  free(k);
  free(h);
  return rc;
}

static size_t radon_fast_bytes(const wrt_job_desc* d) {
  const size_t nb = (size_t)nb_of(d->n);
  const size_t nbr = nb * nb * nb;
  const size_t ndir = (size_t)d->nphi * d->ntheta;
  const size_t ngroups = (ndir + 7) / 8;
  const size_t eb = d->prec == WRT_BF16 ? 2 : 4;
  return nbr * WRT_BRICK * WRT_BRICK * WRT_BRICK * eb + nbr * 4 + 2 * nbr * ngroups * 4 +
         (size_t)d->nshift * ndir * 12 + (size_t)d->n * d->n * d->n * 8 +
         sorted_shift_bytes(d->nshift);
}

/* A side stream (default priority) with the events that fork it from and join it back into
 * a job's stream: a large job's face-brick launch runs there, concurrently with its interior
 * launch, so that neither launch's tail leaves the device idle.  Jobs below SIDE_MIN_N
 * keep one stream: their launches are short, and a second stream per job would make the
 * serve loop's concurrent jobs share the process's few hardware queues more often. */
#define SIDE_MIN_N 192
static int side_init(const wrt_job_desc* d, hipStream_t* side, hipEvent_t* fork,
                     hipEvent_t* join) {
  if (d->n < SIDE_MIN_N) return WRT_OK;
  if (hipStreamCreateWithFlags(side, hipStreamNonBlocking) != hipSuccess) return WRT_E_HIP;
  if (hipEventCreateWithFlags(fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(join, hipEventDisableTiming) != hipSuccess)
    return WRT_E_HIP;
  return WRT_OK;
}

static void side_free(hipStream_t side, hipEvent_t fork, hipEvent_t join) {
  if (side) {
    hipStreamSynchronize(side);
    hipStreamDestroy(side);
  }
  if (fork) hipEventDestroy(fork);
  if (join) hipEventDestroy(join);
}

static void radon_fast_destroy(wrt_job* j) {
  radon_fast* f = (radon_fast*)j->fast;
  if (!f) return;
  side_free(f->side, f->fork, f->join);
  hipFree(f->bricks);
  hipFree(f->flags);
  hipFree(f->tasks);
  hipFree(f->acc);
  hipFree(f->out);
  hipFree(f->staging);
  hipFree(f->sorted);
  free(f->ne_brick);
  free(f->ne_count);
  free(f->ne_bnd);
  free(f);
  j->fast = NULL;
}

static int radon_fast_create(wrt_job* j) {
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
      /* the counts, padding to 8 bytes, the support radius and the largest |value| */
      hipMalloc((void**)&f->flags, f->nbricks * sizeof(int) + 24) != hipSuccess ||
      hipMalloc((void**)&f->tasks, 2 * f->nbricks * ngroups * sizeof(int)) != hipSuccess ||
      hipMalloc((void**)&f->acc, nout * sizeof(long long)) != hipSuccess ||
      hipMalloc((void**)&f->out, nout * sizeof(float)) != hipSuccess ||
      hipMalloc((void**)&f->staging, (size_t)d->n * d->n * d->n * sizeof(double)) != hipSuccess ||
      hipMalloc(&f->sorted, sorted_shift_bytes(d->nshift)) != hipSuccess) {
    radon_fast_destroy(j);
    return WRT_E_NOMEM;
  }
  if (side_init(d, &f->side, &f->fork, &f->join) != WRT_OK) {
    radon_fast_destroy(j);
    return WRT_E_HIP;
  }
  const int rc = sorted_shifts(d, f->sorted, NULL, j->st);
  if (rc != WRT_OK) radon_fast_destroy(j);
  return rc;
}

/* The task list for direction groups of dg: every non-empty brick (heaviest first) x group,
 * interior bricks, then bricks touching the cube's faces; uploaded, the count kept. */
static int radon_fast_tasks(wrt_job* j, int dg) {
  radon_fast* f = (radon_fast*)j->fast;
  const wrt_job_desc* d = &j->d;
  const int ndir = d->nphi * d->ntheta;
  const int ngroups = (ndir + dg - 1) / dg;
  if (dg < 8 || (size_t)f->n_ne * ngroups > f->nbricks * (size_t)((ndir + 7) / 8)) return WRT_E_ARG;
  int* t = (int*)malloc(2 * ((size_t)f->n_ne * ngroups + 1) * sizeof(int));
  if (!t) return WRT_E_NOMEM;
  int k = 0;
  for (int pass = 0; pass < 2; ++pass) {
    for (int i = 0; i < f->n_ne; ++i) {
      if (f->ne_bnd[i] != pass) continue;
      for (int g = 0; g < ngroups; ++g) {
        t[2 * k] = f->ne_brick[i];
        t[2 * k + 1] = g;
        ++k;
      }
    }
    if (pass == 0) f->n_interior = k;
  }
  f->n_boundary = k - f->n_interior;
  f->dgroup = dg;
  int rc = WRT_OK;
  if (k > 0 && (hipMemcpyAsync(f->tasks, t, 2 * (size_t)k * sizeof(int), hipMemcpyHostToDevice,
                               j->st) != hipSuccess ||
                hipStreamSynchronize(j->st) != hipSuccess))
    rc = WRT_E_HIP;
  free(t);
  return rc;
}

static int radon_fast_load(wrt_job* j, const double* in) {
  radon_fast* f = (radon_fast*)j->fast;
  const wrt_job_desc* d = &j->d;
  hipStream_t st = j->st;
  const size_t n3 = (size_t)d->n * d->n * d->n;
  HIPCHK(hipMemcpyAsync(f->staging, in, n3 * sizeof(double), hipMemcpyHostToDevice, st));
  /* flags, then (8-byte aligned) the squared support radius and the largest |value|, as
   * ordered bits */
  const size_t r2off = (f->nbricks * sizeof(int) + 7) & ~(size_t)7;
  unsigned long long* d_r2 = (unsigned long long*)((char*)f->flags + r2off);
  HIPCHK(hipMemsetAsync(f->flags, 0, r2off + 16, st));
  int rc = wrtk_brickify(f->staging, d->n, f->nb, 0, f->nb, d->prec == WRT_BF16, f->bricks,
                         f->flags, d_r2, st);
  if (rc != WRT_OK) return rc;
  int* flags = (int*)malloc(r2off + 16);
  if (!flags) return WRT_E_NOMEM;
  if (hipMemcpyAsync(flags, f->flags, r2off + 16, hipMemcpyDeviceToHost, st) != hipSuccess ||
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
    unsigned long long vb;
    memcpy(&vb, (char*)flags + r2off + 8, sizeof(vb));
    const unsigned vb32 = (unsigned)vb;
    float vmax;
    memcpy(&vmax, &vb32, sizeof(vmax));
    if (!(vmax <= FLT_MAX)) { /* a value that is not finite in FP32 */
      free(flags);
      return WRT_E_ARG;
    }
    f->qstep = wrt_plane_qstep(vmax);
  }
  const int ndir = d->nphi * d->ntheta;
  int nonempty = 0;
  for (size_t b = 0; b < f->nbricks; ++b) nonempty += flags[b] != 0;
  /* enough tasks to fill 256 CUs several times, at least 16 directions (one per wave).
   * More directions per task even out the ends of the 16 waves' shares and stage fewer
   * bricks: 256^3 with 256 phi x 128 theta = 32768 directions and 257 shifts takes 0.728 s
   * per pass with groups of 1024, 0.6755 s with 2048 and 0.6695 s with 4096 (about 400
   * non-empty bricks: 3200 tasks of 4096 directions on 256 CUs).  So at least 8 groups per
   * brick: groups of at most max(512, ndir / 8) directions, and never above 4096 (the tuner
   * also times half the group).  Served jobs run in the persistent batch kernel, whose
   * planner sizes its own groups. */
  long long cap = ndir / 8;
  cap = cap < 512 ? 512 : (cap > 4096 ? 4096 : cap);
  long long want = (long long)ndir * (nonempty ? nonempty : 1) / 2048;
  int dg = (int)(want < 16 ? 16 : (want > cap ? cap : want));
  dg = (dg + 15) & ~15;
  /* bricks by falling non-zero count (the work of a task grows with it), so that the
   * longest tasks start first and the launch ends on short ones */
  int* order = (int*)malloc(f->nbricks * sizeof(int));
  if (!order) { free(flags); return WRT_E_NOMEM; }
  int nne = 0;
  for (size_t b = 0; b < f->nbricks; ++b)
    if (flags[b]) order[nne++] = (int)b;
  for (int i = 1; i < nne; ++i) { /* insertion sort: at most a few thousand bricks */
    const int v = order[i];
    int q = i - 1;
    while (q >= 0 && flags[order[q]] < flags[v]) { order[q + 1] = order[q]; --q; }
    order[q + 1] = v;
  }
  free(f->ne_brick);
  free(f->ne_count);
  free(f->ne_bnd);
  f->ne_brick = (int*)malloc(sizeof(int) * (size_t)(nne + 1));
  f->ne_count = (int*)malloc(sizeof(int) * (size_t)(nne + 1));
  f->ne_bnd = (unsigned char*)malloc((size_t)(nne + 1));
  f->n_ne = nne;
  if (!f->ne_brick || !f->ne_count || !f->ne_bnd) {
    free(order); free(flags);
    return WRT_E_NOMEM;
  }
  for (int i = 0; i < nne; ++i) {
    const int b = order[i];
    const int bx = b / (f->nb * f->nb), by = (b / f->nb) % f->nb, bz = b % f->nb;
    f->ne_brick[i] = b;
    f->ne_count[i] = flags[b];
    f->ne_bnd[i] = bx == 0 || by == 0 || bz == 0 || (bx + 1) * WRT_BRICK >= d->n ||
                   (by + 1) * WRT_BRICK >= d->n || (bz + 1) * WRT_BRICK >= d->n;
  }
  free(order);
  free(flags);
  f->band_rows = 8;
  f->tuned = 0;
  return radon_fast_tasks(j, dg);
}

/* one pass of the brick kernels on stream st (interior on st, faces on the side stream
 * when the job has one), then the FP32 rounding */
static int radon_fast_pass(wrt_job* j, hipStream_t st, int timed) {
  radon_fast* f = (radon_fast*)j->fast;
  const wrt_job_desc* d = &j->d;
  const size_t ndir = (size_t)d->nphi * d->ntheta;
  const size_t nout = (size_t)d->nshift * ndir;
  int rc = wrtk_fill_f64((double*)f->acc, 0.0, nout, st); /* zero bits: integer 0 */
  if (rc != WRT_OK) return rc;
  if (timed) HIPCHK(hipEventRecord(j->e1, st));
  wrt_radon_fast_args a;
  memset(&a, 0, sizeof(a));
  a.bricks = f->bricks;
  a.n = d->n;
  a.nb = f->nb;
  a.bf16 = d->prec == WRT_BF16;
  a.dirs = j->d_dirs;
  a.ndir = (int)ndir;
  a.dgroup = f->dgroup;
  a.shift = (const double*)f->sorted;
  a.sidx = (const int*)((const char*)f->sorted + (size_t)d->nshift * sizeof(double));
  a.nshift = d->nshift;
  a.acc = f->acc;
  a.qmul = (1.0 / d->n) * (1.0 / d->n) / f->qstep;
  a.rho2 = f->rho2;
  a.band_rows = f->band_rows;
  /* face bricks on the side stream, interior bricks on the job's stream, concurrently:
   * both add into the same 64-bit plane sums with integer atomics */
  hipStream_t sb = f->side ? f->side : st;
  if (f->side) {
    HIPCHK(hipEventRecord(f->fork, st));
    HIPCHK(hipStreamWaitEvent(f->side, f->fork, 0));
  }
  wrt_radon_fast_args b = a;
  b.tasks = f->tasks + 2 * (size_t)f->n_interior;
  b.ntasks = f->n_boundary;
  rc = wrtk_radon_fast(&b, 1, sb);
  if (rc != WRT_OK) return rc;
  if (f->side) HIPCHK(hipEventRecord(f->join, f->side));
  a.tasks = f->tasks;
  a.ntasks = f->n_interior;
  rc = wrtk_radon_fast(&a, 0, st);
  if (rc != WRT_OK) return rc;
  if (f->side) HIPCHK(hipStreamWaitEvent(st, f->join, 0));
  if (timed) HIPCHK(hipEventRecord(j->e3, st));
  return wrtk_fix_to_f32(f->acc, f->qstep, f->out, nout, st);
}

/* seconds of one pass with the current band shape and direction group (events of its own:
 * the job's bracket its launches) */
static int radon_fast_time(wrt_job* j, double* sec) {
  hipEvent_t a = NULL, b = NULL;
  int rc = WRT_OK;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess ||
      hipEventRecord(a, j->st) != hipSuccess)
    rc = WRT_E_HIP;
  if (rc == WRT_OK) rc = radon_fast_pass(j, j->st, 0);
  float ms = 0;
  if (rc == WRT_OK && (hipEventRecord(b, j->st) != hipSuccess ||
                       hipEventSynchronize(b) != hipSuccess ||
                       hipEventElapsedTime(&ms, a, b) != hipSuccess))
    rc = WRT_E_HIP;
  if (a) hipEventDestroy(a);
  if (b) hipEventDestroy(b);
  *sec = ms * 1e-3;
  return rc;
}

/* The autotuner of the per-job launch, run at a load's first launch: the band shape (8 x 8,
 * 4 x 16 or 16 x 4 rows x columns) timed at the load's direction group, then half and twice
 * that group at the best shape; the fastest pair is kept for the load and in the tuning cache
 * under the problem's shape (edge, directions, shifts, precision, non-empty bricks), where
 * later loads of that shape find it.  WRT_BAND_ROWS (4, 8, 16) and WRT_RADON_DGROUP (>= 16)
 * fix the shape and the group and skip the tuning; WRT_AUTOTUNE=0 keeps 8 x 8 and the load's
 * group.  A candidate replaces the incumbent only when it is faster by more than
 * WRT_TUNE_MARGIN, so that timing noise between near-equal shapes does not change the pick
 * from one run to the next (the first candidate, 8 x 8 at the load's group, is the default). */
#define WRT_TUNE_MARGIN 0.01
static int radon_fast_tune(wrt_job* j) {
  radon_fast* f = (radon_fast*)j->fast;
  const wrt_job_desc* d = &j->d;
  f->tuned = 1;
  const char* force = getenv("WRT_BAND_ROWS");
  const char* force_g = getenv("WRT_RADON_DGROUP");
  if ((force && *force) || (force_g && *force_g)) {
    if (force && *force) {
      const int br = atoi(force);
      if (br != 4 && br != 8 && br != 16) return WRT_E_ARG;
      f->band_rows = br;
    }
    if (force_g && *force_g) {
      const int g = atoi(force_g);
      if (g < 16 || g % 16) return WRT_E_ARG;
      return radon_fast_tasks(j, g);
    }
    return WRT_OK;
  }
  if (!wrt_tune_enabled() || f->n_ne == 0) return WRT_OK;
  const int dg0 = f->dgroup, ndir = d->nphi * d->ntheta;
  const long long key[7] = {0x72616e64LL /* its own signature space */, d->n, ndir, d->nshift,
                            d->prec, f->n_ne, dg0};
  const unsigned long long sig = wrt_tune_hash(key, 7);
  int br = 8, dg = dg0;
  if (wrt_tune_lookup(sig, &br, &dg) && (br == 4 || br == 8 || br == 16) && dg >= 16) {
    f->band_rows = br;
    return dg != f->dgroup ? radon_fast_tasks(j, dg) : WRT_OK;
  }
  static const int shapes[3] = {8, 4, 16};
  double best = 1e300, sec = 0;
  int best_br = 8, best_dg = dg0;
  for (int q = 0; q < 3; ++q) {
    f->band_rows = shapes[q];
    int rc = radon_fast_time(j, &sec); /* the first run of a variant loads its code ... */
    if (rc == WRT_OK) rc = radon_fast_time(j, &sec); /* ... so the second one is timed */
    if (rc != WRT_OK) return rc;
    if (sec < best * (1.0 - WRT_TUNE_MARGIN)) {
      best = sec;
      best_br = shapes[q];
    }
  }
  f->band_rows = best_br;
  const int cand[2] = {dg0 / 2, dg0 * 2};
  for (int q = 0; q < 2; ++q) {
    const int g = cand[q] & ~15;
    if (g < 16 || g > 4096 || g > ndir) continue;
    int rc = radon_fast_tasks(j, g);
    if (rc == WRT_OK) rc = radon_fast_time(j, &sec);
    if (rc != WRT_OK) return rc;
    if (sec < best * (1.0 - WRT_TUNE_MARGIN)) {
      best = sec;
      best_dg = g;
    }
  }
  if (f->dgroup != best_dg) {
    const int rc = radon_fast_tasks(j, best_dg);
    if (rc != WRT_OK) return rc;
  }
  wrt_tune_store(sig, best_br, best_dg, best);
  return WRT_OK;
}

/* enqueue on the job's stream; e1 / e3 bracket the brick kernels (the first launch of a
 * load first settles the band shape and direction group: radon_fast_tune) */
static int radon_fast_enqueue(wrt_job* j) {
  radon_fast* f = (radon_fast*)j->fast;
  if (!f->tuned) {
    const int rc = radon_fast_tune(j);
    if (rc != WRT_OK) return rc;
    HIPCHK(hipEventRecord(j->e0, j->st)); /* the launch's time leaves the tuning out */
  }
  return radon_fast_pass(j, j->st, 1);
}

/* ------------------------------------------------------------------ RADON_T --- */
typedef struct {
  int nb;
  size_t nbricks;
  int dgroup, ngroups;
  int n_interior, n_boundary;
  int qbits;       /* fixed point: |rint(g * gscale)| <= 2^qbits */
  double gscale;
  int* tasks;      /* device, interior then boundary */
  void* sorted;    /* ascending shifts, then their rows (sorted_shifts) */
  double* staging; /* FP64 plane values [nshift][ndir] */
  float* g32;      /* FP32 plane values */
  long long* part; /* [ngroups][nbricks][WRT_BRICK^3] fixed-point partials */
  float* out;      /* n^3 [x][y][z] */
  hipStream_t side; /* the face-brick launch (side_init) */
  hipEvent_t fork, join;
} radon_T_fast;

#define RT_PART_BUDGET ((size_t)1 << 30) /* bytes of group partials at most */

/* Bits of the transpose's fixed point.  One plane puts at most 16 lattice samples into one
 * cell: its section of the cell lies in a disk of radius sqrt(3) delta (half the cell's
 * diagonal), which holds at most 12 points of the square lattice of spacing delta.  The
 * planes of one direction that meet a cell have shifts within the cell's projection on the
 * normal, an interval of width at most sqrt(3) * 2/n; w counts the most shifts any such
 * window holds.  So no cell receives more than 16 w ndir samples, and with
 * |q| <= 2^qbits no total, over every group, exceeds 2^62. */
static int radon_T_qbits(const wrt_job_desc* d, const double* sorted) {
  const double width = sqrt(3.0) * 2.0 / d->n * (1.0 + 1e-9) + 1e-12;
  int w = 1;
  for (int lo = 0, hi = 0; lo < d->nshift; ++lo) { /* the shifts in ascending order */
    if (hi < lo) hi = lo;
    while (hi + 1 < d->nshift && sorted[hi + 1] - sorted[lo] <= width) ++hi;
    if (hi - lo + 1 > w) w = hi - lo + 1;
  }
  const double samples = 16.0 * w * ((double)d->nphi * d->ntheta);
  return 62 - (int)ceil(log2(samples));
}

/* groups: about 4096 tasks (two half bricks per brick), at least 16 directions per group,
 * partials within the budget */
static void radon_T_groups(const wrt_job_desc* d, size_t nbricks, int* ngroups, int* dgroup) {
  const size_t ndir = (size_t)d->nphi * d->ntheta;
  const size_t slot = nbricks * WRT_BRICK * WRT_BRICK * WRT_BRICK * sizeof(long long);
  size_t g = (4096 + 2 * nbricks - 1) / (2 * nbricks);
  const size_t by_dirs = ndir / 16 ? ndir / 16 : 1;
  const size_t by_mem = RT_PART_BUDGET / slot ? RT_PART_BUDGET / slot : 1;
  if (g > by_dirs) g = by_dirs;
  if (g > by_mem) g = by_mem;
  if (g < 1) g = 1;
  const size_t dg = (ndir + g - 1) / g;
  *dgroup = (int)dg;
  *ngroups = (int)((ndir + dg - 1) / dg);
}

static size_t radon_T_fast_bytes(const wrt_job_desc* d) {
  const size_t nb = (size_t)nb_of(d->n);
  const size_t nbr = nb * nb * nb;
  int ng, dg;
  radon_T_groups(d, nbr, &ng, &dg);
  const size_t nplanes = (size_t)d->nshift * d->nphi * d->ntheta;
  return nplanes * 12 + (size_t)ng * nbr * WRT_BRICK * WRT_BRICK * WRT_BRICK * 8 +
         4 * nbr * (size_t)ng * sizeof(int) + (size_t)d->n * d->n * d->n * 4 +
         sorted_shift_bytes(d->nshift);
}

static void radon_T_fast_destroy(wrt_job* j) {
  radon_T_fast* f = (radon_T_fast*)j->fast;
  if (!f) return;
  side_free(f->side, f->fork, f->join);
  hipFree(f->tasks);
  hipFree(f->sorted);
  hipFree(f->staging);
  hipFree(f->g32);
  hipFree(f->part);
  hipFree(f->out);
  free(f);
  j->fast = NULL;
}

static int radon_T_fast_create(wrt_job* j) {
  const wrt_job_desc* d = &j->d;
  radon_T_fast* f = (radon_T_fast*)calloc(1, sizeof(radon_T_fast));
  if (!f) return WRT_E_NOMEM;
  j->fast = f;
  f->nb = nb_of(d->n);
  f->nbricks = (size_t)f->nb * f->nb * f->nb;
  radon_T_groups(d, f->nbricks, &f->ngroups, &f->dgroup);
  const size_t nplanes = (size_t)d->nshift * d->nphi * d->ntheta;
  const size_t ntasks = 2 * f->nbricks * (size_t)f->ngroups;
  if (ntasks > 0x7fffffff) return WRT_E_ARG;
  if (hipMalloc(&f->sorted, sorted_shift_bytes(d->nshift)) != hipSuccess) return WRT_E_NOMEM;
  {
    double* hs = (double*)malloc(sizeof(double) * (size_t)d->nshift);
    if (!hs) return WRT_E_NOMEM;
    const int rc = sorted_shifts(d, f->sorted, hs, j->st);
    if (rc == WRT_OK) f->qbits = radon_T_qbits(d, hs);
    free(hs);
    if (rc != WRT_OK) return rc;
  }
  if (f->qbits < 24) return WRT_E_ARG; /* more samples per cell than 64 bits can hold */
  if (hipMalloc((void**)&f->tasks, 2 * ntasks * sizeof(int)) != hipSuccess ||
      hipMalloc((void**)&f->staging, nplanes * 8) != hipSuccess ||
      hipMalloc((void**)&f->g32, nplanes * 4) != hipSuccess ||
      hipMalloc((void**)&f->part, (size_t)f->ngroups * f->nbricks * WRT_BRICK * WRT_BRICK * WRT_BRICK * 8) != hipSuccess ||
      hipMalloc((void**)&f->out, (size_t)d->n * d->n * d->n * 4) != hipSuccess)
    return WRT_E_NOMEM;
  if (side_init(d, &f->side, &f->fork, &f->join) != WRT_OK) return WRT_E_HIP;
  /* bricks nearest the centre first (planes with |s| <= 1 cut them from every direction, so
   * their tasks are the longest), interior bricks then bricks on the cube's faces */
  int* order = (int*)malloc(f->nbricks * sizeof(int));
  double* dist = (double*)malloc(f->nbricks * sizeof(double));
  int* t = (int*)malloc(2 * ntasks * sizeof(int));
  if (!order || !dist || !t) {
    free(order);
    free(dist);
    free(t);
    return WRT_E_NOMEM;
  }
  const int nb = f->nb;
  for (size_t b = 0; b < f->nbricks; ++b) {
    const int bx = (int)(b / ((size_t)nb * nb)), by = (int)((b / nb) % nb), bz = (int)(b % nb);
    const double c[3] = {bx + 0.5, by + 0.5, bz + 0.5};
    double r2 = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double x = -1.0 + 2.0 * WRT_BRICK * c[k] / d->n;
      r2 += x * x;
    }
    order[b] = (int)b;
    dist[b] = r2;
  }
  for (size_t i = 1; i < f->nbricks; ++i) { /* insertion sort by distance */
    const int v = order[i];
    long q = (long)i - 1;
    while (q >= 0 && dist[order[q]] > dist[v]) { order[q + 1] = order[q]; --q; }
    order[q + 1] = v;
  }
  size_t k = 0;
  for (int pass = 0; pass < 2; ++pass) {
    for (size_t i = 0; i < f->nbricks; ++i) {
      const int b = order[i];
      const int bx = b / (nb * nb), by = (b / nb) % nb, bz = b % nb;
      for (int zh = 0; zh < 2; ++zh) {
        const int z0 = bz * WRT_BRICK + zh * (WRT_BRICK / 2), z1 = z0 + WRT_BRICK / 2;
        if (z0 >= d->n) continue; /* the half brick lies outside the cube */
        const int bnd = bx == 0 || by == 0 || z0 == 0 || (bx + 1) * WRT_BRICK >= d->n ||
                        (by + 1) * WRT_BRICK >= d->n || z1 >= d->n;
        if (bnd != pass) continue;
        for (int g = 0; g < f->ngroups; ++g) {
          t[2 * k] = 2 * b + zh;
          t[2 * k + 1] = g;
          ++k;
        }
      }
    }
    if (pass == 0) f->n_interior = (int)k;
  }
  f->n_boundary = (int)k - f->n_interior;
  int rc = WRT_OK;
  if (hipMemcpyAsync(f->tasks, t, 2 * k * sizeof(int), hipMemcpyHostToDevice, j->st) != hipSuccess ||
      hipStreamSynchronize(j->st) != hipSuccess)
    rc = WRT_E_HIP;
  free(order);
  free(dist);
  free(t);
  return rc;
}

static int radon_T_fast_load(wrt_job* j, const double* in) {
  radon_T_fast* f = (radon_T_fast*)j->fast;
  const size_t nplanes = (size_t)j->d.nshift * j->d.nphi * j->d.ntheta;
  /* fixed-point scale from the largest |value| as FP32 holds it: |rint(g * gscale)| is at
   * most 2^qbits */
  float gmax = 0.f;
  int bad = 0;
#pragma omp parallel for reduction(max : gmax) reduction(| : bad)
  for (long long i = 0; i < (long long)nplanes; ++i) {
    const float v = fabsf((float)in[i]);
    if (!isfinite(v)) bad = 1;
    else if (v > gmax) gmax = v;
  }
  if (bad) return WRT_E_ARG;
  f->gscale = gmax > 0.f ? ldexp(1.0, f->qbits) / (double)gmax : 1.0;
  HIPCHK(hipMemcpyAsync(f->staging, in, nplanes * 8, hipMemcpyHostToDevice, j->st));
  const int rc = wrtk_f64_to_f32(f->staging, f->g32, nplanes, j->st);
  if (rc != WRT_OK) return rc;
  HIPCHK(hipStreamSynchronize(j->st));
  return WRT_OK;
}

static int radon_T_fast_enqueue(wrt_job* j) {
  radon_T_fast* f = (radon_T_fast*)j->fast;
  const wrt_job_desc* d = &j->d;
  hipStream_t st = j->st;
  wrt_radon_fast_args a;
  memset(&a, 0, sizeof(a));
  a.n = d->n;
  a.nb = f->nb;
  a.dirs = j->d_dirs;
  a.ndir = d->nphi * d->ntheta;
  a.dgroup = f->dgroup;
  a.shift = (const double*)f->sorted;
  a.sidx = (const int*)((const char*)f->sorted + (size_t)d->nshift * sizeof(double));
  a.nshift = d->nshift;
  a.rho2 = 4.0; /* no support clipping: every cell of the cube receives */
  a.g = f->g32;
  a.gscale = f->gscale;
  a.part = f->part;
  HIPCHK(hipEventRecord(j->e1, st));
  /* face half bricks on the side stream, concurrently with the interior ones: every task
   * writes its own partial slot */
  hipStream_t sb = f->side ? f->side : st;
  if (f->side) {
    HIPCHK(hipEventRecord(f->fork, st));
    HIPCHK(hipStreamWaitEvent(f->side, f->fork, 0));
  }
  wrt_radon_fast_args b = a;
  b.tasks = f->tasks + 2 * (size_t)f->n_interior;
  b.ntasks = f->n_boundary;
  int rc = wrtk_radon_T_fast(&b, 1, sb);
  if (rc != WRT_OK) return rc;
  if (f->side) HIPCHK(hipEventRecord(f->join, f->side));
  a.tasks = f->tasks;
  a.ntasks = f->n_interior;
  rc = wrtk_radon_T_fast(&a, 0, st);
  if (rc != WRT_OK) return rc;
  if (f->side) HIPCHK(hipStreamWaitEvent(st, f->join, 0));
  HIPCHK(hipEventRecord(j->e3, st));
  const double delta = (1.0 / (d->n));
  return wrtk_radon_T_reduce(f->part, f->ngroups, d->n, f->nb, delta * delta / f->gscale,
                             f->out, st);
}

/* -------------------------------------------------------------------- RAY_T --- */
typedef struct {
  double* staging; /* FP64 ray values [z][s][phi] */
  float* g32;
  int* zi;         /* [nz] slice index (-1: |z| > 1) */
  double* zc;      /* [nz] slice fraction */
  double* acc;     /* n^3 FP64 [x][y][z] */
  float* out;      /* n^3 FP32 */
} ray_T_fast;

static size_t ray_T_fast_bytes(const wrt_job_desc* d) {
  const size_t n3 = (size_t)d->n * d->n * d->n;
  return (size_t)d->nz * d->nshift * d->nphi * 12 + n3 * 12 + (size_t)d->nz * 12;
}

static void ray_T_fast_destroy(wrt_job* j) {
  ray_T_fast* f = (ray_T_fast*)j->fast;
  if (!f) return;
  hipFree(f->staging);
  hipFree(f->g32);
  hipFree(f->zi);
  hipFree(f->zc);
  hipFree(f->acc);
  hipFree(f->out);
  free(f);
  j->fast = NULL;
}

static int ray_T_fast_create(wrt_job* j) {
  const wrt_job_desc* d = &j->d;
  ray_T_fast* f = (ray_T_fast*)calloc(1, sizeof(ray_T_fast));
  if (!f) return WRT_E_NOMEM;
  j->fast = f;
  const size_t n3 = (size_t)d->n * d->n * d->n;
  const size_t nr = (size_t)d->nz * d->nshift * d->nphi;
  if (hipMalloc((void**)&f->staging, nr * 8) != hipSuccess ||
      hipMalloc((void**)&f->g32, nr * 4) != hipSuccess ||
      hipMalloc((void**)&f->zi, (size_t)d->nz * sizeof(int)) != hipSuccess ||
      hipMalloc((void**)&f->zc, (size_t)d->nz * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&f->acc, n3 * 8) != hipSuccess ||
      hipMalloc((void**)&f->out, n3 * 4) != hipSuccess)
    return WRT_E_NOMEM;
  return slice_decisions(d, f->zi, f->zc, j->st);
}

static int ray_T_fast_load(wrt_job* j, const double* in) {
  ray_T_fast* f = (ray_T_fast*)j->fast;
  const size_t nr = (size_t)j->d.nz * j->d.nshift * j->d.nphi;
  HIPCHK(hipMemcpyAsync(f->staging, in, nr * 8, hipMemcpyHostToDevice, j->st));
  const int rc = wrtk_f64_to_f32(f->staging, f->g32, nr, j->st);
  if (rc != WRT_OK) return rc;
  HIPCHK(hipStreamSynchronize(j->st));
  return WRT_OK;
}

static int ray_T_fast_enqueue(wrt_job* j) {
  ray_T_fast* f = (ray_T_fast*)j->fast;
  const wrt_job_desc* d = &j->d;
  const size_t n3 = (size_t)d->n * d->n * d->n;
  int rc = wrtk_fill_f64(f->acc, 0.0, n3, j->st);
  if (rc != WRT_OK) return rc;
  HIPCHK(hipEventRecord(j->e1, j->st));
  rc = wrtk_ray_T_f32(f->g32, d->n, f->zi, f->zc, d->nz, j->d_shift, d->nshift, j->d_dirs,
                      d->nphi, f->acc, j->st);
  if (rc != WRT_OK) return rc;
  HIPCHK(hipEventRecord(j->e3, j->st));
  return wrtk_acc_to_f32(f->acc, f->out, n3, j->st);
}

/* dispatch by operation; each job owns one fast state */
size_t wrt_fast_bytes(const wrt_job_desc* d) {
  if (!wrt_fast_supported(d)) return 0;
  switch (d->op) {
    case WRT_OP_RADON: return radon_fast_bytes(d);
    case WRT_OP_RAY: return ray_fast_bytes(d);
    case WRT_OP_RADON_T: return radon_T_fast_bytes(d);
    case WRT_OP_RAY_T: return ray_T_fast_bytes(d);
    default: return adjoint_fast_bytes(d);
  }
}

void wrt_fast_destroy(wrt_job* j) {
  if (!j->fast) return;
  switch (j->d.op) {
    case WRT_OP_RADON: radon_fast_destroy(j); break;
    case WRT_OP_RAY: ray_fast_destroy(j); break;
    case WRT_OP_RADON_T: radon_T_fast_destroy(j); break;
    case WRT_OP_RAY_T: ray_T_fast_destroy(j); break;
    default: adjoint_fast_destroy(j); break;
  }
}

int wrt_fast_create(wrt_job* j) {
  int rc;
  switch (j->d.op) {
    case WRT_OP_RADON: rc = radon_fast_create(j); break;
    case WRT_OP_RAY: rc = ray_fast_create(j); break;
    case WRT_OP_RADON_T: rc = radon_T_fast_create(j); break;
    case WRT_OP_RAY_T: rc = ray_T_fast_create(j); break;
    default: rc = adjoint_fast_create(j); break;
  }
  if (rc != WRT_OK && j->fast) wrt_fast_destroy(j);
  return rc;
}

int wrt_fast_load(wrt_job* j, const double* in) {
  switch (j->d.op) {
    case WRT_OP_RADON: return radon_fast_load(j, in);
    case WRT_OP_RAY: return ray_fast_load(j, in);
    case WRT_OP_RADON_T: return radon_T_fast_load(j, in);
    case WRT_OP_RAY_T: return ray_T_fast_load(j, in);
    default: return adjoint_fast_load(j, in);
  }
}

int wrt_fast_enqueue(wrt_job* j) {
  switch (j->d.op) {
    case WRT_OP_RADON: return radon_fast_enqueue(j);
    case WRT_OP_RAY: return ray_fast_enqueue(j);
    case WRT_OP_RADON_T: return radon_T_fast_enqueue(j);
    case WRT_OP_RAY_T: return ray_T_fast_enqueue(j);
    default: return adjoint_fast_enqueue(j);
  }
}

int wrt_fast_fetch(wrt_job* j, double* out) {
  const wrt_job_desc* d = &j->d;
  const size_t n3 = (size_t)d->n * d->n * d->n;
  switch (d->op) {
    case WRT_OP_RADON:
      return fetch_f32(j, ((radon_fast*)j->fast)->out, (size_t)d->nshift * d->nphi * d->ntheta, out);
    case WRT_OP_RAY:
      return fetch_f32(j, ((ray_fast*)j->fast)->out, (size_t)d->nz * d->nshift * d->nphi, out);
    case WRT_OP_RADON_T: return fetch_f32(j, ((radon_T_fast*)j->fast)->out, n3, out);
    case WRT_OP_RAY_T: return fetch_f32(j, ((ray_T_fast*)j->fast)->out, n3, out);
    default: return fetch_f32(j, ((adjoint_fast*)j->fast)->out, n3, out);
  }
}

int wrt_fast_tuning(const wrt_job* j, int* band_rows, int* dgroup) {
  if (!j || !j->fast || j->d.op != WRT_OP_RADON || j->d.prec == WRT_F64) return WRT_E_ARG;
  const radon_fast* f = (const radon_fast*)j->fast;
  if (!f->tuned) return WRT_E_STATE;
  if (band_rows) *band_rows = f->band_rows;
  if (dgroup) *dgroup = f->dgroup;
  return WRT_OK;
}

/* the FP32 plane values of a fast radon job from its fixed-point plane sums, i.e. the
 * rounding the per-job launch does on the device, done here for the rounds of the serve loop
 * (which leave the FP32 buffer alone) */
int wrt_fast_fetch_acc(wrt_job* j, double* out) {
  if (!j || !j->fast || !out || j->d.op != WRT_OP_RADON) return WRT_E_ARG;
  const wrt_job_desc* d = &j->d;
  const radon_fast* f = (const radon_fast*)j->fast;
  const size_t count = (size_t)d->nshift * d->nphi * d->ntheta;
  /* the 8-byte sums land in out's slots and are converted in place */
  HIPCHK(hipMemcpyAsync(out, f->acc, count * sizeof(long long), hipMemcpyDeviceToHost, j->st));
  HIPCHK(hipStreamSynchronize(j->st));
  for (size_t i = 0; i < count; ++i) {
    long long v;
    memcpy(&v, &out[i], sizeof(v));
    out[i] = (double)(float)((double)v * f->qstep);
  }
  return WRT_OK;
}

/* a loaded fast job as an item of the persistent batch launch (ra.dgroup left to the
 * planner; fd->dgroup is the per-job launch's choice), with the work the planner splits
 * into tasks */
int wrt_fast_describe(wrt_job* j, wrt_fast_desc* fd) {
  if (!j || !j->fast || !fd) return WRT_E_ARG;
  const wrt_job_desc* d = &j->d;
  memset(fd, 0, sizeof(*fd));
  wrt_batch_item* it = &fd->item;
  if (d->op == WRT_OP_RADON) {
    radon_fast* f = (radon_fast*)j->fast;
    it->kind = WRT_ITEM_RADON;
    it->pow2 = (d->n & (d->n - 1)) == 0;
    it->h = d->n / 2.0;
    it->dc = 2.0 / d->n;
    it->inv_dc = 1.0 / it->dc;
    it->ra.bricks = f->bricks;
    it->ra.n = d->n;
    it->ra.nb = f->nb;
    it->ra.bf16 = d->prec == WRT_BF16;
    it->ra.dirs = j->d_dirs;
    it->ra.ndir = d->nphi * d->ntheta;
    it->ra.shift = (const double*)f->sorted;
    it->ra.sidx = (const int*)((const char*)f->sorted + (size_t)d->nshift * sizeof(double));
    it->ra.nshift = d->nshift;
    it->ra.acc = f->acc;
    it->ra.qmul = (1.0 / d->n) * (1.0 / d->n) / f->qstep;
    it->ra.rho2 = f->rho2;
    fd->brick = f->ne_brick;
    fd->weight = f->ne_count;
    fd->bnd = f->ne_bnd;
    fd->nbrick = f->n_ne;
    fd->dgroup = f->dgroup;
    fd->acc = f->acc;
    fd->qstep = f->qstep;
    fd->out = f->out;
    fd->nout = (long long)d->nshift * d->nphi * d->ntheta;
  } else {
    adjoint_fast* f = (adjoint_fast*)j->fast;
    it->kind = WRT_ITEM_ADJOINT;
    it->adirs = f->dirs;
    it->atrig = f->trig;
    it->acub = f->cub;
    it->aout = f->out;
    it->aorder = f->order;
    it->andir = d->nphi * d->ntheta;
    it->ans = d->nshift;
    it->an = d->n;
    it->as0 = f->s0;
    it->ads = f->ds;
    fd->nvox = (long long)d->n * d->n * d->n;
  }
  return WRT_OK;
}
