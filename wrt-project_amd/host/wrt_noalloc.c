/*
 * wrt_noalloc.c -- wrt_radon_grid_noalloc: the FP32 plane integrals of wrt_radon_grid (the
 * same brick kernel, direction tables and index decisions) without allocating anything: no
 * hipMalloc / hipFree, no host heap.  Every device buffer lives in the library's fixed
 * scratch (WRT_NOALLOC_BYTES per GPU, csrc/k_noalloc.hip) and every host-side table in
 * static storage below.  The caller passes host buffers as for wrt_radon_grid.
 *
 * Per call, on the context stream:
 *   clear the brick counts, the support radius and the task counts;
 *   direction tables (host libm, the reference's bits) and the ascending shift grid with
 *   each position's row, staged through static host arrays;
 *   the volume one brick row (32 x-planes) at a time into the scratch's slab, each row cut
 *   into FP32 bricks;
 *   read back the support radius, build the task list on the device, read back its counts;
 *   zero the FP64 plane accumulators, run the interior and face brick launches, round every
 *   plane to FP32 and copy the planes (as doubles) to the caller.
 * The call returns once the output is in the caller's buffer.  The scratch keeps the last
 * call's contents; calls that overlap on one GPU (two host threads, two contexts) share it
 * and race, as do overlapping calls from several threads on the static host tables.
 */
#include <float.h>
#include <math.h>
#include <string.h>

#include "wrt_internal.h"

#define HIPCHK(x)                            \
  do {                                       \
    if ((x) != hipSuccess) return WRT_E_HIP; \
  } while (0)
#define RCHK(x)                        \
  do {                                 \
    const int r__ = (x);               \
    if (r__ != WRT_OK) return r__;     \
  } while (0)

#define NA_MAX_DIRS 65536
#define NA_MAX_SHIFTS 8192
static double na_dirs[9 * (size_t)NA_MAX_DIRS];
static double na_sv[NA_MAX_SHIFTS];
static int na_si[NA_MAX_SHIFTS];

/* byte offsets of the scratch regions (256-byte aligned) */
typedef struct {
  size_t flags, r2, counts, tasks, dirs, shifts, sidx, acc, bricks, slab, end;
} na_layout;

static size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

static void na_plan(int n, int ndir, int nshift, int ngroups, na_layout* L) {
  const size_t nb = (size_t)(n + WRT_BRICK - 1) / WRT_BRICK, nbr = nb * nb * nb;
  size_t o = 0;
  L->flags = o;
  o = up256(o + nbr * sizeof(int));
  L->r2 = o;
  o += 16; /* the squared support radius, then the largest |value| */
  L->counts = o;
  o = up256(o + 2 * sizeof(int));
  L->tasks = o;
  o = up256(o + 2 * nbr * (size_t)ngroups * sizeof(int));
  L->dirs = o;
  o = up256(o + 9 * (size_t)ndir * sizeof(double));
  L->shifts = o;
  o = up256(o + (size_t)nshift * sizeof(double));
  L->sidx = o;
  o = up256(o + (size_t)nshift * sizeof(int));
  L->acc = o;
  o = up256(o + (size_t)nshift * ndir * sizeof(double));
  L->bricks = o;
  o = up256(o + nbr * WRT_BRICK * WRT_BRICK * WRT_BRICK * sizeof(float));
  L->slab = o;
  o = up256(o + (size_t)WRT_BRICK * n * n * sizeof(double));
  L->end = o;
}

/* ascending shifts (ties in the caller's order) and each position's row; Shell sort with
 * Knuth's gaps, in place, so that nothing is allocated */
static void na_sort_shifts(const double* s, int ns) {
  for (int j = 0; j < ns; ++j) {
    na_sv[j] = s[j];
    na_si[j] = j;
  }
  int gap = 1;
  while (gap < ns / 3) gap = 3 * gap + 1;
  for (; gap > 0; gap /= 3)
    for (int i = gap; i < ns; ++i) {
      const double v = na_sv[i];
      const int r = na_si[i];
      int k = i;
      while (k >= gap && (na_sv[k - gap] > v || (na_sv[k - gap] == v && na_si[k - gap] > r))) {
        na_sv[k] = na_sv[k - gap];
        na_si[k] = na_si[k - gap];
        k -= gap;
      }
      na_sv[k] = v;
      na_si[k] = r;
    }
}

int wrt_radon_grid_noalloc(wrt_ctx* c, const double* vol, int n, const double* phi, int nphi,
                           const double* theta, int ntheta, const double* shift, int nshift,
                           double* out, wrt_stats* st) {
  if (!c || !vol || !out || n < 1 || n > 4096 || wrt_check_grid(phi, nphi, 1) ||
      wrt_check_grid(theta, ntheta, 1) || wrt_check_grid(shift, nshift, 1))
    return WRT_E_ARG;
  const long long ndir_ll = (long long)nphi * ntheta;
  const int nb = (n + WRT_BRICK - 1) / WRT_BRICK;
  if (ndir_ll > NA_MAX_DIRS || nshift > NA_MAX_SHIFTS ||
      (long long)nb * nb * nb > WRT_NOALLOC_MAX_BRICKS)
    return WRT_E_SCRATCH;
  const int ndir = (int)ndir_ll;
  const int nbr = nb * nb * nb;
  /* direction groups as the per-job form chooses them, with every brick counted */
  long long cap = ndir / 16;
  cap = cap < 512 ? 512 : (cap > 2048 ? 2048 : cap);
  long long want = (long long)ndir * nbr / 2048;
  int dg = (int)(want < 16 ? 16 : (want > cap ? cap : want));
  dg = (dg + 15) & ~15;
  const int ngroups = (ndir + dg - 1) / dg;
  na_layout L;
  na_plan(n, ndir, nshift, ngroups, &L);
  if (L.end > WRT_NOALLOC_BYTES) return WRT_E_SCRATCH;
  char* base = (char*)wrtk_noalloc_scratch(c->device);
  if (!base) return WRT_E_STATE;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const size_t nplanes = (size_t)nshift * ndir;

  /* brick counts, support radius and task counts start at zero */
  RCHK(wrtk_noalloc_zero(base + L.flags, (L.tasks - L.flags) / 8, s));
  /* tables through the static host arrays (each copy completes before they are reused) */
  wrt_host_radon_dirs(n, phi, nphi, theta, ntheta, na_dirs);
  na_sort_shifts(shift, nshift);
  HIPCHK(hipMemcpyAsync(base + L.dirs, na_dirs, 9 * (size_t)ndir * sizeof(double),
                        hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(base + L.shifts, na_sv, (size_t)nshift * sizeof(double),
                        hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(base + L.sidx, na_si, (size_t)nshift * sizeof(int),
                        hipMemcpyHostToDevice, s));
  /* the volume one brick row at a time through the slab */
  for (int bx = 0; bx < nb; ++bx) {
    const int x0 = bx * WRT_BRICK, x1 = x0 + WRT_BRICK < n ? x0 + WRT_BRICK : n;
    HIPCHK(hipMemcpyAsync(base + L.slab, vol + (size_t)x0 * n * n,
                          (size_t)(x1 - x0) * n * n * sizeof(double), hipMemcpyHostToDevice, s));
    RCHK(wrtk_brickify((const double*)(base + L.slab), n, nb, bx, 1, 0, base + L.bricks,
                       (int*)(base + L.flags), (unsigned long long*)(base + L.r2), s));
  }
  /* the support radius and the fixed-point unit, as the per-job form derives them */
  unsigned long long r2bits[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(r2bits, base + L.r2, sizeof(r2bits), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  double r2;
  memcpy(&r2, &r2bits[0], sizeof(r2));
  const unsigned vb32 = (unsigned)r2bits[1];
  float vmax;
  memcpy(&vmax, &vb32, sizeof(vmax));
  if (!(vmax <= FLT_MAX)) return WRT_E_ARG; /* a value that is not finite in FP32 */
  const double qstep = wrt_plane_qstep(vmax);
  const double rho = sqrt(r2) + sqrt(3.0) / n;
  const double rho2 = rho * rho * (1.0 + 1e-12) + 1e-12;
  /* task list */
  RCHK(wrtk_noalloc_tasks((const int*)(base + L.flags), n, nb, ngroups, (int*)(base + L.tasks),
                          (int*)(base + L.counts), s));
  int counts[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(counts, base + L.counts, sizeof(counts), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));

  HIPCHK(hipEventRecord(c->ev2, s));
  RCHK(wrtk_noalloc_zero(base + L.acc, nplanes, s));
  wrt_radon_fast_args a;
  memset(&a, 0, sizeof(a));
  a.bricks = base + L.bricks;
  a.n = n;
  a.nb = nb;
  a.bf16 = 0;
  a.dirs = (const double*)(base + L.dirs);
  a.ndir = ndir;
  a.dgroup = dg;
  a.shift = (const double*)(base + L.shifts);
  a.sidx = (const int*)(base + L.sidx);
  a.nshift = nshift;
  a.acc = (long long*)(base + L.acc);
  a.qmul = (1.0 / n) * (1.0 / n) / qstep;
  a.rho2 = rho2;
  HIPCHK(hipEventRecord(c->ev0, s));
  a.tasks = (const int*)(base + L.tasks);
  a.ntasks = counts[0];
  RCHK(wrtk_radon_fast(&a, 0, s));
  a.tasks = (const int*)(base + L.tasks) + 2 * (size_t)counts[0];
  a.ntasks = counts[1];
  RCHK(wrtk_radon_fast(&a, 1, s));
  HIPCHK(hipEventRecord(c->ev1, s));
  RCHK(wrtk_noalloc_round((long long*)(base + L.acc), qstep, nplanes, s));
  HIPCHK(hipEventRecord(c->ev3, s));
  HIPCHK(hipMemcpyAsync(out, base + L.acc, nplanes * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (st) {
    float ms = 0, mk = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev2, c->ev3));
    HIPCHK(hipEventElapsedTime(&mk, c->ev0, c->ev1));
    st->seconds = ms * 1e-3;
    st->kernel_seconds = mk * 1e-3;
    st->logical_samples = 16.0 * n * (double)n * (double)nplanes;
    st->alg_bytes = 0;
    st->outputs = (long long)nplanes;
  }
  return WRT_OK;
}
