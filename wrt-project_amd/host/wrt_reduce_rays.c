/*
 * wrt_reduce_rays.c -- the ray -> Radon reduction on independent ray and plane grids.
 *
 * The reference's radon_reduction reads its rays with the output's phi index and spline
 * nodes (radon_reduction/slicing.c:50-60,114), so rays on another grid (for instance 180
 * angles in [0, pi) and 257 shifts, reduced to 64 azimuths in [0, 2 pi) and 129 offsets)
 * cannot be passed to it.  Here the ray rows at each plane azimuth are formed first:
 *
 *   P(z, s, phi + pi) = P(z, -s, phi)      (the line at phi + pi through -s is the same line;
 *                                            ray_grid/src/integral.c:11-50 samples it with
 *                                            the same symmetric parameter set)
 *
 * maps every azimuth into the ray grid's period, and 4-point Lagrange interpolation over
 * the periodic ray-angle grid (with s -> -s wherever the stencil crosses a half turn) gives
 * the row.  Because the natural cubic spline is linear in its data and the ray shift grid
 * is symmetric about 0, the spline of the combined row equals the same combination of the
 * rows' splines, each evaluated at +-s.  The shift integral is then the reference's
 * (radon_reduction/slicing.c:80-119) with the spline nodes on the ray shift grid and the
 * plane offsets on the output grid.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "wrt_internal.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* 4-point Lagrange weights on the nodes -1, 0, 1, 2 at u in [0, 1) */
static void lagrange4(double u, double w[4]) {
  w[0] = -u * (u - 1.0) * (u - 2.0) / 6.0;
  w[1] = (u + 1.0) * (u - 1.0) * (u - 2.0) / 2.0;
  w[2] = -(u + 1.0) * u * (u - 2.0) / 2.0;
  w[3] = (u + 1.0) * u * (u - 1.0) / 6.0;
}

/* stencil[j][12]: weights, ray-angle indices, flags (1: the row at -s) */
static void build_stencil(const double* phi, int nphi, int nrphi, int half_turn, double* sten) {
  const double period = half_turn ? M_PI : 2.0 * M_PI;
  const double dphi = period / nrphi;
  for (int j = 0; j < nphi; ++j) {
    double a = fmod(phi[j], 2.0 * M_PI);
    if (a < 0) a += 2.0 * M_PI;
    int flip = 0;
    if (half_turn && a >= M_PI) {
      a -= M_PI;
      flip = 1;
    }
    const double u = a / dphi;
    double k0 = floor(u), fr = u - k0;
    /* an azimuth on a ray angle (up to the rounding of the two grids) takes that row alone */
    if (fr < 1e-12) {
      fr = 0.0;
    } else if (fr > 1.0 - 1e-12) {
      fr = 0.0;
      k0 += 1.0;
    }
    double w[4];
    lagrange4(fr, w);
    double* s = sten + 12 * (size_t)j;
    for (int m = 0; m < 4; ++m) {
      long idx = (long)k0 - 1 + m;
      int f = flip;
      while (idx < 0) {
        idx += nrphi;
        f ^= half_turn;
      }
      while (idx >= nrphi) {
        idx -= nrphi;
        f ^= half_turn;
      }
      s[m] = w[m];
      s[4 + m] = (double)idx;
      s[8 + m] = (double)f;
    }
  }
}

static int increasing(const double* g, int n) {
  for (int i = 1; i < n; ++i)
    if (!(g[i] > g[i - 1])) return 0;
  return 1;
}

int wrt_radon_reduction_rays(wrt_ctx* c, const double* rays, int nz, const double* rshift,
                             int nrshift, int nrphi, int half_turn, const double* shift,
                             int nshift, const double* phi, int nphi, const double* theta,
                             int ntheta, double* out, wrt_stats* st) {
  if (!c || !rays || !out || nz < 3 || nrphi < 4 || (half_turn != 0 && half_turn != 1) ||
      wrt_check_grid(rshift, nrshift, 3) || !increasing(rshift, nrshift) ||
      wrt_check_grid(shift, nshift, 1) || wrt_check_grid(phi, nphi, 1) ||
      wrt_check_grid(theta, ntheta, 1))
    return WRT_E_ARG;
  if (half_turn)
    for (int k = 0; k < nrshift; ++k)
      if (fabs(rshift[k] + rshift[nrshift - 1 - k]) > 1e-12) return WRT_E_ARG;
  if (hipSetDevice(c->device) != hipSuccess) return WRT_E_HIP;

  const size_t nin = (size_t)nz * nrshift * nrphi;
  const size_t nrow = (size_t)nz * nphi * nrshift;
  const size_t nout = (size_t)nshift * nphi * ntheta;
  /* tables: rshift, spline factor, plane offsets, theta trig, stencil */
  const size_t o_fac = nrshift, o_p = 3 * (size_t)nrshift, o_tt = o_p + nshift,
               o_st = o_tt + 2 * (size_t)ntheta, ntab = o_st + 12 * (size_t)nphi;
  double* h = (double*)calloc(ntab, sizeof(double));
  if (!h) return WRT_E_NOMEM;
  memcpy(h, rshift, sizeof(double) * nrshift);
  memcpy(h + o_p, shift, sizeof(double) * nshift);
  for (int i = 0; i < ntheta; ++i) {
    h[o_tt + 2 * i] = cos(theta[i]);
    h[o_tt + 2 * i + 1] = sin(theta[i]);
  }
  build_stencil(phi, nphi, nrphi, half_turn, h + o_st);

  double *d_tab = NULL, *d_in = NULL, *d_row = NULL, *d_spl = NULL, *d_out = NULL;
  int rc = wrt_hip(hipMalloc((void**)&d_tab, ntab * 8));
  if (rc == WRT_OK) rc = wrt_hip(hipMalloc((void**)&d_in, nin * 8));
  if (rc == WRT_OK) rc = wrt_hip(hipMalloc((void**)&d_row, nrow * 8));
  if (rc == WRT_OK) rc = wrt_hip(hipMalloc((void**)&d_spl, 2 * nrow * 8));
  if (rc == WRT_OK) rc = wrt_hip(hipMalloc((void**)&d_out, nout * 8));
  if (rc == WRT_OK) rc = wrt_hip(hipMemcpyAsync(d_tab, h, ntab * 8, hipMemcpyHostToDevice, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipMemcpyAsync(d_in, rays, nin * 8, hipMemcpyHostToDevice, c->stream));
  if (rc == WRT_OK) rc = wrtk_spline_factor(d_tab, nrshift, d_tab + o_fac, c->stream);
  if (rc == WRT_OK) rc = wrt_hip(hipEventRecord(c->ev0, c->stream));
  if (rc == WRT_OK)
    rc = wrtk_ray_phi_resample(d_in, nz, nrshift, nrphi, d_tab + o_st, nphi, d_row, c->stream);
  if (rc == WRT_OK)
    rc = wrtk_spline_rows(d_tab, nrshift, d_tab + o_fac, d_row, d_spl, nz * nphi, 1,
                          nrshift, 1, c->stream);
  if (rc == WRT_OK) rc = wrt_hip(hipEventRecord(c->ev2, c->stream));
  if (rc == WRT_OK)
    rc = wrtk_reduction(d_spl, nz, d_tab, nrshift, d_tab + o_p, nshift, nphi, d_tab + o_tt,
                        ntheta, d_out, c->stream);
  if (rc == WRT_OK) rc = wrt_hip(hipEventRecord(c->ev1, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipMemcpyAsync(out, d_out, nout * 8, hipMemcpyDeviceToHost, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipStreamSynchronize(c->stream));
  if (rc == WRT_OK && st) {
    float ms = 0, msk = 0;
    rc = wrt_hip(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (rc == WRT_OK) rc = wrt_hip(hipEventElapsedTime(&msk, c->ev2, c->ev1));
    st->seconds = ms * 1e-3;
    st->kernel_seconds = msk * 1e-3;
    st->logical_samples = (4.0 * nz + 1.0) * (double)nout;
    st->alg_bytes = 0;
    st->outputs = (long long)nout;
  }
  hipFree(d_tab);
  hipFree(d_in);
  hipFree(d_row);
  hipFree(d_spl);
  hipFree(d_out);
  free(h);
  return rc;
}
