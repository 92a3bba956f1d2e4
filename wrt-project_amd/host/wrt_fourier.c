/*
 * wrt_fourier.c -- host side of the Fourier-slice (Chang-type) inversions
 * (radon_inversion/chang3d.m:3-30 = rtft3d.m:3-137 then nfft_reconstruct_3d.m:19-53, and
 * chang2d.m:3-26 = rtft2d.m:3-62 then nfft_reconstruct_2d.m:19-45): the grid constants and
 * tables of rtft3d / rtft2d, the kernels of k_fourier.hip, FP64.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "wrt_internal.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

int wrtk_rtft(const double* radon, int nshift, int ndir, const void* tw, const void* corr,
              const double* jw, const double* rw, int n, void* coef, hipStream_t st);
int wrtk_fourier_adjoint(const void* coef, int n, double dfreq, const void* dirs, int ndir, int nx,
                         int ny, int nz, int ngrid, int nslice, double inv_w0, double* part,
                         double* out, hipStream_t st);

/* direction slices of the adjoint sum: enough waves to fill the chip (about 16 per SIMD)
 * when the grid alone has too few, within a 512 MiB buffer of partial sums */
static int adjoint_slices(int nx, int ny, int nz, int ndir) {
  const double nvox = (double)nx * ny * nz;
  const double waves = (double)nx * ny * ((nz + 3) / 4) / 64.0;
  double ns = ceil(16384.0 / waves);
  const double cap = floor(512.0 * 1024 * 1024 / (nvox * 8));
  if (ns > cap) ns = cap;
  if (ns > ndir) ns = ndir;
  if (ns > 65535) ns = 65535;
  return ns < 1 ? 1 : (int)ns;
}

/* the padded length of rtft: ntotal = padding * floor(1 / dshift), an even integer
 * (rtft3d.m:31-33, rtft2d.m:22-24); 0 when it is not */
static int rtft_length(int nshift, double rsupp, double padding, double* dshift, double* dfreq) {
  *dshift = 2 * rsupp / (nshift - 1);
  const double ntd = padding * floor(1 / *dshift);
  if (!(ntd >= 2) || ntd > 65536 || ntd != floor(ntd) || fmod(ntd, 2.0) != 0.0) return 0;
  *dfreq = 1 / (ntd * *dshift);
  return (int)ntd;
}

static int common_args(wrt_ctx* c, const double* radon, double* out, int nshift, double rsupp,
                       double padding, double w0, int ngrid, int gmax) {
  return !c || !radon || !out || nshift < 2 || nshift > 8192 || ngrid < 2 || ngrid > gmax ||
         (ngrid & 1) || !isfinite(rsupp) || rsupp <= 0.0 || !isfinite(padding) ||
         padding <= 0.0 || !isfinite(w0) || w0 == 0.0;
}

/* Tables, upload, the two kernels, download.  dirs [ndir][4] and jw [ndir] are filled by the
 * caller in the host block after the fixed tables; radial: 2 for rho^2, 1 for |rho|. */
typedef struct {
  int n, ndir, radial;
  double dfreq, rsupp;
  char* h;
  size_t nb_tw, nb_rw, nb_jw, nb_dir;
} fourier_tabs;

static int tabs_init(fourier_tabs* t, int n, double dfreq, double rsupp, int ndir, int radial) {
  t->n = n;
  t->ndir = ndir;
  t->radial = radial;
  t->dfreq = dfreq;
  t->rsupp = rsupp;
  t->nb_tw = (size_t)n * 16;
  t->nb_rw = (size_t)n * 8;
  t->nb_jw = (size_t)ndir * 8;
  t->nb_dir = (size_t)ndir * 32;
  t->h = (char*)malloc(2 * t->nb_tw + t->nb_rw + t->nb_jw + t->nb_dir);
  if (!t->h) return WRT_E_NOMEM;
  double* tw = (double*)t->h;
  double* corr = (double*)(t->h + t->nb_tw);
  double* rw = (double*)(t->h + 2 * t->nb_tw);
  for (int q = 0; q < n; ++q) { /* exp(-2 pi i q / n), from the exact fraction q / n */
    const long double a = 2.0L * 3.141592653589793238462643383279502884L * q / n;
    tw[2 * q] = (double)cosl(a);
    tw[2 * q + 1] = (double)-sinl(a);
  }
  for (int k = 0; k < n; ++k) { /* exp(2 pi i rsupp f_k), f_k = k dfreq uncentred (rtft3d.m:48) */
    const double a = (2 * M_PI * rsupp) * (k * dfreq);
    corr[2 * k] = cos(a);
    corr[2 * k + 1] = sin(a);
  }
  for (int m = 0; m < n; ++m) { /* rho_m: centered_frequencies after fftshift */
    const double rho = (m - n / 2) * dfreq;
    rw[m] = radial == 2 ? rho * rho : fabs(rho);
  }
  return WRT_OK;
}

static double* tabs_jw(fourier_tabs* t) { return (double*)(t->h + 2 * t->nb_tw + t->nb_rw); }
static double* tabs_dirs(fourier_tabs* t) {
  return (double*)(t->h + 2 * t->nb_tw + t->nb_rw + t->nb_jw);
}

static int fourier_run(wrt_ctx* c, fourier_tabs* t, const double* radon, int nshift, int nx,
                       int ny, int nz, int ngrid, double w0, double* out, wrt_stats* st) {
  if (hipSetDevice(c->device) != hipSuccess) return WRT_E_HIP;
  const int n = t->n, ndir = t->ndir;
  const size_t nb_tab = 2 * t->nb_tw + t->nb_rw + t->nb_jw + t->nb_dir;
  const size_t nrad = (size_t)nshift * ndir, nvox = (size_t)nx * ny * nz;
  const int nslice = adjoint_slices(nx, ny, nz, ndir);
  char* d_tab = NULL;
  double *d_rad = NULL, *d_out = NULL, *d_part = NULL;
  void* d_coef = NULL;
  int rc = wrt_hip(hipMalloc((void**)&d_tab, nb_tab));
  if (rc == WRT_OK) rc = wrt_hip(hipMalloc((void**)&d_rad, nrad * 8));
  if (rc == WRT_OK) rc = wrt_hip(hipMalloc(&d_coef, (size_t)ndir * (n / 2 + 1) * 16));
  if (rc == WRT_OK) rc = wrt_hip(hipMalloc((void**)&d_out, nvox * 8));
  if (rc == WRT_OK && nslice > 1) rc = wrt_hip(hipMalloc((void**)&d_part, nvox * 8 * nslice));
  if (rc == WRT_OK) rc = wrt_hip(hipMemcpyAsync(d_tab, t->h, nb_tab, hipMemcpyHostToDevice, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipMemcpyAsync(d_rad, radon, nrad * 8, hipMemcpyHostToDevice, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipEventRecord(c->ev0, c->stream));
  if (rc == WRT_OK)
    rc = wrtk_rtft(d_rad, nshift, ndir, d_tab, d_tab + t->nb_tw,
                   (const double*)(d_tab + 2 * t->nb_tw + t->nb_rw),
                   (const double*)(d_tab + 2 * t->nb_tw), n, d_coef, c->stream);
  if (rc == WRT_OK)
    rc = wrtk_fourier_adjoint(d_coef, n, t->dfreq, d_tab + 2 * t->nb_tw + t->nb_rw + t->nb_jw,
                              ndir, nx, ny, nz, ngrid, nslice, 1.0 / w0, d_part, d_out, c->stream);
  if (rc == WRT_OK) rc = wrt_hip(hipEventRecord(c->ev1, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipMemcpyAsync(out, d_out, nvox * 8, hipMemcpyDeviceToHost, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipStreamSynchronize(c->stream));
  if (rc == WRT_OK && st) {
    float ms = 0;
    rc = wrt_hip(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    st->seconds = st->kernel_seconds = ms * 1e-3;
    st->logical_samples = (double)nvox * ndir * n; /* terms of the reference's adjoint sum */
    st->alg_bytes = 0;
    st->outputs = (long long)nvox;
  }
  hipFree(d_tab);
  hipFree(d_rad);
  hipFree(d_coef);
  hipFree(d_out);
  hipFree(d_part);
  return rc;
}

int wrt_fourier_inversion(wrt_ctx* c, const double* radon, const double* phi, int nphi,
                          const double* theta, const double* theta_w, int ntheta, int nshift,
                          double rsupp, double padding, double w0, int ngrid, double* out,
                          wrt_stats* st) {
  if (common_args(c, radon, out, nshift, rsupp, padding, w0, ngrid, 2048) ||
      wrt_check_grid(phi, nphi, 1) || wrt_check_grid(theta, ntheta, 1) ||
      wrt_check_grid(theta_w, ntheta, 1))
    return WRT_E_ARG;
  double dshift, dfreq;
  const int n = rtft_length(nshift, rsupp, padding, &dshift, &dfreq);
  if (!n) return WRT_E_ARG;
  const double dphi = 2 * M_PI / nphi; /* rtft3d.m:29 */
  fourier_tabs t;
  int rc = tabs_init(&t, n, dfreq, rsupp, nphi * ntheta, 2);
  if (rc) return rc;
  double* jw = tabs_jw(&t);
  double* dirs = tabs_dirs(&t);
  for (int ip = 0; ip < nphi; ++ip)
    for (int it = 0; it < ntheta; ++it) {
      const int d = ip * ntheta + it;
      jw[d] = 0.5 * dshift * dfreq * dphi * theta_w[it]; /* rtft3d.m:93 without rho^2 */
      dirs[4 * d] = sin(theta[it]) * cos(phi[ip]);
      dirs[4 * d + 1] = sin(theta[it]) * sin(phi[ip]);
      dirs[4 * d + 2] = cos(theta[it]);
      dirs[4 * d + 3] = 0.0;
    }
  rc = fourier_run(c, &t, radon, nshift, ngrid, ngrid, ngrid, ngrid, w0, out, st);
  free(t.h);
  return rc;
}

int wrt_fourier_inversion2d(wrt_ctx* c, const double* radon, const double* phi, int nphi,
                            int nshift, double rsupp, double padding, double w0, int ngrid,
                            double* out, wrt_stats* st) {
  if (common_args(c, radon, out, nshift, rsupp, padding, w0, ngrid, 32768) ||
      wrt_check_grid(phi, nphi, 1))
    return WRT_E_ARG;
  double dshift, dfreq;
  const int n = rtft_length(nshift, rsupp, padding, &dshift, &dfreq);
  if (!n) return WRT_E_ARG;
  const double dphi = 2 * M_PI / nphi; /* rtft2d.m:20 */
  fourier_tabs t;
  int rc = tabs_init(&t, n, dfreq, rsupp, nphi, 1);
  if (rc) return rc;
  double* jw = tabs_jw(&t);
  double* dirs = tabs_dirs(&t);
  for (int ip = 0; ip < nphi; ++ip) {
    /* rtft2d.m:41-48: weight 0.5 dfreq dphi |rho|, values scaled by dshift */
    jw[ip] = 0.5 * dshift * dfreq * dphi;
    /* the grid runs as 1 x ngrid x ngrid: (x, y) on the kernel's (y, z) axes */
    dirs[4 * ip] = 0.0;
    dirs[4 * ip + 1] = cos(phi[ip]);
    dirs[4 * ip + 2] = sin(phi[ip]);
    dirs[4 * ip + 3] = 0.0;
  }
  rc = fourier_run(c, &t, radon, nshift, 1, ngrid, ngrid, ngrid, w0, out, st);
  free(t.h);
  return rc;
}
