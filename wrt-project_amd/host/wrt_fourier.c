/*
 * wrt_fourier.c -- host side of the Fourier-slice (Chang-type) inversion
 * (radon_inversion/chang3d.m:3-30 = rtft3d.m:3-137 then nfft_reconstruct_3d.m:19-53): the
 * grid constants and tables of rtft3d, the two kernels of k_fourier.hip, FP64.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "wrt_internal.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

int wrtk_rtft(const double* radon, int nshift, int ndir, const void* tw, const void* corr,
              const double* jw, const double* rho, int n, void* coef, hipStream_t st);
int wrtk_fourier_adjoint(const void* coef, int n, double dfreq, const void* dirs, int ndir,
                         int ngrid, int nslice, double inv_w0, double* part, double* out,
                         hipStream_t st);

/* direction slices of the adjoint sum: enough waves to fill the chip (about 16 per SIMD)
 * when the grid alone has too few, within a 512 MiB buffer of partial sums */
static int adjoint_slices(int ngrid, int ndir) {
  const double nvox = (double)ngrid * ngrid * ngrid;
  const double waves = (double)ngrid * ngrid * ((ngrid + 3) / 4) / 64.0;
  double ns = ceil(16384.0 / waves);
  const double cap = floor(512.0 * 1024 * 1024 / (nvox * 8));
  if (ns > cap) ns = cap;
  if (ns > ndir) ns = ndir;
  if (ns > 65535) ns = 65535;
  return ns < 1 ? 1 : (int)ns;
}

int wrt_fourier_inversion(wrt_ctx* c, const double* radon, const double* phi, int nphi,
                          const double* theta, const double* theta_w, int ntheta, int nshift,
                          double rsupp, double padding, double w0, int ngrid, double* out,
                          wrt_stats* st) {
  if (!c || !radon || !out || nshift < 2 || nshift > 8192 || ngrid < 2 || ngrid > 2048 ||
      (ngrid & 1) || !isfinite(rsupp) || rsupp <= 0.0 || !isfinite(padding) || padding <= 0.0 ||
      !isfinite(w0) || w0 == 0.0 || wrt_check_grid(phi, nphi, 1) ||
      wrt_check_grid(theta, ntheta, 1) || wrt_check_grid(theta_w, ntheta, 1))
    return WRT_E_ARG;
  /* rtft3d.m:29-35 */
  const double dphi = 2 * M_PI / nphi;
  const double dshift = 2 * rsupp / (nshift - 1);
  const double ntd = padding * floor(1 / dshift);
  if (ntd < 2 || ntd > 65536 || ntd != floor(ntd) || fmod(ntd, 2.0) != 0.0) return WRT_E_ARG;
  const int n = (int)ntd;
  const double dfreq = 1 / (n * dshift);
  const int ndir = nphi * ntheta;
  if (hipSetDevice(c->device) != hipSuccess) return WRT_E_HIP;

  /* tables: twiddles, shift correction, frequencies, jacobian factors, directions */
  const size_t nb_tw = (size_t)n * 16, nb_rho = (size_t)n * 8, nb_jw = (size_t)ndir * 8,
               nb_dir = (size_t)ndir * 32;
  const size_t nb_tab = 2 * nb_tw + nb_rho + nb_jw + nb_dir;
  char* h = (char*)malloc(nb_tab);
  if (!h) return WRT_E_NOMEM;
  double* tw = (double*)h;
  double* corr = (double*)(h + nb_tw);
  double* rho = (double*)(h + 2 * nb_tw);
  double* jw = (double*)(h + 2 * nb_tw + nb_rho);
  double* dirs = (double*)(h + 2 * nb_tw + nb_rho + nb_jw);
  for (int q = 0; q < n; ++q) { /* exp(-2 pi i q / n), from the exact fraction q / n */
    const long double a = 2.0L * 3.141592653589793238462643383279502884L * q / n;
    tw[2 * q] = (double)cosl(a);
    tw[2 * q + 1] = (double)-sinl(a);
  }
  for (int k = 0; k < n; ++k) { /* exp(2 pi i rsupp f_k), f_k = k dfreq (rtft3d.m:48) */
    const double a = (2 * M_PI * rsupp) * (k * dfreq);
    corr[2 * k] = cos(a);
    corr[2 * k + 1] = sin(a);
  }
  for (int m = 0; m < n; ++m) rho[m] = (m - n / 2) * dfreq; /* centered_frequencies, fftshift */
  for (int ip = 0; ip < nphi; ++ip)
    for (int it = 0; it < ntheta; ++it) {
      const int d = ip * ntheta + it;
      jw[d] = 0.5 * dshift * dfreq * dphi * theta_w[it]; /* rtft3d.m:93 without rho^2 */
      dirs[4 * d] = sin(theta[it]) * cos(phi[ip]);
      dirs[4 * d + 1] = sin(theta[it]) * sin(phi[ip]);
      dirs[4 * d + 2] = cos(theta[it]);
      dirs[4 * d + 3] = 0.0;
    }

  const size_t nrad = (size_t)nshift * ndir, nvox = (size_t)ngrid * ngrid * ngrid;
  char* d_tab = NULL;
  double *d_rad = NULL, *d_out = NULL;
  void* d_coef = NULL;
  int rc = wrt_hip(hipMalloc((void**)&d_tab, nb_tab));
  if (rc == WRT_OK) rc = wrt_hip(hipMalloc((void**)&d_rad, nrad * 8));
  const int nslice = adjoint_slices(ngrid, ndir);
  double* d_part = NULL;
  if (rc == WRT_OK) rc = wrt_hip(hipMalloc(&d_coef, (size_t)ndir * (n / 2 + 1) * 16));
  if (rc == WRT_OK) rc = wrt_hip(hipMalloc((void**)&d_out, nvox * 8));
  if (rc == WRT_OK && nslice > 1) rc = wrt_hip(hipMalloc((void**)&d_part, nvox * 8 * nslice));
  if (rc == WRT_OK) rc = wrt_hip(hipMemcpyAsync(d_tab, h, nb_tab, hipMemcpyHostToDevice, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipMemcpyAsync(d_rad, radon, nrad * 8, hipMemcpyHostToDevice, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipEventRecord(c->ev0, c->stream));
  if (rc == WRT_OK)
    rc = wrtk_rtft(d_rad, nshift, ndir, d_tab, d_tab + nb_tw, (const double*)(d_tab + 2 * nb_tw + nb_rho),
                   (const double*)(d_tab + 2 * nb_tw), n, d_coef, c->stream);
  if (rc == WRT_OK)
    rc = wrtk_fourier_adjoint(d_coef, n, dfreq, d_tab + 2 * nb_tw + nb_rho + nb_jw, ndir, ngrid,
                              nslice, 1.0 / w0, d_part, d_out, c->stream);
  if (rc == WRT_OK) rc = wrt_hip(hipEventRecord(c->ev1, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipMemcpyAsync(out, d_out, nvox * 8, hipMemcpyDeviceToHost, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipStreamSynchronize(c->stream));
  if (rc == WRT_OK && st) {
    float ms = 0;
    rc = wrt_hip(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    st->seconds = st->kernel_seconds = ms * 1e-3;
    st->logical_samples = (double)nvox * ndir * n; /* terms of the adjoint sum */
    st->alg_bytes = 0;
    st->outputs = (long long)nvox;
  }
  hipFree(d_tab);
  hipFree(d_rad);
  hipFree(d_coef);
  hipFree(d_out);
  hipFree(d_part);
  free(h);
  return rc;
}
