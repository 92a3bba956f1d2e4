/*
 * wrt_spectral.c -- the spectral Laplacians that follow the backprojection
 * (radon_inversion/backprojection/laplace2dfft.m, laplace3dfft.m): real-to-complex DFT
 * (hipFFT), the frequency factor (k_spectral.hip), complex-to-real DFT.  FP64 throughout, as
 * the reference's Octave code computes.  The plan pair is made per call and freed at its end
 * (see the end of laplace()).
 */
#include <hipfft/hipfft.h>
#include <math.h>
#include <stdlib.h>

#include "wrt_internal.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

int wrtk_laplace_factor(void* spec, int n, int dims, double coef, double scale, hipStream_t st);

typedef struct {
  hipfftHandle fwd, inv;
  int n, dims;
} spectral_plan;

void wrt_spectral_release(wrt_ctx* c) {
  spectral_plan* p = (spectral_plan*)c->spectral;
  if (!p) return;
  hipfftDestroy(p->fwd);
  hipfftDestroy(p->inv);
  free(p);
  c->spectral = NULL;
}

static int get_plan(wrt_ctx* c, int n, int dims, spectral_plan** out) {
  spectral_plan* p = (spectral_plan*)c->spectral;
  if (p && p->n == n && p->dims == dims) {
    *out = p;
    return WRT_OK;
  }
  wrt_spectral_release(c);
  p = (spectral_plan*)calloc(1, sizeof(spectral_plan));
  if (!p) return WRT_E_NOMEM;
  hipfftResult r1, r2;
  if (dims == 3) {
    r1 = hipfftPlan3d(&p->fwd, n, n, n, HIPFFT_D2Z);
    r2 = r1 == HIPFFT_SUCCESS ? hipfftPlan3d(&p->inv, n, n, n, HIPFFT_Z2D) : r1;
  } else {
    r1 = hipfftPlan2d(&p->fwd, n, n, HIPFFT_D2Z);
    r2 = r1 == HIPFFT_SUCCESS ? hipfftPlan2d(&p->inv, n, n, HIPFFT_Z2D) : r1;
  }
  if (r1 != HIPFFT_SUCCESS || r2 != HIPFFT_SUCCESS) {
    if (r1 == HIPFFT_SUCCESS) hipfftDestroy(p->fwd);
    free(p);
    return r1 == HIPFFT_ALLOC_FAILED || r2 == HIPFFT_ALLOC_FAILED ? WRT_E_NOMEM : WRT_E_HIP;
  }
  p->n = n;
  p->dims = dims;
  if (hipfftSetStream(p->fwd, c->stream) != HIPFFT_SUCCESS ||
      hipfftSetStream(p->inv, c->stream) != HIPFFT_SUCCESS) {
    c->spectral = p;
    wrt_spectral_release(c);
    return WRT_E_HIP;
  }
  c->spectral = p;
  *out = p;
  return WRT_OK;
}

static int laplace(wrt_ctx* c, const double* u, int n, int dims, double period, double* out,
                   wrt_stats* st) {
  if (!c || !u || !out || n < 1 || n > (dims == 3 ? 2048 : 65536) || !isfinite(period) ||
      period <= 0.0)
    return WRT_E_ARG;
  if (hipSetDevice(c->device) != hipSuccess) return WRT_E_HIP;
  const size_t count = dims == 3 ? (size_t)n * n * n : (size_t)n * n;
  const size_t ccount = (dims == 3 ? (size_t)n * n : (size_t)n) * (size_t)(n / 2 + 1);
  spectral_plan* p = NULL;
  int rc = get_plan(c, n, dims, &p);
  if (rc != WRT_OK) return rc;
  double* d_u = NULL;
  void* d_spec = NULL;
  rc = wrt_hip(hipMalloc((void**)&d_u, count * sizeof(double)));
  if (rc == WRT_OK) rc = wrt_hip(hipMalloc(&d_spec, ccount * 2 * sizeof(double)));
  if (rc == WRT_OK) rc = wrt_hip(hipMemcpyAsync(d_u, u, count * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipEventRecord(c->ev0, c->stream));
  if (rc == WRT_OK && hipfftExecD2Z(p->fwd, d_u, (hipfftDoubleComplex*)d_spec) != HIPFFT_SUCCESS) rc = WRT_E_HIP;
  /* U = -(2 pi / period)^2 |k|^2 (laplace3dfft.m:21), times the inverse DFT's 1/n^d */
  const double w = 2 * M_PI / period;
  if (rc == WRT_OK) rc = wrtk_laplace_factor(d_spec, n, dims, -(w * w), 1.0 / (double)count, c->stream);
  if (rc == WRT_OK && hipfftExecZ2D(p->inv, (hipfftDoubleComplex*)d_spec, d_u) != HIPFFT_SUCCESS) rc = WRT_E_HIP;
  if (rc == WRT_OK) rc = wrt_hip(hipEventRecord(c->ev1, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipMemcpyAsync(out, d_u, count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (rc == WRT_OK) rc = wrt_hip(hipStreamSynchronize(c->stream));
  if (rc == WRT_OK && st) {
    float ms = 0;
    rc = wrt_hip(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    st->seconds = st->kernel_seconds = ms * 1e-3;
    st->logical_samples = (double)count;
    st->alg_bytes = (double)count * 16 + (double)ccount * 32;
    st->outputs = (long long)count;
  }
  hipFree(d_u);
  hipFree(d_spec);
  /* The plans are not kept past the call: with a live 256^3 D2Z/Z2D plan pair in the
   * context, the serve loop's rounds of the persistent kernel started later from that
   * context ran measurably slower (8-QPS stream: p99 395 -> 630 ms, burst rate -17%; a new
   * context, or freeing the plans, restores both). */
  wrt_spectral_release(c);
  return rc;
}

int wrt_laplace3d(wrt_ctx* c, const double* u, int n, double period, double* out, wrt_stats* st) {
  return laplace(c, u, n, 3, period, out, st);
}

int wrt_laplace2d(wrt_ctx* c, const double* u, int n, double period, double* out, wrt_stats* st) {
  return laplace(c, u, n, 2, period, out, st);
}
