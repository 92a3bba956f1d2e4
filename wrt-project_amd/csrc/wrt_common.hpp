// wrt_common.hpp -- device helpers shared by the FP64 kernels.
//
// Compiled with -ffp-contract=off: every + - * / below rounds once, in the order written,
// which is the reference's C evaluation order.  fma() appears only where it is exact.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/wrt.h"

#define WRT_HOSTDEV __host__ __device__ __forceinline__

// Address spaces for pointers a kernel reads from memory (the persistent batch kernel's
// argument blocks), where the compiler cannot infer them as it does for kernel arguments:
// global (vector loads and atomics without the flat path's extra waits) and constant
// (wave-uniform reads of tables the kernel never writes become scalar loads); and LDS, for
// a staged table handed to code that also reads tables from memory (LDS reads, not flat).
// (The host pass, which only checks device code, sees them as plain pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
#define WRT_GLOBAL __attribute__((address_space(1)))
#define WRT_CONST __attribute__((address_space(4)))
#define WRT_LDS __attribute__((address_space(3)))
#else
#define WRT_GLOBAL
#define WRT_CONST
#define WRT_LDS
#endif

// fmod(a, d) for a >= 0, d > 0, exact (equal to the C library's fmod).  q = trunc(a/d) is
// the true quotient k or k+1 (a rounded-up quotient); a - k*d is representable, so the
// fma that forms it is exact.
WRT_HOSTDEV double wrt_fmod_pos(double a, double d) {
  const double q = trunc(a / d);
  double r = fma(-q, d, a);
  if (r < 0.0) r = fma(-(q - 1.0), d, a);
  return r;
}

// zero_interp (radon_grid/src/interp.c:63-88): cube test, cell index by truncated quotient
// with the cell width 2/n, index n clamped to n-1.
__device__ __forceinline__ double wrt_zero_interp(const double* __restrict__ vol, int n,
                                                  double dc, double x, double y, double z) {
  if ((fabs(x) > 1.0) || (fabs(y) > 1.0) || (fabs(z) > 1.0)) return 0.0;
  int ix = (int)((x + 1.0) / dc), iy = (int)((y + 1.0) / dc), iz = (int)((z + 1.0) / dc);
  if (ix == n) ix -= 1;
  if (iy == n) iy -= 1;
  if (iz == n) iz -= 1;
  return vol[((size_t)ix * n + iy) * n + iz];
}

WRT_HOSTDEV double wrt_lerp(double t, double a, double b) { return ((1.0 - t) * a + t * b); }

// cube_trilinear_interp (ray_grid/src/interp.c:25-60), including its quirk: the index is a
// truncated quotient, the fraction is fmod/delta.
__device__ __forceinline__ double wrt_trilinear(const double* __restrict__ vol, int n,
                                                double dg, double x, double y, double z) {
  if ((fabs(x) > 1.0) || (fabs(y) > 1.0) || (fabs(z) > 1.0)) return 0.0;
  const double ax = x + 1.0, ay = y + 1.0, az = z + 1.0;
  int ix = (int)(ax / dg), iy = (int)(ay / dg), iz = (int)(az / dg);
  double cx = wrt_fmod_pos(ax, dg) / dg, cy = wrt_fmod_pos(ay, dg) / dg,
         cz = wrt_fmod_pos(az, dg) / dg;
  if (ix == n - 1) { ix -= 1; cx = 1.0; }
  if (iy == n - 1) { iy -= 1; cy = 1.0; }
  if (iz == n - 1) { iz -= 1; cz = 1.0; }
  const size_t nn = (size_t)n * n;
  const double* p = vol + (size_t)ix * nn + (size_t)iy * n + iz;
  const double v1 = p[0], v2 = p[nn], v3 = p[nn + n], v4 = p[n];
  const double v5 = p[1], v6 = p[nn + 1], v7 = p[nn + n + 1], v8 = p[n + 1];
  // cell_trilinear_interp: x-lerps, then y, then z (ray_grid/src/interp.c:3-21)
  const double c1 = wrt_lerp(cy, wrt_lerp(cx, v1, v2), wrt_lerp(cx, v4, v3));
  const double c2 = wrt_lerp(cy, wrt_lerp(cx, v5, v6), wrt_lerp(cx, v8, v7));
  return wrt_lerp(cz, c1, c2);
}

// Interval of a sorted, uniform-ish grid x[0..ns-1]: the largest i <= ns-2 with
// x[i] <= v (0 when v < x[0]); the same index a bisection search returns.
__device__ __forceinline__ int wrt_interval(const double* __restrict__ x, int ns, double x0,
                                            double inv_h, double v) {
  int i = (int)((v - x0) * inv_h);
  i = i < 0 ? 0 : (i > ns - 2 ? ns - 2 : i);
  while (i > 0 && x[i] > v) --i;
  while (i < ns - 2 && x[i + 1] <= v) ++i;
  return i;
}

// Natural cubic spline evaluation on interval i (textbook b, c, d form; c = half the second
// derivative).  Row data y and coefficients c are strided by es.
__device__ __forceinline__ double wrt_spline_eval(const double* __restrict__ x,
                                                  const double* __restrict__ y,
                                                  const double* __restrict__ c, long long es,
                                                  int i, double v) {
  const double xl = x[i], xh = x[i + 1];
  const double yl = y[(long long)i * es], yh = y[(long long)(i + 1) * es];
  const double cl = c[(long long)i * es], ch = c[(long long)(i + 1) * es];
  const double h = xh - xl;
  const double t = v - xl;
  const double b = (yh - yl) / h - h * (ch + 2.0 * cl) / 3.0;
  const double d = (ch - cl) / (3.0 * h);
  return yl + t * (b + t * (cl + t * d));
}

// Newton divided-difference quadratic through three points (GSL "polynomial", size 3).
WRT_HOSTDEV double wrt_poly3(double x0, double x1, double x2, double y0, double y1, double y2,
                             double x) {
  const double d1 = (y1 - y0) / (x1 - x0);
  const double d2a = (y2 - y1) / (x2 - x1);
  const double d2 = (d2a - d1) / (x2 - x0);
  double r = d2;
  r = d1 + (x - x1) * r;
  r = y0 + (x - x0) * r;
  return r;
}

// 64-lane wavefront sum (CDNA wave64).
__device__ __forceinline__ double wrt_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ __forceinline__ float wrt_wave_sumf(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

#define WRT_LAUNCH_CHECK()                                 \
  do {                                                     \
    hipError_t e_ = hipGetLastError();                     \
    if (e_ != hipSuccess) return WRT_E_HIP;                \
  } while (0)
