// k_adjoint_fast.hip -- FP32 form of the backprojection radon3d_adjoint
// (radon_inversion/backprojection/C/src/integral.c:6-24): at every voxel centre, the sum over
// directions of (2pi/nphi) w_theta S_{phi,theta}(x . n), with S the natural cubic spline of the
// direction's Radon row over the shift grid (zero for |s| > 1).
//
// The spline is solved in FP64 (k_spline_rows); here each interval's cubic is rounded to FP32
// as (y_i, b_i, c_i, d_i) so that an evaluation is one 16-byte load and three FMAs.  s is
// formed in FP32; an interval decision that differs from the FP64 one near a knot changes the
// value only at the rounding level, because the spline is C2 across knots.  The cut |s| <= 1
// is not continuous (rows need not vanish at s = +-1), so wherever the FP32 s lies within
// 1e-5 of +-1 the decision is taken again in FP64 with the reference's expression
// ((x sin t) cos p + (y sin t) sin p) + z cos t.  Sums over directions are Kahan-compensated
// in FP32.
#include "wrt_common.hpp"
#include "wrt_dev.h"

// (y, c) packed FP64 rows [row][ns][2] -> per-interval FP32 cubics [row][ns-1] (y, b, c, d)
__global__ __launch_bounds__(256) void k_adjoint_cubics(const double* __restrict__ x, int ns,
                                                        const double* __restrict__ tab,
                                                        float4* __restrict__ cub, long long rows) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = rows * (ns - 1);
  for (long long k = t; k < total; k += (long long)gridDim.x * blockDim.x) {
    const long long r = k / (ns - 1);
    const int i = (int)(k % (ns - 1));
    const double* row = tab + r * 2LL * ns;
    const double yl = row[2 * i], cl = row[2 * i + 1], yh = row[2 * i + 2], ch = row[2 * i + 3];
    const double h = x[i + 1] - x[i];
    const double b = (yh - yl) / h - h * (ch + 2.0 * cl) / 3.0;
    const double d = (ch - cl) / (3.0 * h);
    cub[k] = make_float4((float)yl, (float)b, (float)cl, (float)d);
  }
}

extern "C" int wrtk_adjoint_cubics(const double* x, int ns, const double* tab, void* cub,
                                   long long rows, hipStream_t st) {
  const long long total = rows * (ns - 1);
  if (total <= 0) return WRT_OK;
  long long blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  k_adjoint_cubics<<<(unsigned)blocks, 256, 0, st>>>(x, ns, tab, (float4*)cub, rows);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

// dirs[k] = (n.x, n.y, n.z, weight) for direction k = phi*ntheta + theta, in FP32;
// trig[k] = (sin t, cos p, sin p, cos t) in FP64, read only for the cut decisions near |s| = 1.
// cub: [ndir][ns-1] float4 cubics over the uniform shift grid s_i = s0 + i*ds.
__global__ __launch_bounds__(256) void k_adjoint_f32(const float4* __restrict__ dirs,
                                                     const double4* __restrict__ trig, int ndir,
                                                     const float4* __restrict__ cub, int ns,
                                                     float s0, float inv_ds, float ds, int n,
                                                     float* __restrict__ out) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)n * n * n;
  if (k >= total) return;
  const int iy = (int)(k % n);
  const int ix = (int)((k / n) % n);
  const int iz = (int)(k / ((long long)n * n));
  const double xd = (-1.0) + (1.0 / n) + ix * (2.0 / n);
  const double yd = (-1.0) + (1.0 / n) + iy * (2.0 / n);
  const double zd = (-1.0) + (1.0 / n) + iz * (2.0 / n);
  const float x = (float)xd, y = (float)yd, z = (float)zd;
  const int last = ns - 2;
  float acc = 0.f, comp = 0.f;
  for (int d = 0; d < ndir; ++d) {
    const float4 D = dirs[d];  // uniform across the wavefront
    const float s = fmaf(x, D.x, fmaf(y, D.y, z * D.z));
    bool in = fabsf(s) <= 1.0f;
    if (fabsf(fabsf(s) - 1.0f) <= 1e-5f) {
      const double4 T = trig[d];
      const double sd = xd * T.x * T.y + yd * T.x * T.z + zd * T.w;
      in = !(fabs(sd) > 1.0);
    }
    if (in) {
      int i = (int)((s - s0) * inv_ds);
      i = i < 0 ? 0 : (i > last ? last : i);
      const float t = s - fmaf((float)i, ds, s0);
      const float4 C = cub[(long long)d * (ns - 1) + i];
      const float v = fmaf(fmaf(fmaf(C.w, t, C.z), t, C.y), t, C.x);
      const float yk = fmaf(D.w, v, -comp);
      const float tk = acc + yk;
      comp = (tk - acc) - yk;
      acc = tk;
    }
  }
  out[k] = acc;
}

extern "C" int wrtk_adjoint_f32(const void* dirs, const void* trig, int ndir, const void* cub,
                                int ns, double s0, double ds, int n, float* out, hipStream_t st) {
  const long long total = (long long)n * n * n;
  if (total <= 0) return WRT_OK;
  k_adjoint_f32<<<dim3((unsigned)((total + 255) / 256)), 256, 0, st>>>(
      (const float4*)dirs, (const double4*)trig, ndir, (const float4*)cub, ns, (float)s0, (float)(1.0 / ds), (float)ds,
      n, out);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}
