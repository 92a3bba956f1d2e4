// adjoint_voxel.hpp -- the FP32 backprojection at one voxel, shared by the per-job launch
// (k_adjoint_fast.hip) and the persistent batch launch (k_batch.hip).
//
// radon_inversion/backprojection/C/src/integral.c:6-24 at the voxel centre
// (-1 + 1/n + i*2/n, radon3d_ajoint.c:248-250): the sum over directions of
// (2pi/nphi) w_theta S(x . n), S the natural cubic spline of the direction's row over the
// shift grid and zero for |s| > 1.  Each interval's cubic (y, b, c, d) is an FP32 float4
// rounded from the FP64 spline; s is formed in FP32.  An interval decision that differs
// from the FP64 one near a knot changes the value only at the rounding level (the spline is
// C2 across knots); the cut |s| <= 1 is not continuous (rows need not vanish at s = +-1), so
// where the FP32 s lies within 1e-5 of +-1 the decision is taken again in FP64 with the
// reference's expression ((x sin t) cos p + (y sin t) sin p) + z cos t.  The sum over
// directions is Kahan-compensated in FP32.
#ifndef WRT_ADJOINT_VOXEL_HPP
#define WRT_ADJOINT_VOXEL_HPP
#include "wrt_common.hpp"

// voxel k of the [z][x][y] output; dirs / trig / cub as in wrt_batch_item
__device__ __forceinline__ float adjoint_voxel_f32(const float4* __restrict__ dirs,
                                                   const double4* __restrict__ trig, int ndir,
                                                   const float4* __restrict__ cub, int ns,
                                                   float s0, float inv_ds, float ds, int n,
                                                   long long k) {
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
  return acc;
}

#endif
