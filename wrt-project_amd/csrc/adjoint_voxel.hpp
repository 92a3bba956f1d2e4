// adjoint_voxel.hpp -- the FP32 backprojection at one voxel, shared by the per-job launch
// (k_adjoint_fast.hip) and the persistent batch launch (k_batch.hip).
//
// radon_inversion/backprojection/C/src/integral.c:6-24 at the voxel centre
// (-1 + 1/n + i*2/n, radon3d_ajoint.c:248-250): the sum over directions of
// (2pi/nphi) w_theta S(x . n), S the natural cubic spline of the direction's row over the
// shift grid and zero for |s| > 1.  Each interval's cubic is an FP32 float4 rounded from the
// FP64 spline, in the interval's unit variable tt = (s - x_i)/h; s is formed in FP32, and
// u = (s - s0)/ds with one fma gives the interval i (truncated, clamped) and tt = u - i.  An
// interval decision that differs from the FP64 one near a knot changes the value only at
// the rounding level (the spline is C2 across knots); the cut |s| <= 1 is not continuous
// (rows need not vanish at s = +-1), so where the FP32 s lies within 1e-5 of +-1 the
// decision is taken again in FP64 with the reference's expression
// ((x sin t) cos p + (y sin t) sin p) + z cos t.  The sum over directions adds four
// consecutive directions in FP32, then Kahan-compensates the sum of those groups.
#ifndef WRT_ADJOINT_VOXEL_HPP
#define WRT_ADJOINT_VOXEL_HPP
#include "wrt_common.hpp"

// V voxels per thread, k[0..V-1] of the [z][x][y] output (valid[v] = 0 skips one: its
// result is unspecified and must not be stored); dirs / trig / cub as in wrt_batch_item.
// Directions go in fours and every voxel's four cubic gathers are issued before any is
// used, so a thread keeps 4V table loads in flight; each voxel's sum still runs in
// direction order, so the result of a voxel does not depend on V.  The pointer types are
// parameters so that a caller whose tables come from memory (the batch kernel's items) can
// pass them in the constant and global address spaces, which the kernel arguments of the
// per-job launch have by themselves: the wave-uniform direction reads then are scalar
// loads and the cubic gathers global (not flat) loads.
template <int V, class DirP, class TrigP, class CubP>
__device__ __forceinline__ void adjoint_voxels_f32(DirP __restrict__ dirs, TrigP __restrict__ trig,
                                                   int ndir, CubP __restrict__ cub, int ns,
                                                   float s0, float inv_ds, float ds, int n,
                                                   const long long* k, float* out) {
  double xd[V], yd[V], zd[V];
  float x[V], y[V], z[V], acc[V], comp[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const int iy = (int)(k[v] % n);
    const int ix = (int)((k[v] / n) % n);
    const int iz = (int)(k[v] / ((long long)n * n));
    xd[v] = (-1.0) + (1.0 / n) + ix * (2.0 / n);
    yd[v] = (-1.0) + (1.0 / n) + iy * (2.0 / n);
    zd[v] = (-1.0) + (1.0 / n) + iz * (2.0 / n);
    x[v] = (float)xd[v];
    y[v] = (float)yd[v];
    z[v] = (float)zd[v];
    acc[v] = 0.f;
    comp[v] = 0.f;
  }
  const int last = ns - 2;
  const float u0 = -s0 * inv_ds;  // u = s * inv_ds + u0 = (s - s0) / ds
  int d = 0;
  for (; d + 4 <= ndir; d += 4) {
    float4 D[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) D[u] = dirs[d + u];  // uniform across the wavefront
    float sv[V][4];
    int iv[V][4];
    bool inv[V][4];
    float4 Cv[V][4];
#pragma unroll
    for (int v = 0; v < V; ++v)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float s = fmaf(x[v], D[u].x, fmaf(y[v], D[u].y, z[v] * D[u].z));
        bool in = fabsf(s) <= 1.0f;
        if (fabsf(fabsf(s) - 1.0f) <= 1e-5f) {
          const double4 T = trig[d + u];
          const double sd = xd[v] * T.x * T.y + yd[v] * T.x * T.z + zd[v] * T.w;
          in = !(fabs(sd) > 1.0);
        }
        const float uu = fmaf(s, inv_ds, u0);
        int i = (int)uu;
        i = i < 0 ? 0 : (i > last ? last : i);  // a valid interval even where s is cut
        sv[v][u] = uu - (float)i;  // tt in [0, 1] where the sample is kept
        iv[v][u] = i;
        inv[v][u] = in;
      }
#pragma unroll
    for (int v = 0; v < V; ++v)
#pragma unroll
      for (int u = 0; u < 4; ++u) Cv[v][u] = cub[(long long)(d + u) * (ns - 1) + iv[v][u]];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      float g = 0.f;  // the four directions' terms, in direction order
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float t = sv[v][u];
        const float4 C = Cv[v][u];
        const float val = fmaf(fmaf(fmaf(C.w, t, C.z), t, C.y), t, C.x);
        g = fmaf(D[u].w, inv[v][u] ? val : 0.f, g);
      }
      const float yk = g - comp[v];
      const float tk = acc[v] + yk;
      comp[v] = (tk - acc[v]) - yk;
      acc[v] = tk;
    }
  }
  for (; d < ndir; ++d) {
    const float4 D = dirs[d];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const float s = fmaf(x[v], D.x, fmaf(y[v], D.y, z[v] * D.z));
      bool in = fabsf(s) <= 1.0f;
      if (fabsf(fabsf(s) - 1.0f) <= 1e-5f) {
        const double4 T = trig[d];
        const double sd = xd[v] * T.x * T.y + yd[v] * T.x * T.z + zd[v] * T.w;
        in = !(fabs(sd) > 1.0);
      }
      if (in) {
        const float uu = fmaf(s, inv_ds, u0);
        int i = (int)uu;
        i = i < 0 ? 0 : (i > last ? last : i);
        const float t = uu - (float)i;
        const float4 C = cub[(long long)d * (ns - 1) + i];
        const float val = fmaf(fmaf(fmaf(C.w, t, C.z), t, C.y), t, C.x);
        const float yk = fmaf(D.w, val, -comp[v]);
        const float tk = acc[v] + yk;
        comp[v] = (tk - acc[v]) - yk;
        acc[v] = tk;
      }
    }
  }
#pragma unroll
  for (int v = 0; v < V; ++v) out[v] = acc[v];
}

#endif
