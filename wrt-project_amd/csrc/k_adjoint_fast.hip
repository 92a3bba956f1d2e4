// k_adjoint_fast.hip -- FP32 form of the backprojection radon3d_adjoint
// (radon_inversion/backprojection/C/src/integral.c:6-24): the per-interval FP32 cubics of
// the FP64 spline rows, and the per-job launch of the voxel body (adjoint_voxel.hpp).
#include "adjoint_voxel.hpp"
#include "wrt_dev.h"
#include "wrt_fast.h"

// (y, c) packed FP64 rows [row][ns][2] -> per-interval FP32 cubics [row][ns-1] in the
// interval's unit variable tt = (s - x_i) / h: y + (b h) tt + (c h^2) tt^2 + (d h^3) tt^3
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
    cub[k] = make_float4((float)yl, (float)(b * h), (float)(cl * h * h), (float)(d * h * h * h));
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

// voxels per thread of the per-job launch: the persistent kernel's, whose tiling the job's
// tile order lists
#define ADJ_VPT WRT_ADJ_VPT

// one tile of 4 x 4 x 4*ADJ_VPT voxels per wavefront (ADJ_VPT voxels per thread), four per
// workgroup
__global__ __launch_bounds__(256) void k_adjoint_f32(const float4* __restrict__ dirs,
                                                     const double4* __restrict__ trig, int ndir,
                                                     const float4* __restrict__ cub, int ns,
                                                     float s0, float inv_ds, float ds, int n,
                                                     long long ntiles, float* __restrict__ out) {
  const long long tile = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile >= ntiles) return;  // the whole wavefront
  long long k[ADJ_VPT];
  bool valid[ADJ_VPT];
  adjoint_tile_voxels<ADJ_VPT>(tile, threadIdx.x & 63, n, k, valid);
  float r[ADJ_VPT];
  adjoint_voxels_f32<ADJ_VPT>(dirs, trig, ndir, cub, ns, s0, inv_ds, ds, n, k, r);
#pragma unroll
  for (int v = 0; v < ADJ_VPT; ++v)
    if (valid[v]) out[k[v]] = r[v];
}

// the LDS-staged form (adjoint_staged_tiles): one tile per wavefront, 16 per workgroup
__global__ __launch_bounds__(ADJ_LDS_WAVES * 64) void k_adjoint_f32_lds(
    const float4* __restrict__ dirs, const double4* __restrict__ trig, int ndir,
    const float4* __restrict__ cub, int ns, float s0, float inv_ds, float ds, int n,
    const int* __restrict__ order, long long ntiles, int chunk, float* __restrict__ out) {
  __shared__ float4 tab[ADJ_LDS_F4];
  const long long tile0 = (long long)blockIdx.x * ADJ_LDS_WAVES;
  adjoint_staged_tiles<ADJ_VPT, true, 8>(dirs, trig, ndir, cub, ns, s0, inv_ds, ds, n, order,
                                         tile0, (int)min((long long)ADJ_LDS_WAVES, ntiles - tile0),
                                         chunk, (WRT_LDS float4*)tab, out);
}

extern "C" int wrtk_adjoint_f32(const void* dirs, const void* trig, int ndir, const void* cub,
                                int ns, double s0, double ds, int n, const int* order, float* out,
                                hipStream_t st) {
  if (n <= 0) return WRT_OK;
  const long long ntiles = wrt_adj_tiles(n, ADJ_VPT);
  // directions per chunk: a multiple of 4 whose rows fit the LDS table
  const int chunk = adjoint_chunk(ns, 8);
  if (chunk >= 4) {
    k_adjoint_f32_lds<<<dim3((unsigned)((ntiles + ADJ_LDS_WAVES - 1) / ADJ_LDS_WAVES)),
                        ADJ_LDS_WAVES * 64, 0, st>>>(
        (const float4*)dirs, (const double4*)trig, ndir, (const float4*)cub, ns, (float)s0,
        (float)(1.0 / ds), (float)ds, n, order, ntiles, chunk, out);
  } else {  // rows longer than 2048 intervals
    k_adjoint_f32<<<dim3((unsigned)((ntiles + 3) / 4)), 256, 0, st>>>(
        (const float4*)dirs, (const double4*)trig, ndir, (const float4*)cub, ns, (float)s0,
        (float)(1.0 / ds), (float)ds, n, ntiles, out);
  }
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}
