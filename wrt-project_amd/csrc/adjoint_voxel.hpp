// adjoint_voxel.hpp -- the FP32 backprojection at one voxel, shared by the per-job launch
// (k_adjoint_fast.hip) and the persistent batch launch (k_batch.hip).
//
// radon_inversion/backprojection/C/src/integral.c:6-24 at the voxel centre
// (-1 + 1/n + i*2/n, radon3d_ajoint.c:248-250): the sum over directions of
// (2pi/nphi) w_theta S(x . n), S the natural cubic spline of the direction's row over the
// shift grid and zero for |s| > 1.  Each interval's cubic is an FP32 float4 rounded from the
// FP64 spline, in the interval's unit variable tt = (s - x_i)/h; s is formed in FP32 as
// z n_z + (x n_x + y n_y) (the bracket shared by a thread's voxels, which differ in z), and
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

#include <type_traits>

// The voxels go to wavefronts in tiles of 4 (x) x 4 (y) x 4V (z): lane l takes x offset
// (l >> 2) & 3, y offset l & 3 and z offsets (l >> 4) + 4v, v < V.  A tile is compact, so
// for each direction its voxels' shifts fall in a few neighbouring spline intervals (the
// wavefront's cubic gathers touch few cache lines), and most tiles lie wholly inside or
// wholly outside the unit ball.  Tiles are numbered z-major like the output:
// ((tz * ntx) + tx) * ntx + ty with ntx = ceil(n / 4) tiles along x and y and
// ceil(n / 4V) along z (wrt_adj_tiles).  A tile past the volume's edge clamps its coordinates
// into the volume (computed, not stored: valid[v] = false).
template <int V>
__device__ __forceinline__ void adjoint_tile_voxels(long long tile, int lane, int n, long long* k,
                                                    bool* valid) {
  const int ntx = (n + 3) >> 2;
  const long long txy = (long long)ntx * ntx;
  const int tz = (int)(tile / txy);
  const int rem = (int)(tile - (long long)tz * txy);
  const int tx = rem / ntx, ty = rem - tx * ntx;
  const int ix = 4 * tx + ((lane >> 2) & 3), iy = 4 * ty + (lane & 3);
#pragma unroll
  for (int v = 0; v < V; ++v) {
    const int iz = 4 * V * tz + (lane >> 4) + 4 * v;
    valid[v] = ix < n && iy < n && iz < n;
    k[v] = ((long long)min(iz, n - 1) * n + min(ix, n - 1)) * n + min(iy, n - 1);
  }
}

// The sum over directions for V voxels per thread.  MODE 0: the general case.  MODE >= 1
// (INSIDE): every voxel of the wavefront lies within radius 1 - 1e-4 of the origin, where
// |s| <= 1 holds for every direction and |s| is farther than 1e-5 from 1 (the FP32 s errs by
// under 1e-6), so the cut and its FP64 re-decision are skipped.  MODE 2: besides, the shift
// grid covers [-1 + 5e-5, 1 - 5e-5], so u = (s - s0)/ds of such an s lies in [0, ns - 1) and
// needs no clamp: the interval is trunc(u) and tt = fract(u) (exact: u - floor(u) is
// representable).  Every mode gives the same result, bit for bit, with fewer instructions.
//
// The directions [d_begin, d_end) of the sum, with the running sums acc and their Kahan
// compensations comp carried in and out: d_begin is a multiple of 4, and d_end is too unless
// it is ndir (the single directions past the last group of four are added only there), so
// that a sum taken in several ranges is the sum taken at once, bit for bit.  The cubic row
// of direction d is cub + (d - cub_d0) * (ns - 1): the whole table (cub_d0 = 0), or a chunk
// of rows staged in LDS that starts at direction cub_d0.
template <int V, int MODE, class DirP, class TrigP, class CubP>
__device__ __forceinline__ void adjoint_dir_range(DirP __restrict__ dirs, TrigP __restrict__ trig,
                                                  int d_begin, int d_end, int ndir,
                                                  CubP __restrict__ cub, int cub_d0, int ns,
                                                  float s0, float inv_ds, bool covers,
                                                  double xd, double yd, const double* zd,
                                                  float x, float y, const float* z, float* acc,
                                                  float* comp) {
  const int last = ns - 2;
  const float lastf = (float)last;  // exact: ns < 2^24
  const float u0 = -s0 * inv_ds;  // u = s * inv_ds + u0 = (s - s0) / ds
  int d = d_begin;
  for (; d + 4 <= d_end; d += 4) {
    float4 D[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) D[u] = dirs[d + u];  // uniform across the wavefront
    float sa[V][4];
    float sxy[4];  // x n_x + y n_y, shared by the thread's voxels
#pragma unroll
    for (int u = 0; u < 4; ++u) sxy[u] = fmaf(x, D[u].x, y * D[u].y);
#pragma unroll
    for (int v = 0; v < V; ++v)
#pragma unroll
      for (int u = 0; u < 4; ++u) sa[v][u] = fmaf(z[v], D[u].z, sxy[u]);
    // the four directions in mode EM (the modes above; EM >= 1 only where every s of the
    // wavefront is in range and far from the cut)
    auto group = [&](auto em_tag) {
      constexpr int EM = decltype(em_tag)::value;
      constexpr bool INSIDE = EM >= 1;
      float sv[V][4];
      int iv[V][4];
      bool inv[V][4];
      float4 Cv[V][4];
#pragma unroll
      for (int v = 0; v < V; ++v)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float s = sa[v][u];
          bool in = true;
          if (!INSIDE) {
            in = fabsf(s) <= 1.0f;
            // the FP64 re-decision as a branch of the whole wave on the ballot (a scalar
            // branch, no exec mask saved and restored per evaluation); each lane takes its
            // result only if its s is near the cut
            const bool nr = fabsf(fabsf(s) - 1.0f) <= 1e-5f;
            if (__builtin_expect(__builtin_amdgcn_ballot_w64(nr) != 0ull, 0)) {
              const double4 T = trig[d + u];
              const double sd = xd * T.x * T.y + yd * T.x * T.z + zd[v] * T.w;
              in = nr ? !(fabs(sd) > 1.0) : in;
            }
          }
          const float uu = fmaf(s, inv_ds, u0);
          if (EM == 2) {
            iv[v][u] = (int)uu;
            sv[v][u] = __builtin_amdgcn_fractf(uu);
          } else {
            // the interval, truncated and clamped into [0, last] (a valid interval even where
            // s is cut): truncating the clamped uu is the same, in one med3 and one conversion
            const int i = (int)__builtin_amdgcn_fmed3f(uu, 0.0f, lastf);
            sv[v][u] = uu - (float)i;  // tt in [0, 1] where the sample is kept
            iv[v][u] = i;
          }
          inv[v][u] = in;
        }
      // each direction's row of cubics from its scalar base, the interval as a 32-bit offset
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const auto row = cub + (long long)(d + u - cub_d0) * (ns - 1);
#pragma unroll
        for (int v = 0; v < V; ++v) Cv[v][u] = row[(unsigned)iv[v][u]];
      }
#pragma unroll
      for (int v = 0; v < V; ++v) {
        float g = 0.f;  // the four directions' terms, in direction order
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float t = sv[v][u];
          const float4 C = Cv[v][u];
          const float val = fmaf(fmaf(fmaf(C.w, t, C.z), t, C.y), t, C.x);
          g = fmaf(D[u].w, INSIDE || inv[v][u] ? val : 0.f, g);
        }
        const float yk = g - comp[v];
        const float tk = acc[v] + yk;
        comp[v] = (tk - acc[v]) - yk;
        acc[v] = tk;
      }
    };
    if constexpr (MODE >= 1) {
      group(std::integral_constant<int, MODE>());
    } else {
      // In the general mode a group whose every s (FP32, error under 1e-6) lies within
      // 1 - 2e-4 of 0 keeps every sample, is not near the cut, and on a covering grid has u
      // in [0, ns - 1): it takes mode 2's instructions, with the same result bit for bit.
      // A group whose every s lies beyond 1 + 1e-5 (so the cut holds, FP64 decision
      // included) adds g = 0: only its Kahan step with g = 0 is taken.
      bool far = covers, out = true;
#pragma unroll
      for (int v = 0; v < V; ++v)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          far = far && fabsf(sa[v][u]) < 1.0f - 2e-4f;
          out = out && fabsf(sa[v][u]) > 1.0f + 1e-5f;
        }
      if (__builtin_amdgcn_ballot_w64(!far) == 0ull) {
        group(std::integral_constant<int, 2>());
      } else if (__builtin_amdgcn_ballot_w64(!out) == 0ull) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const float yk = 0.f - comp[v];
          const float tk = acc[v] + yk;
          comp[v] = (tk - acc[v]) - yk;
          acc[v] = tk;
        }
      } else {
        group(std::integral_constant<int, 0>());
      }
    }
  }
  if (d_end < ndir) return;
  constexpr bool INSIDE = MODE >= 1;
  for (; d < ndir; ++d) {
    const float4 D = dirs[d];
    const float sxy = fmaf(x, D.x, y * D.y);
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const float s = fmaf(z[v], D.z, sxy);
      bool in = true;
      if (!INSIDE) {
        in = fabsf(s) <= 1.0f;
        const bool nr = fabsf(fabsf(s) - 1.0f) <= 1e-5f;
        if (__builtin_expect(__builtin_amdgcn_ballot_w64(nr) != 0ull, 0)) {
          const double4 T = trig[d];
          const double sd = xd * T.x * T.y + yd * T.x * T.z + zd[v] * T.w;
          in = nr ? !(fabs(sd) > 1.0) : in;
        }
      }
      if (in) {
        const float uu = fmaf(s, inv_ds, u0);
        const int i = (int)__builtin_amdgcn_fmed3f(uu, 0.0f, lastf);
        const float t = uu - (float)i;
        const float4 C = cub[(long long)(d - cub_d0) * (ns - 1) + (unsigned)i];
        const float val = fmaf(fmaf(fmaf(C.w, t, C.z), t, C.y), t, C.x);
        const float yk = fmaf(D.w, val, -comp[v]);
        const float tk = acc[v] + yk;
        comp[v] = (tk - acc[v]) - yk;
        acc[v] = tk;
      }
    }
  }
}

// The voxels' coordinates (the thread's voxels share x and y: the tiles give them one column
// of z) and the mode of the wavefront: 2 if every voxel lies within radius 1 - 1e-4 and the
// grid covers [-1 + 5e-5, 1 - 5e-5], 1 if only the former (and MODE1), else 0.
template <int V, bool MODE1 = true>
struct adjoint_voxel_set {
  double xd, yd, zd[V];
  float x, y, z[V];
  int mode;
  bool covers;
  __device__ __forceinline__ adjoint_voxel_set(const long long* k, int n, int ns, float s0,
                                               float ds) {
    const int iy = (int)(k[0] % n);
    const int ix = (int)((k[0] / n) % n);
    xd = (-1.0) + (1.0 / n) + ix * (2.0 / n);
    yd = (-1.0) + (1.0 / n) + iy * (2.0 / n);
    x = (float)xd;
    y = (float)yd;
    bool inside = true;
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int iz = (int)(k[v] / ((long long)n * n));
      zd[v] = (-1.0) + (1.0 / n) + iz * (2.0 / n);
      z[v] = (float)zd[v];
      inside = inside && xd * xd + yd * yd + zd[v] * zd[v] <= (1.0 - 1e-4) * (1.0 - 1e-4);
    }
    // the grid covers [-1 + 5e-5, 1 - 5e-5] (wave-uniform)
    covers = s0 <= -1.0f + 5e-5f && fmaf((float)(ns - 1), ds, s0) >= 1.0f - 5e-5f;
    const bool all_inside = __builtin_amdgcn_ballot_w64(!inside) == 0ull;
    mode = all_inside && covers ? 2 : (MODE1 && all_inside ? 1 : 0);
  }
  // the directions [d_begin, d_end) in the wavefront's mode
  template <class DirP, class TrigP, class CubP>
  __device__ __forceinline__ void range(DirP dirs, TrigP trig, int d_begin, int d_end, int ndir,
                                        CubP cub, int cub_d0, int ns, float s0, float inv_ds,
                                        float* acc, float* comp) const {
    if (mode == 2)
      adjoint_dir_range<V, 2>(dirs, trig, d_begin, d_end, ndir, cub, cub_d0, ns, s0, inv_ds,
                              covers, xd, yd, zd, x, y, z, acc, comp);
    else if (MODE1 && mode == 1)
      adjoint_dir_range<V, 1>(dirs, trig, d_begin, d_end, ndir, cub, cub_d0, ns, s0, inv_ds,
                              covers, xd, yd, zd, x, y, z, acc, comp);
    else
      adjoint_dir_range<V, 0>(dirs, trig, d_begin, d_end, ndir, cub, cub_d0, ns, s0, inv_ds,
                              covers, xd, yd, zd, x, y, z, acc, comp);
  }
};

// V voxels per thread, k[0..V-1] of the [z][x][y] output, all with the x and y of k[0] (as
// adjoint_tile_voxels gives them); dirs / trig / cub as in wrt_batch_item.  Called by all 64
// lanes of a wavefront (the inside test is a ballot).
// Directions go in fours and every voxel's four cubic gathers are issued before any is
// used, so a thread keeps 4V table loads in flight; each voxel's sum still runs in
// direction order, so the result of a voxel does not depend on V or on the tiling.  The
// pointer types are parameters so that a caller whose tables come from memory (the batch
// kernel's items) can pass them in the constant and global address spaces, which the
// kernel arguments of the per-job launch have by themselves: the wave-uniform direction
// reads then are scalar loads and the cubic gathers global (not flat) loads.
// MODE1 = false leaves out mode 1 (inside tiles on a grid that does not cover [-1, 1] take
// the general mode), for a caller short of registers.
template <int V, bool MODE1 = true, class DirP, class TrigP, class CubP>
__device__ __forceinline__ void adjoint_voxels_f32(DirP __restrict__ dirs, TrigP __restrict__ trig,
                                                   int ndir, CubP __restrict__ cub, int ns,
                                                   float s0, float inv_ds, float ds, int n,
                                                   const long long* k, float* out) {
  const adjoint_voxel_set<V, MODE1> vs(k, n, ns, s0, ds);
  float acc[V], comp[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    acc[v] = 0.f;
    comp[v] = 0.f;
  }
  vs.range(dirs, trig, 0, ndir, ndir, cub, 0, ns, s0, inv_ds, acc, comp);
#pragma unroll
  for (int v = 0; v < V; ++v) out[v] = acc[v];
}

// The LDS-staged form over a workgroup of ADJ_LDS_WAVES wavefronts (all of its threads call
// it): wavefront w < ntile takes tile order[tile0 + w], and the workgroup walks the directions in
// chunks of `chunk` directions (a multiple of 4 with chunk * (ns - 1) <= PF * 1024) whose
// cubic rows it copies into tab, so that every evaluation's 16-byte gather is an LDS read
// (gathers from memory, one vector load through the CU's L1 per evaluation, bound the
// unstaged form).  Each thread loads its PF float4 of the next chunk into registers while
// the current one is evaluated and stores them between two barriers.  The sums are the
// unstaged form's, bit for bit (adjoint_dir_range).  Returns with the barriers passed and
// the outputs stored; a caller that reuses tab waits first.
#define ADJ_LDS_F4 8192
#define ADJ_LDS_WAVES 16
template <int V, bool MODE1, int PF, class DirP, class TrigP, class CubP, class OrdP,
          class OutP>
__device__ __forceinline__ void adjoint_staged_tiles(DirP dirs, TrigP trig, int ndir, CubP cub,
                                                     int ns, float s0, float inv_ds, float ds,
                                                     int n, OrdP order, long long tile0,
                                                     int ntile, int chunk, WRT_LDS float4* tab,
                                                     OutP out) {
  constexpr int STRIDE = ADJ_LDS_WAVES * 64;
  static_assert(PF * STRIDE <= ADJ_LDS_F4, "a chunk fits the table");
  const int tid = threadIdx.x;
  const int w = tid >> 6;
  // wave-uniform: a wavefront past the last tile stages and waits with the others
  const bool active = w < ntile;
  long long k[V];
  bool valid[V];
  adjoint_tile_voxels<V>(order[tile0 + (active ? w : 0)], tid & 63, n, k, valid);
  const adjoint_voxel_set<V, MODE1> vs(k, n, ns, s0, ds);
  float acc[V], comp[V];
#pragma unroll
  for (int v = 0; v < V; ++v) {
    acc[v] = 0.f;
    comp[v] = 0.f;
  }
  const int row = ns - 1;
  // the thread's share of a chunk: elements tid + p * 1024, p < PF, in registers q0..q7
  // (named, not an array, which the compiler keeps in scratch memory); an index past the
  // chunk reads its last element and is not stored
  static_assert(PF == 4 || PF == 8, "4 or 8 float4 per thread and chunk");
  float4 q0, q1, q2, q3, q4, q5, q6, q7;
  auto fetch = [&](decltype(cub) src, int cnt) {
    q0 = src[min(tid + 0 * STRIDE, cnt - 1)];
    q1 = src[min(tid + 1 * STRIDE, cnt - 1)];
    q2 = src[min(tid + 2 * STRIDE, cnt - 1)];
    q3 = src[min(tid + 3 * STRIDE, cnt - 1)];
    if constexpr (PF == 8) {
      q4 = src[min(tid + 4 * STRIDE, cnt - 1)];
      q5 = src[min(tid + 5 * STRIDE, cnt - 1)];
      q6 = src[min(tid + 6 * STRIDE, cnt - 1)];
      q7 = src[min(tid + 7 * STRIDE, cnt - 1)];
    }
  };
  auto store = [&](int p, const float4& v, int cnt) {
    if (tid + p * STRIDE < cnt) tab[tid + p * STRIDE] = v;
  };
  fetch(cub, min(chunk, ndir) * row);
  for (int d0 = 0; d0 < ndir; d0 += chunk) {
    const int d1 = min(d0 + chunk, ndir);
    const int cnt = (d1 - d0) * row;
    __syncthreads();  // every wavefront is done with the previous chunk
    store(0, q0, cnt);
    store(1, q1, cnt);
    store(2, q2, cnt);
    store(3, q3, cnt);
    if constexpr (PF == 8) {
      store(4, q4, cnt);
      store(5, q5, cnt);
      store(6, q6, cnt);
      store(7, q7, cnt);
    }
    __syncthreads();
    // the next chunk, in flight while this one is evaluated
    if (d1 < ndir) fetch(cub + (long long)d1 * row, (min(d1 + chunk, ndir) - d1) * row);
    if (active)
      vs.range(dirs, trig, d0, d1, ndir, (const WRT_LDS float4*)tab, d0, ns, s0, inv_ds, acc,
               comp);
  }
  if (active) {
#pragma unroll
    for (int v = 0; v < V; ++v)
      if (valid[v]) out[k[v]] = acc[v];
  }
}

// directions per chunk of the staged form with PF float4 per thread, for rows of ns shifts
// (0: the rows do not fit)
__host__ __device__ inline int adjoint_chunk(int ns, int pf) {
  return ns < 2 ? 0 : 4 * (pf * ADJ_LDS_WAVES * 64 / (4 * (ns - 1)));
}

#endif
