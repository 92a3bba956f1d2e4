// radon_brick.hpp -- the brick task of the fast plane integrals (radon_grid, FP32 or BF16
// values), shared by the per-job launches (k_fast.hip) and the persistent batch launch
// (k_batch.hip).
//
// Work unit: one workgroup per (non-empty 32^3 brick, group of directions).  The brick is
// staged once into LDS as FP32; every plane (shift, direction) that cuts the brick is a
// "section".  A wavefront walks a section in bands of BR rows (BR = 8 by default; 4 or 16,
// with BC = 64 / BR columns per step): lane (r, k) = (lane / BC, lane % BC) handles lattice
// row i1 = band + r and columns i2 = c0 + k, c0 stepping by BC across the row's
// conservative column interval.  Per sample the cell index is decided exactly as
// radon_grid/src/integral.c:47-51 + interp.c:70-72 decide it in binary64
//   x = ((s*n) + (a*i1)) + (b*i2),  ix = (int)((x + 1) / (2/n)),  index n -> n-1,
// (the per-column products b*i2 are formed in registers, the division is an exact
// power-of-two scaling or a corrected reciprocal), so only the value rounding (FP32/BF16)
// and the summation order differ from the plain form.  Samples whose cell lies in the
// brick add its FP32 value to the lane's FP32 sum; a section's lane sums are reduced over the
// wavefront in FP32, rounded to the plane's 64-bit fixed-point unit and added to the plane's
// integer accumulator (one atomic per section and brick).
//
// Transpose mode works on half bricks (32 x 32 x 16 cells) of 64-bit integer accumulators:
// every sample whose cell lies in the half brick adds the section's value, quantized to a
// fixed point q = rint(g * scale), to the cell with an LDS 64-bit integer atomic.  Integer
// sums are exact and associative, so the result does not depend on the order in which waves
// reach a cell; the scale keeps every cell's total within 2^62 (the caller's bound).  (On
// gfx950 ds_add_f32 runs lane by lane, 193 clocks per wave instruction per CU against 7 for
// ds_add_u64 and 15 for a read + ds_cmpst_rtn_b32 loop, distinct addresses.)
#ifndef WRT_RADON_BRICK_HPP
#define WRT_RADON_BRICK_HPP
#include "wrt_common.hpp"
#include "wrt_fast.h"

#include <limits.h>

#define FB WRT_BRICK
#define FR_WAVES 16
// skewed LDS pitches: bank = (5*lx + ly + lz) mod 32, so a small patch of a plane spreads
// over the banks for every orientation except a few measure-zero ones
#define FR_PY 33
#define FR_PX (33 * 32 + 5)

static_assert(FB == 32, "the band walk assumes 32^3 bricks");

// RN(u / d) for d = RN(2/n) when n is not a power of two: reciprocal estimate plus one
// exact-residual correction (Markstein); equal to the IEEE quotient for these operands
// (checked on the host against the C division for every n <= 1024 in the tests).
__device__ __forceinline__ double div_rn(double u, double d, double inv) {
  const double y = u * inv;
  const double r = fma(-y, d, u);
  return fma(r, inv, y);
}

// a*b + c on the 24-bit multiplier (full rate); a, b < 2^24, b wave-uniform (held in an
// SGPR, which keeps loop-invariant constants out of the VGPR budget).  Written out because
// the compiler turns small known-range products into the quarter-rate 32-bit multiply.
__device__ __forceinline__ unsigned mad_u24(unsigned a, unsigned b, unsigned c) {
  unsigned r;
  asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
  return r;
}

// floor(x) and floor(-x) converted to int32 in one instruction each (x within int32 range)
__device__ __forceinline__ int cvt_floor(float x) {
  int r;
  asm("v_cvt_flr_i32_f32_e32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}
__device__ __forceinline__ int cvt_floor_neg(float x) {
  int r;
  asm("v_cvt_flr_i32_f32_e64 %0, -%1" : "=v"(r) : "v"(x));
  return r;
}

// max over each aligned group of G (4, 8 or 16) lanes, every lane of the group ends with it:
// quad_perm [1,0,3,2] (xor 1), quad_perm [2,3,0,1] (xor 2), then row_half_mirror (lane i of
// a half row with lane 7 - i, which lies in the other quad) and row_mirror (lane i of a row
// with lane 15 - i, in the other half row).  Every lane has a source lane in these patterns,
// so the moves need no old value and each step combines into one DPP max.
template <int CTRL>
__device__ __forceinline__ int dpp_perm(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, true);
}
template <int G>
__device__ __forceinline__ int group_max(int v) {
  static_assert(G == 4 || G == 8 || G == 16, "groups of 4, 8 or 16 lanes");
  v = max(v, dpp_perm<0xB1>(v));
  v = max(v, dpp_perm<0x4E>(v));
  if constexpr (G >= 8) v = max(v, dpp_perm<0x141>(v));
  if constexpr (G >= 16) v = max(v, dpp_perm<0x140>(v));
  return v;
}

// Wavefront sum of a float with DPP lane moves (no LDS round trips): Hillis-Steele scan
// within 16-lane rows (row_shr 1, 2, 4, 8, zero fill), then row broadcasts 15 and 31; the
// total ends in lane 63.  Each step is one DPP add; six roundings of 2^-24 on partial sums.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_move_f32(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWMASK, 0xf, true));
}
__device__ __forceinline__ float wave_sum_f32(float v) {
  v += dpp_move_f32<0x111, 0xf>(v);  // row_shr:1
  v += dpp_move_f32<0x112, 0xf>(v);  // row_shr:2
  v += dpp_move_f32<0x114, 0xf>(v);  // row_shr:4
  v += dpp_move_f32<0x118, 0xf>(v);  // row_shr:8
  // The broadcasts as one DPP add each, written out: the rows outside the row mask keep
  // their value (the compiler forms these as a zero move, a DPP move and an add).  The
  // s_nop gives the DPP read of v the two wait states it needs after v's last write.
  asm volatile(
      "s_nop 1\n\t"
      "v_add_f32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_add_f32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf"
      : "+v"(v));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// a wave-uniform double moved into scalar registers (frees vector registers)
__device__ __forceinline__ double uniform_f64(double v) {
  return __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(v)),
                          __builtin_amdgcn_readfirstlane(__double2loint(v)));
}

__device__ __forceinline__ float uniform_f32(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// The LDS a brick task works in: the brick in the skewed layout plus one zero slot that
// rejected samples read (branch-free accumulation), per wave the packed first/last column
// of every row of its current section, and the workgroup's next direction.
struct radon_lds {
  float sb[FB * FR_PX + 1];
  int rowlo[FR_WAVES][136];
  int next_dir;
};

// Transpose mode: the half brick's 64-bit accumulators, cell (lx, ly, lz) at
// lx * FT_PX + ly * FT_PY + lz (odd pitches spread a plane patch over the banks).
#define FT_Z 16
#define FT_PY 17
#define FT_PX (17 * 32 + 1)
struct radon_T_lds {
  unsigned long long acc[FB * FT_PX];
  int rowlo[FR_WAVES][136];
  int next_dir;
};

// One task: brick `brick` against directions [grp*dgroup, (grp+1)*dgroup).
// Forward (T = false, L a radon_lds): the brick of a.bricks is staged and every section's
// sum is added to its plane's fixed-point accumulator.  Transpose (T = true, L a radon_T_lds): the
// half brick zhalf (cells with z in [16*zhalf, 16*zhalf + 16) of the brick) starts at zero
// and every sample of a section whose cell lies in it adds the section's quantized value
// rint(a.g[plane] * a.gscale) to that cell, so that after the task the accumulators hold
// sum_planes q_p * (number of the plane's samples in the cell); the caller stores them.
// Called by every thread of a FR_WAVES*64-thread workgroup; returns with the workgroup's
// updates issued (no barrier at the end: a caller that reads or reuses L waits first).
// The variant flags are wave-uniform: BND the brick touches a cube face, BF16 the bricks
// are stored as BF16, POW2 the edge n is a power of two.  The per-job launches pass them as
// compile-time constants (one specialised kernel per variant); the persistent batch kernel
// passes them per task at run time, which costs one scalar branch at the stage and in the
// rare exact branch, so that one code path (one register allocation) serves every task.
template <bool T = false, int BR = 8, class LDS = radon_lds>
__device__ __forceinline__ void radon_brick_task(const wrt_radon_fast_args& a, int brick, int grp,
                                                 double h, double dc, double inv_dc, LDS& L,
                                                 bool BND, bool BF16, bool POW2, int zhalf = 0) {
  constexpr int ZN = T ? FT_Z : FB;  // the box's extent in z
  static_assert(BR == 4 || BR == 8 || BR == 16, "bands of 4, 8 or 16 rows");
  static_assert(!T || BR == 8, "transpose mode walks 8-row bands");
  constexpr int BC = 64 / BR;                          // columns per step
  constexpr int LBC = BC == 16 ? 4 : (BC == 8 ? 3 : 2);  // log2(BC)
  constexpr int NBH = 64 / BR;                         // bands per 64-row half
  int (*const rowlo)[136] = L.rowlo;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // The job's tables, which no kernel writes while it runs, in the constant address space:
  // their wave-uniform reads become scalar loads through the scalar cache (vector loads of a
  // uniform address waited on at every section otherwise; the batch kernel reads the argument
  // block from memory, where the compiler does not know the address space at all).
  const WRT_CONST double* const g_dirs = (const WRT_CONST double*)a.dirs;
  const WRT_CONST double* const g_shift = (const WRT_CONST double*)a.shift;
  const WRT_CONST int* const g_sidx = (const WRT_CONST int*)a.sidx;
  const int nb = a.nb, n = a.n;
  const int bxi = brick / (nb * nb), byi = (brick / nb) % nb, bzi = brick % nb;
  const int b0x = bxi * FB, b0y = byi * FB, b0z = bzi * FB + (T ? FT_Z * zhalf : 0);

  // stage the brick into the skewed LDS layout (transpose: clear the accumulators)
  if constexpr (T) {
    for (int i = threadIdx.x; i < FB * FT_PX; i += blockDim.x) L.acc[i] = 0ull;
    if (threadIdx.x == 0) L.next_dir = grp * a.dgroup;
  } else {
    // 32-bit words: one FP32 voxel, or two BF16 voxels (z even and odd of one row; the
    // low half is the even one)
    float* const sb = L.sb;
    const int nw = BF16 ? FB * FB * FB / 2 : FB * FB * FB;
    const WRT_GLOBAL unsigned* const src = (const WRT_GLOBAL unsigned*)a.bricks + (size_t)brick * nw;
    for (int i = threadIdx.x; i < nw; i += blockDim.x) {
      const unsigned w = src[i];
      const int v = BF16 ? 2 * i : i;  // the (first) voxel's index in the brick
      float* const p = sb + 1 + (v >> 10) * FR_PX + ((v >> 5) & 31) * FR_PY + (v & 31);
      if (BF16) {
        p[0] = __uint_as_float(w << 16);
        p[1] = __uint_as_float(w & 0xffff0000u);
      } else {
        p[0] = __uint_as_float(w);
      }
    }
    if (threadIdx.x == 0) {
      sb[0] = 0.f;
      L.next_dir = grp * a.dgroup;
    }
  }
  __syncthreads();

  const double delta = (1.0 / (n));
  const double lo[3] = {-1.0 + b0x * dc, -1.0 + b0y * dc, -1.0 + b0z * dc};
  const double hi[3] = {-1.0 + (b0x + FB) * dc, -1.0 + (b0y + FB) * dc, -1.0 + (b0z + ZN) * dc};
  const double fb0x = (double)b0x, fb0y = (double)b0y, fb0z = (double)b0z;
  const int L0 = -2 * n, L1 = 2 * n - 1;  // lattice index range of i1 and i2 (bounds of R0..R1, C0..C1)
  // Faces of the cube the brick touches.  Only along those axes can a sample whose cell is
  // in the brick fail the reference's cube test |x| <= 1.  The fixed-point decisions below
  // floor the cell coordinate, so a sample with x < -1 gets cell -1 or less, outside every
  // brick, and one with x > 1 gets cell n or more, past the brick or in its zero padding
  // (cells past n hold 0 in the forward and are never read back in the transpose): both add
  // nothing, as the reference's cube test gives.  Only the exact binary64 decision near a
  // cell boundary truncates as zero_interp does, where -1 < q < 0 (x just below -1) would
  // give cell 0, so on face bricks it applies the cube test itself; there x == 1 exactly
  // gives index n, which the reference maps to n - 1.
  const bool fhi[3] = {b0x + FB >= n, b0y + FB >= n, b0z + ZN >= n};
  const int cxh = fhi[0] ? n - 1 : INT_MAX, cyh = fhi[1] ? n - 1 : INT_MAX,
            czh = fhi[2] ? n - 1 : INT_MAX;
  const int r_lane = lane >> LBC, k_lane = lane & (BC - 1);
  // the brick's farthest corner lies within the support sphere
  bool inside;
  {
    double f2c = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) f2c += fmax(lo[c] * lo[c], hi[c] * hi[c]);
    inside = f2c * (1.0 + 1e-12) < a.rho2;
  }
  // as a scalar, so that branches on it stay scalar compares
  const int inside_s = __builtin_amdgcn_readfirstlane((int)inside);
  // The rows' candidate column intervals are computed in FP32 against a slab widened by
  // RI_TOL in coordinates and a support sphere whose rho^2 (<= 3) is widened by
  // 1e-4 (1 + rho^2).  Errors are measured in coordinates: a column error times |b_c| for a
  // slab bound, times |b| for the sphere's, both at most 1/n.  A slab face F of axis c bounds
  // the columns of row r at (F - (s n_c + a_c r)) / b_c, evaluated as kF - s kn + beta r with
  // the per-direction products kF = F / b_c, kn = n_c / b_c, beta = -a_c / b_c rounded to FP32
  // and two FMAs: each rounding is at most 2^-24 times one of |F|, |s n_c|, |a_c r|,
  // |F - s n_c| or the bound's own |t b_c| (|t| <= 4n + 2 within the clamps below), all at
  // most 6, so together under 2e-6 < RI_TOL whatever b_c.  The sphere |p|^2 <= rho^2 meets
  // row r (p = s n + a r + b t, with n, a, b orthogonal) in the columns |t - t_c| <= sqrt(u)
  // / |b|, u = rho^2 - s^2 - |a|^2 r^2, t_c = -(s n.b + r a.b) / |b|^2 (zero up to the
  // rounding of the basis).  Where u >= 0, FP32 moves u by under 2e-6, much less than the
  // widening, which therefore lengthens the half-width by at least (1e-4 - 2e-6) /
  // (2 sqrt(3.001)) > 2.5e-5, beyond every rounding of t_c, of the half-width and of the ends
  // (each under 2e-6).  The intervals therefore hold every column whose sample can fall in
  // the brick and the sphere; which samples count is decided per sample, exactly.
  constexpr float RI_TOL = 4e-6f;
  float lof[3], hcf[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    lof[c] = uniform_f32((float)lo[c] - RI_TOL);
    hcf[c] = uniform_f32((float)fmin(hi[c], 1.0) + RI_TOL);
  }
  const float rho2wf = uniform_f32((float)(a.rho2 + 1e-4 * (1.0 + a.rho2)));
  const double rho_w = sqrt(a.rho2) * (1.0 + 1e-12);

  const int d_end = min((grp + 1) * a.dgroup, a.ndir);
  // waves take the group's directions one at a time from a shared counter, so that they
  // finish together however unequal the directions' sections are
  for (;;) {
    int d = 0;
    if (lane == 0) d = atomicAdd(&L.next_dir, 1);
    d = __builtin_amdgcn_readfirstlane(d);
    if (d >= d_end) break;
    const WRT_CONST double* D = g_dirs + 9 * (size_t)d;
    const double nrm[3] = {uniform_f64(D[0]), uniform_f64(D[1]), uniform_f64(D[2])};
    const double av[3] = {uniform_f64(D[3]), uniform_f64(D[4]), uniform_f64(D[5])};
    const double bv[3] = {uniform_f64(D[6]), uniform_f64(D[7]), uniform_f64(D[8])};
    // projections of the brick box on a, b and the normal (box corners extremise them)
    double pa0 = 0, pa1 = 0, pb0 = 0, pb1 = 0, pn0 = 0, pn1 = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      pa0 += fmin(lo[c] * av[c], hi[c] * av[c]);
      pa1 += fmax(lo[c] * av[c], hi[c] * av[c]);
      pb0 += fmin(lo[c] * bv[c], hi[c] * bv[c]);
      pb1 += fmax(lo[c] * bv[c], hi[c] * bv[c]);
      pn0 += fmin(lo[c] * nrm[c], hi[c] * nrm[c]);
      pn1 += fmax(lo[c] * nrm[c], hi[c] * nrm[c]);
    }
    const double aa = av[0] * av[0] + av[1] * av[1] + av[2] * av[2];
    const double bb = bv[0] * bv[0] + bv[1] * bv[1] + bv[2] * bv[2];
    // The direction's row window R0..R1 and column window C0..C1 over the brick box.  The
    // box spans (pa1 - pa0) / |a|^2 = 64 sum_c |a_c| / |a| <= 64 sqrt(3) < 111 rows (a box
    // edge of 32 cells is 64 lattice steps of |a| = half a cell; 16 cells in z in transpose
    // mode), so with the margins R1 - R0 < 118 and the rows fit the two 64-row passes.
    const int R0 = max(L0, (int)floor(pa0 / aa) - 2), R1 = min(L1, (int)ceil(pa1 / aa) + 2);
    const int C0 = max(L0, (int)floor(pb0 / bb) - 2), C1 = min(L1, (int)ceil(pb1 / bb) + 2);
    if (R0 > R1 || C0 > C1) continue;
    const int nrow = R1 - R0 + 1;
    // shifts with s in [pn0 - eps, pn1 + eps] (the ascending copy of the grid, bisected).  A
    // plane that misses the closed box has every lattice point outside it, and a point's
    // binary64 rounding (below 1e-15 here) is all that could bring its cell into the brick.
    // The range is also clipped to |s| <= rho (slightly widened): a plane farther from the
    // origin than the support radius meets only zero cells.
    const double eps = 1e-9;
    const double slo = fmax(pn0 - eps, -rho_w), shi = fmin(pn1 + eps, rho_w);
    int j0, j1;
    {
      int l = 0, r = a.nshift;
      while (l < r) { const int m = (l + r) >> 1; if (g_shift[m] < slo) l = m + 1; else r = m; }
      j0 = l;
      l = 0; r = a.nshift;
      while (l < r) { const int m = (l + r) >> 1; if (g_shift[m] <= shi) l = m + 1; else r = m; }
      j1 = l - 1;
    }
    if (j0 > j1) continue;
    // FP32 constants of the rows' column bounds (see above): for an axis with b_c != 0 the
    // lower and upper slab faces give kL, kU; an axis with b_c == 0 (to 1e-9 of a cell over
    // the lattice) bounds the row instead, tested per row, and its column bounds are the
    // neutral -3e38 and 3e38 (kn = beta = 0), so that every row takes the max and min of
    // three bounds with no per-axis selection.  kL and kU stay in vector registers: with kn
    // and the row in scalar ones, each bound is one FMA within the one-scalar-operand limit.
    float kL[3], kU[3], kn[3], beta[3], avf[3], nf[3];
    bool zb[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double ibc = (fabs(bv[c]) * (4.0 * n) < 1e-9 * dc) ? 0.0 : 1.0 / bv[c];
      const float ib = (float)ibc;
      zb[c] = ibc == 0.0;
      const float Lf = bv[c] > 0.0 ? lof[c] : hcf[c], Uf = bv[c] > 0.0 ? hcf[c] : lof[c];
      kL[c] = zb[c] ? -3.0e38f : Lf * ib;
      kU[c] = zb[c] ? 3.0e38f : Uf * ib;
      kn[c] = uniform_f32((float)nrm[c] * ib);
      beta[c] = uniform_f32(-(float)av[c] * ib);
      avf[c] = uniform_f32((float)av[c]);
      nf[c] = uniform_f32((float)nrm[c]);
    }
    const int anyzb_s = __builtin_amdgcn_readfirstlane((int)(zb[0] || zb[1] || zb[2]));
    const float aaf = uniform_f32((float)aa), ibsf = uniform_f32((float)(1.0 / sqrt(bb)));
    // exact: |C| <= 2n + 1
    const float C0f = (float)C0, C1f = (float)C1, C0m1 = (float)(C0 - 1), C1p1 = (float)(C1 + 1);
    const float kab = uniform_f32((float)(-(av[0] * bv[0] + av[1] * bv[1] + av[2] * bv[2]) / bb));
    const float knb = uniform_f32((float)(-(nrm[0] * bv[0] + nrm[1] * bv[1] + nrm[2] * bv[2]) / bb));
    // Fixed-point steps of the cell coordinate q_c = (x_c + 1) h (h = n/2, 32 fraction bits)
    // per row (a_c h) and per column (b_c h); |a_c h|, |b_c h| <= 1/2, clamped into int32
    // (error at most 2^-32 each).  Eight columns for the column walk.
    int Afx[3], Dfx[3], nDfx[3];
    long long DS[3];  // one step of BC columns
#pragma unroll
    for (int c = 0; c < 3; ++c) {  // wave-uniform: held in scalar registers
      Afx[c] = __builtin_amdgcn_readfirstlane(
          (int)fmin(fmax(rint(av[c] * h * 4294967296.0), -2147483648.0), 2147483647.0));
      Dfx[c] = __builtin_amdgcn_readfirstlane(
          (int)fmin(fmax(rint(bv[c] * h * 4294967296.0), -2147483647.0), 2147483647.0));
      nDfx[c] = -Dfx[c];
      DS[c] = (long long)BC * Dfx[c];
    }
    // The fixed-point base of axis c at (row R0, column C0) for each section is one FMA onto
    // 1.5 * 2^52: s * (n_c h 2^32) + (K_c 2^32 + 1.5 * 2^52 + 1024), K_c = ((a_c R0 + b_c C0)
    // + 1) h - b0_c.  The base q lies within 200 cells of the brick (|q 2^32| < 2^40), so the
    // FMA rounds to an integer-valued double in [2^52, 2^53) whose bit pattern less the
    // constant's is the 64-bit fixed-point value with its 1024-unit bias (below); with the
    // constant's own rounding it errs by at most 2^-32 cells.
    const double fb0[3] = {fb0x, fb0y, fb0z};
    double nh32[3], K32[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      nh32[c] = nrm[c] * h * 4294967296.0;  // exact scalings by powers of two
      K32[c] = ((((av[c] * (double)R0 + bv[c] * (double)C0) + 1.0) * h - fb0[c]) * 4294967296.0 +
                6755399441055744.0) + 1024.0;
    }
    for (int j = j0; j <= j1; ++j) {
      const int jo = g_sidx[j];  // the plane's row in the caller's shift order
      unsigned long long gq = 0ull;  // transpose: the section's quantized value
      if constexpr (T) {
        const float gv = ((const WRT_CONST float*)a.g)[(size_t)jo * a.ndir + d];
        if (gv == 0.f) continue;
        gq = (unsigned long long)(long long)rint((double)gv * a.gscale);
        if (gq == 0ull) continue;
      }
      const double s = g_shift[j];
      const float sf = (float)s;
      float aL[3], aU[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        aL[c] = fmaf(-sf, kn[c], kL[c]);
        aU[c] = fmaf(-sf, kn[c], kU[c]);
      }
      const float rc2f = fmaf(-sf, sf, rho2wf), tcs = sf * knb;
      // the flags re-read as scalars in the loop, so that their tests stay scalar compares
      // (hoisted out of the loop, the compiler keeps them as lane masks rebuilt per section)
      int ins = inside_s, azb = anyzb_s;
      asm volatile("" : "+s"(ins), "+s"(azb));
      // Candidate column interval of every row R0 + t of the window (t = lane, lane + 64):
      // first and last column into the wave's row table, the BC-column step count reduced to
      // a per-band maximum (bands of the rows BR k .. BR k + BR - 1 of the window), and the
      // rows that meet the brick and the sphere as a ballot
      int bsteps[2];
      unsigned long long live[2];
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        if (64 * half >= nrow) {  // uniform: the window has at most 64 rows
          bsteps[half] = -1;
          live[half] = 0ull;
          continue;
        }
        const int t = lane + 64 * half;
        // branch-free except for the rare cases: every lane forms its row's bounds.  A row
        // past the window has an empty interval or columns whose samples all lie outside
        // the brick (the window holds the box's rows), so it needs no mask: a walked sample
        // is tested exactly whatever the interval.
        const float rf = (float)(R0 + t);  // exact: |row| <= 2n + 130
        float tlo = fmaxf(fmaxf(fmaf(beta[0], rf, aL[0]), fmaf(beta[1], rf, aL[1])),
                          fmaf(beta[2], rf, aL[2]));
        float thi = fminf(fminf(fmaf(beta[0], rf, aU[0]), fmaf(beta[1], rf, aU[1])),
                          fmaf(beta[2], rf, aU[2]));
        if (azb) {  // rare: a basis component vanishes (structured directions)
          bool dead = false;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const float w = fmaf(avf[c], rf, sf * nf[c]);
            if (zb[c] && (w < lof[c] || w > hcf[c])) dead = true;
          }
          if (dead) thi = -3.0e38f;
        }
        // Nothing to clip when the whole brick lies inside the sphere.  A row that misses the
        // sphere (u < 0) gets the half-width 0: at most the one column at t_c, harmlessly.
        if (!ins) {
          const float u = fmaf(-aaf * rf, rf, rc2f);
          // the hardware square root (1 ulp): far inside the slack
          const float hw = __builtin_amdgcn_sqrtf(fmaxf(u, 0.0f)) * ibsf;
          const float tc = fmaf(rf, kab, tcs);
          // max and min as med3 against +-3e38 (fmaxf would first quiet its operands)
          tlo = __builtin_amdgcn_fmed3f(tlo, tc - hw, 3.0e38f);
          thi = __builtin_amdgcn_fmed3f(thi, tc + hw, -3.0e38f);
        }
        // First and last column: rl = ceil of the lower bound clamped into [C0, C1 + 1], as
        // its negation -rl = floor(-x) (one conversion); rh = floor of the upper bound clamped
        // into [C0 - 1, C1].  An empty row ends with rh < rl.
        const int nrl = cvt_floor_neg(__builtin_amdgcn_fmed3f(tlo, C0f, C1p1));
        const int rh = cvt_floor(__builtin_amdgcn_fmed3f(thi, C0m1, C1f));
        // rh - rl, and the row's steps of BC columns less one, (rh - rl) / BC rounded down:
        // negative for an empty row (the bands add the one back)
        const int span = rh + nrl;
        const int nst = span >> LBC;
        // -rl and rh, 16 bits each (|i2| <= 2n + 1 < 2^15)
        rowlo[wv][t] = (int)__builtin_amdgcn_perm((unsigned)rh, (unsigned)nrl, 0x05040100u);
        // max over the BR consecutive rows of a band (lanes BR b .. BR b + BR - 1)
        bsteps[half] = group_max<BR>(nst);
        live[half] = __ballot(span >= 0);
      }
      if ((live[0] | live[1]) == 0ull) continue;  // no row meets the brick and the sphere
      // the bands that hold the first and the last live row
      const int t_first = live[0] ? __builtin_ctzll(live[0]) : 64 + __builtin_ctzll(live[1]);
      const int t_last = live[1] ? 127 - __builtin_clzll(live[1]) : 63 - __builtin_clzll(live[0]);
      const int k_first = t_first / BR, k_last = t_last / BR;
      // both halves' band maxima in one register: band b of half 0 at lane BR b, of half 1
      // at lane BR b + 1 (a band's BR lanes hold the same maximum)
      const int bsteps01 = (lane & (BR - 1)) == 1 ? bsteps[1] : bsteps[0];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // Cell decisions in fixed point.  Q_c holds (x_c + 1) h - b0_c with 32 fraction bits,
      // where x_c is the sample's coordinate: Q = Qb + t A + (col - C0) D, Qb the value at
      // (row R0, column C0) converted from binary64.  Against the reference's quotient q
      // (x rounded as radon_grid/src/integral.c:47-51 rounds it, then (x + 1) / (2/n)) the
      // error is at most (1 + t + col - C0) 2^-32 (t < 118 rows; col - C0 < 160: a brick's
      // section spans at most 64 sqrt(3) + 6 columns, plus a step's overshoot) plus the
      // roundings of q (below 1e-13) and of Qb (2^-32): under 2^-23.  Q carries a bias of
      // 2^-22 (1024 units), so wherever its fraction is at least 2^-21 the reference's
      // quotient lies inside (k + 2^-23, k + 1 - 2^-23) for the integer part k of Q, which
      // is then the reference's truncated index (less b0); the few samples nearer a cell
      // boundary take the exact binary64 decision instead.
      long long Qb[3];
#pragma unroll
      for (int c = 0; c < 3; ++c)
        Qb[c] = __double_as_longlong(fma(s, nh32[c], K32[c])) - 0x4338000000000000ll;
      // One sample: Q the lane's fixed-point coordinates, R0 + t its row and cb - nrl + step*m
      // its column (for the exact decision); ok and off give its cell (below).  The forward
      // walk adds the cell's value to its lane sum; the transpose adds the section's quantized
      // value to the cell.
      // The cell of brick offsets (lx, ly, lz): all three in [0, 32) (z: [0, 16) in transpose
      // mode, tested as 2*lz < 32; a negative offset sets high bits of the OR).  ok and the
      // cell's byte offset in the LDS brick (forward: slot 0 holds the zero that rejected
      // samples read, the brick starts at slot 1, and a rejected sample gets offset 0).
      auto cell = [&](int lx, int ly, int lz, bool& ok, unsigned& off) {
        ok = ((unsigned)lx | (unsigned)ly | ((unsigned)lz << (T ? 1 : 0))) < (unsigned)FB;
        if constexpr (T) {
          off = mad_u24((unsigned)lx, 8u * FT_PX, mad_u24((unsigned)ly, 8u * FT_PY, (unsigned)lz * 8u));
        } else {
          off = mad_u24((unsigned)lx, 4u * FR_PX, mad_u24((unsigned)ly, 4u * FR_PY, (unsigned)lz * 4u + 4u));
          off = ok ? off : 0u;
        }
      };
      auto locate = [&](const long long (&Q)[3], int t, int cb, int nrl, int step, int m,
                        bool& ok, unsigned& off) {
        const unsigned u = min(min((unsigned)Q[0], (unsigned)Q[1]), (unsigned)Q[2]);
        // the fixed point's cell (away from a boundary the fixed point floors where the
        // reference truncates: they differ only for -1 < q < 0, x < -1, where the floor's
        // cell -1 lies outside every brick and the reference's cube test gives 0) ...
        cell((int)(Q[0] >> 32), (int)(Q[1] >> 32), (int)(Q[2] >> 32), ok, off);
        const bool near = u < 2048u;
        // a branch of the whole wave on the ballot (a scalar branch on VCC, no exec mask to
        // save and restore in every step); each lane takes the exact result only if it is near
        if (__builtin_expect(__builtin_amdgcn_ballot_w64(near) != 0ull, 0)) {
          // ... replaced near a cell boundary (fraction of Q below 2^-21) by the reference's
          // rounding, in binary64 (the per-column products b*i2 rounded as the reference
          // rounds delta*v2*i2); the step index through readfirstlane keeps the column out
          // of the common path.  Row and column are formed here from t and nrl passed through
          // an empty asm, so that the compiler does not hoist them, or the row's binary64
          // terms below, out of the sample loop into every band's setup.
          int tv = t, nv = nrl;
          asm volatile("" : "+v"(tv), "+v"(nv));
          const double f2 = (double)(cb - nv + step * __builtin_amdgcn_readfirstlane(m));
          const double f1 = (double)(R0 + tv);
          // the shift, opaque here too, so that its products stay in this branch rather than
          // in every section's setup
          double sv = s;
          asm volatile("" : "+s"(sv));
          const double Rx = sv * nrm[0] + av[0] * f1, Ry = sv * nrm[1] + av[1] * f1,
                       Rz = sv * nrm[2] + av[2] * f1;
          const double x = Rx + bv[0] * f2, y = Ry + bv[1] * f2, z = Rz + bv[2] * f2;
          double qx, qy, qz;
          if (POW2) {
            // RN(x + 1) * h == RN(x*h + h) for h = n/2 a power of two: scaling by 2^k
            // commutes with rounding, so one fma gives the reference's quotient exactly
            qx = fma(x, h, h); qy = fma(y, h, h); qz = fma(z, h, h);
          } else {
            qx = div_rn(x + 1.0, dc, inv_dc);
            qy = div_rn(y + 1.0, dc, inv_dc);
            qz = div_rn(z + 1.0, dc, inv_dc);
          }
          // truncation as in zero_interp; for an interior brick a negative quotient (x < -1)
          // truncates to 0 <= b0 - 32 and a quotient above n lands past the brick, so the
          // brick test also is the cube test.  Face bricks apply the cube test itself, and
          // x == 1 (index n) maps to n - 1.
          int ix = (int)qx, iy = (int)qy, iz = (int)qz;
          if (BND) {
            ix = min(ix, cxh);
            iy = min(iy, cyh);
            iz = min(iz, czh);
          }
          bool okx;
          unsigned offx;
          cell(ix - b0x, iy - b0y, iz - b0z, okx, offx);
          if (BND && (fabs(x) > 1.0 || fabs(y) > 1.0 || fabs(z) > 1.0)) {
            okx = false;
            if constexpr (!T) offx = 0u;
          }
          ok = near ? okx : ok;
          off = near ? offx : off;
        }
      };
      auto sample = [&](const long long (&Q)[3], int t, int cb, int nrl, int step, int m) {
        bool ok;
        unsigned off;
        locate(Q, t, cb, nrl, step, m, ok, off);
        if constexpr (T) {
          if (ok)
            __hip_atomic_fetch_add(
                (unsigned long long*)__builtin_assume_aligned((char*)L.acc + off, 8), gq,
                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      };
      // Rows past R1 and columns past the lattice end need no test: a plane's cube section
      // lies within sqrt(3) < 2 of the plane centre, inside the [-2n, 2n) lattice, so only
      // the brick test can reject a sample.
      float bs = 0.f;  // forward: the lane's FP32 sum over the section's samples
      // Forward: the lane's fixed-point coordinates at row R0 and column k_lane (Qb + (k_lane
      // - C0) D); a band adds t A and rl D (as (-rl)(-D), the row table's form).
      long long Tb[3];
      if constexpr (!T) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          Tb[c] = Qb[c] + (long long)(k_lane - C0) * Dfx[c];
          // opaque to the compiler, which would otherwise merge the two column terms into
          // one 64-bit product per band
          asm volatile("" : "+v"(Tb[c]));
        }
      }
      // Bands of BR rows: lane (r, k) takes row BR k' + r of band k' and the columns rl + k +
      // BC m (rl the row's first column).  Transpose mode (BR = 8) takes the columns rl + dc +
      // 2k + 16m in two phases dc = 0, 1:
      // the lanes of one instruction then lie 2 columns apart and fewer of them add to one
      // cell at once (the LDS integer atomic serialises those, which costs time, not
      // results).  Measured at 256^3: 2.66 s against 2.92 s for the forward's columns and
      // 2.90 s for bands of 16 rows taken 2 apart (their longer rows' steps waste more).
      int rlo_next = rowlo[wv][BR * k_first + r_lane];
      for (int k = k_first; k <= k_last; ++k) {
        // this lane's row within the window, passed through an empty asm so that its
        // products below stay signed 32 x 32 -> 64 multiply-adds (one instruction each)
        // instead of the two unsigned ones the compiler picks for a known non-negative factor
        int t = BR * k + r_lane;
        asm volatile("" : "+v"(t));
        const int nrl = (rlo_next << 16) >> 16;  // minus the band row's first column
        if (k < k_last) rlo_next = rowlo[wv][t + BR];
        const int steps = 1 + __builtin_amdgcn_readlane(bsteps01, BR * (k % NBH) + k / NBH);
        if (steps <= 0) continue;
        if constexpr (T) {
          const int steps2 = (steps + 1) >> 1;  // 16-column steps cover the 8-column ones
#pragma unroll
          for (int dcol = 0; dcol < 2; ++dcol) {
            const int cb = dcol + 2 * k_lane, c0 = cb - nrl;  // c0: first column i2 of the lane
            long long Q[3];
#pragma unroll
            for (int c = 0; c < 3; ++c)
              Q[c] = Qb[c] + (long long)t * Afx[c] + (long long)(c0 - C0) * Dfx[c];
            for (int m = 0; m < steps2; ++m) {
              sample(Q, t, cb, nrl, 16, m);
#pragma unroll
              for (int c = 0; c < 3; ++c) Q[c] += 2 * DS[c];
            }
          }
        } else {
          long long Q[3];
#pragma unroll
          for (int c = 0; c < 3; ++c)
            Q[c] = Tb[c] + (long long)t * Afx[c] + (long long)nrl * nDfx[c];
          // Software-pipelined: step m's LDS read is issued before step m + 1's cell decision
          // and added after it, so that the read's latency overlaps that work.  A rejected
          // sample reads the zero slot (off = 0).
          bool ok;
          unsigned off;
          locate(Q, t, k_lane, nrl, BC, 0, ok, off);
          // opaque, so that the loop steps Q instead of forming step 1 from the band's terms
#pragma unroll
          for (int c = 0; c < 3; ++c) asm volatile("" : "+v"(Q[c]));
          for (int m = 1; m < steps; ++m) {
            const float v = *(const float*)((const char*)L.sb + off);
#pragma unroll
            for (int c = 0; c < 3; ++c) Q[c] += DS[c];
            locate(Q, t, k_lane, nrl, BC, m, ok, off);
            bs += v;
          }
          bs += *(const float*)((const char*)L.sb + off);
        }
      }
      if constexpr (!T) {
        // the section's lane sums reduced in FP32 (one DPP add per step); its total, in
        // units of qstep, rounded to an integer by one FMA onto 1.5 * 2^52 (|tot * qmul| <
        // 2^50, so the result is an integer-valued double whose bits less the constant's are
        // that integer), added to the plane's 64-bit sum: integer adds are exact and
        // associative, so the plane's sum does not depend on the order of the sections
        const float tot = wave_sum_f32(bs);
        if (lane == 0 && tot != 0.0f) {
          const double y = fma((double)tot, a.qmul, 6755399441055744.0);
          const long long q = __double_as_longlong(y) - 0x4338000000000000ll;
          __hip_atomic_fetch_add((WRT_GLOBAL unsigned long long*)a.acc + (size_t)jo * a.ndir + d,
                                 (unsigned long long)q, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
  }
}

#endif
