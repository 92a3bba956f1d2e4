// k_fast.hip -- fast form of the plane integrals (radon_grid, FP32 or BF16 values).
//
// Work unit: one workgroup per (non-empty 32^3 brick, group of directions).  The brick is
// staged once into LDS as FP32; every plane (shift, direction) that cuts the brick is a
// "section".  A wavefront walks a section in 8-row bands: lane (r, k) = (lane>>3, lane&7)
// handles lattice row i1 = band + r and columns i2 = c0 + k, c0 stepping by 8 across the
// band's conservative column interval.  Per sample the cell index is decided exactly as
// radon_grid/src/integral.c:47-51 + interp.c:70-72 decide it in binary64
//   x = ((s*n) + (a*i1)) + (b*i2),  ix = (int)((x + 1) / (2/n)),  index n -> n-1,
// (the per-column products b*i2 are formed in registers, the division is an exact
// power-of-two scaling or a corrected reciprocal), so only the value rounding (FP32/BF16)
// and the summation order differ from the plain form.  Samples whose cell lies in the
// brick add its FP32 value to a per-band partial, folded into a Kahan-compensated FP32
// lane sum; a section's lane sums are reduced over the wavefront in FP64 and added to the
// plane's FP64 accumulator (one atomic per section and brick).
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

// a*b + c on the 24-bit multiplier (full rate); a, b < 2^24.  Written out because the
// compiler turns small known-range products into the quarter-rate 32-bit multiply.
__device__ __forceinline__ unsigned mad_u24(unsigned a, unsigned b, unsigned c) {
  unsigned r;
  asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// Wavefront sum of a double with DPP lane moves (no LDS round trips): Hillis-Steele scan
// within 16-lane rows (row_shr 1, 2, 4, 8, zero fill), then row broadcasts 15 and 31; the
// total ends in lane 63 and is returned to every lane.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xf, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_move<0x111, 0xf>(v);  // row_shr:1
  v += dpp_move<0x112, 0xf>(v);  // row_shr:2
  v += dpp_move<0x114, 0xf>(v);  // row_shr:4
  v += dpp_move<0x118, 0xf>(v);  // row_shr:8
  v += dpp_move<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_move<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ float bf16_to_f32(unsigned short v) {
  return __uint_as_float(((unsigned)v) << 16);
}

template <bool BND, bool BF16, bool POW2>
__global__ __launch_bounds__(FR_WAVES * 64) void k_radon_fast(wrt_radon_fast_args a, double h,
                                                              double dc, double inv_dc) {
  // the brick, plus one zero slot that rejected samples read (branch-free accumulation)
  __shared__ float sb[FB * FR_PX + 1];
  // per wave: first column of every row of the current section
  __shared__ int rowlo[FR_WAVES][136];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int brick = a.tasks[2 * blockIdx.x], grp = a.tasks[2 * blockIdx.x + 1];
  const int nb = a.nb, n = a.n;
  const int bxi = brick / (nb * nb), byi = (brick / nb) % nb, bzi = brick % nb;
  const int b0x = bxi * FB, b0y = byi * FB, b0z = bzi * FB;

  // stage the brick into the skewed LDS layout
  {
    const size_t base = (size_t)brick * (FB * FB * FB);
    for (int i = threadIdx.x; i < FB * FB * FB; i += blockDim.x) {
      const int lx = i >> 10, ly = (i >> 5) & 31, lz = i & 31;
      float v;
      if (BF16) v = bf16_to_f32(((const unsigned short*)a.bricks)[base + i]);
      else v = ((const float*)a.bricks)[base + i];
      sb[lx * FR_PX + ly * FR_PY + lz] = v;
    }
    if (threadIdx.x == 0) sb[FB * FR_PX] = 0.f;
  }
  __syncthreads();
  const unsigned zero_off = 4u * (FB * FR_PX);

  const double delta = (1.0 / (n));
  const double lo[3] = {-1.0 + b0x * dc, -1.0 + b0y * dc, -1.0 + b0z * dc};
  const double hi[3] = {-1.0 + (b0x + FB) * dc, -1.0 + (b0y + FB) * dc, -1.0 + (b0z + FB) * dc};
  const double fb0x = (double)b0x, fb0y = (double)b0y, fb0z = (double)b0z;
  const int L0 = -2 * n, L1 = 2 * n - 1;  // lattice index range of i1 and i2 (bounds of R0..R1, C0..C1)
  // axes along which the brick touches a face of the cube (only those need the cube test;
  // along the others x < -1 truncates outside the brick and x > 1 lands past it)
  const bool fx = b0x == 0 || b0x + FB >= n, fy = b0y == 0 || b0y + FB >= n,
             fz = b0z == 0 || b0z + FB >= n;
  const int r_lane = lane >> 3, k_lane = lane & 7;

  const int d_begin = grp * a.dgroup;
  const int d_end = min(d_begin + a.dgroup, a.ndir);
  for (int d = d_begin + wv; d < d_end; d += FR_WAVES) {
    const double* D = a.dirs + 9 * (size_t)d;
    const double nrm[3] = {D[0], D[1], D[2]};
    const double av[3] = {D[3], D[4], D[5]};
    const double bv[3] = {D[6], D[7], D[8]};
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
    const int R0 = max(L0, (int)floor(pa0 / aa) - 2), R1 = min(L1, (int)ceil(pa1 / aa) + 2);
    const int C0 = max(L0, (int)floor(pb0 / bb) - 2), C1 = min(L1, (int)ceil(pb1 / bb) + 2);
    if (R0 > R1 || C0 > C1) continue;
    // shifts with s in [pn0 - eps, pn1 + eps] (sorted shift grid, binary search)
    const double eps = 4.0 * delta;
    int j0, j1;
    {
      int l = 0, r = a.nshift;
      while (l < r) { const int m = (l + r) >> 1; if (a.shift[m] < pn0 - eps) l = m + 1; else r = m; }
      j0 = l;
      l = 0; r = a.nshift;
      while (l < r) { const int m = (l + r) >> 1; if (a.shift[m] <= pn1 + eps) l = m + 1; else r = m; }
      j1 = l - 1;
    }
    if (j0 > j1) continue;
    const int nrows = R1 - R0 + 1;  // <= 122 by geometry (brick diameter + margins)
    // reciprocals of the column steps (for the conservative row intervals)
    double ib[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) ib[c] = (fabs(bv[c]) * (4.0 * n) < 1e-9 * dc) ? 0.0 : 1.0 / bv[c];
    const int nbands = (nrows + 7) >> 3;
    for (int j = j0; j <= j1; ++j) {
      const double s = a.shift[j];
      const double S[3] = {s * nrm[0], s * nrm[1], s * nrm[2]};
      // conservative column interval of every row (rows R0 + lane and R0 + 64 + lane): first
      // column into the wave's row table, 8-column step count reduced to a per-band maximum
      int bsteps[2];
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        const int t = lane + 64 * half;
        int rl = C0, nst = 0;
        if (t < nrows) {
          const double f1 = (double)(R0 + t);
          double tlo = -1e300, thi = 1e300;
          bool empty = false;
          // the row's points p(t) = w + b t within the support sphere |p|^2 <= rho2 (every
          // sample outside it reads a zero cell, so clipping the row there changes no sum)
          {
            const double wx = S[0] + av[0] * f1, wy = S[1] + av[1] * f1, wz = S[2] + av[2] * f1;
            const double qb = wx * bv[0] + wy * bv[1] + wz * bv[2];
            const double qc = wx * wx + wy * wy + wz * wz - a.rho2;
            const double disc = qb * qb - bb * qc;
            if (disc < 0.0) {
              empty = true;
            } else {
              const double sq = sqrt(disc);
              tlo = (-qb - sq) / bb;
              thi = (-qb + sq) / bb;
            }
          }
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const double w = S[c] + av[c] * f1;
            const double tol = 1e-9 * dc;
            if (ib[c] == 0.0) {
              if (w < lo[c] - tol || w > hi[c] + tol) empty = true;
            } else if (bv[c] > 0) {
              tlo = fmax(tlo, (lo[c] - tol - w) * ib[c]);
              thi = fmin(thi, (hi[c] + tol - w) * ib[c]);
            } else {
              tlo = fmax(tlo, (hi[c] + tol - w) * ib[c]);
              thi = fmin(thi, (lo[c] - tol - w) * ib[c]);
            }
          }
          if (!empty && tlo <= thi + 2.0) {
            // the continuous bounds are accurate to ~1e-12 columns (and the slab tolerance
            // widens them), so the integer columns within 1e-6 of them are the candidates
            tlo = fmax(tlo, (double)C0 - 1.0);
            thi = fmin(thi, (double)C1 + 1.0);
            rl = max(C0, (int)ceil(tlo - 1e-6));
            const int rh = min(C1, (int)floor(thi + 1e-6));
            nst = rh >= rl ? ((rh - rl + 8) >> 3) : 0;
          }
        }
        rowlo[wv][t] = rl;
        // max over the 8 consecutive rows of a band (lanes 8b .. 8b+7)
        nst = max(nst, __shfl_xor(nst, 1, 64));
        nst = max(nst, __shfl_xor(nst, 2, 64));
        nst = max(nst, __shfl_xor(nst, 4, 64));
        bsteps[half] = nst;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      float ks = 0.f, kc = 0.f;  // Kahan-compensated lane sum of the section
      int rlo_next = rowlo[wv][r_lane];
      for (int k = 0; k < nbands; ++k) {
        const int t = 8 * k + r_lane;  // this lane's row within the section
        const int rlo = rlo_next;
        if (k + 1 < nbands) rlo_next = rowlo[wv][t + 8];
        const int steps = (k < 8) ? __builtin_amdgcn_readlane(bsteps[0], 8 * k)
                                  : __builtin_amdgcn_readlane(bsteps[1], 8 * (k - 8));
        if (steps == 0) continue;
        const int i1 = R0 + t;
        // Rows past R1 and columns past the lattice end need no test: a plane's cube section
        // lies within sqrt(3) < 2 of the plane centre, inside the [-2n, 2n) lattice, so only
        // the brick test below can reject a sample.
        const double f1 = (double)i1;
        const double Rx = S[0] + av[0] * f1, Ry = S[1] + av[1] * f1, Rz = S[2] + av[2] * f1;
        float bs = 0.f;
        int c = rlo + k_lane;
        for (int m = 0; m < steps; ++m, c += 8) {
          // per-column products b*i2, rounded as the reference rounds delta*v2*i2
          const double f2 = (double)c;
          const double x = Rx + bv[0] * f2, y = Ry + bv[1] * f2, z = Rz + bv[2] * f2;
          bool ok = true;
          if (BND) {  // cube test on the axes along which the brick touches a face
            if (fx) ok = ok && (fabs(x) <= 1.0);
            if (fy) ok = ok && (fabs(y) <= 1.0);
            if (fz) ok = ok && (fabs(z) <= 1.0);
          }
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
          // brick test below also is the cube test; bricks on the faces test |x| <= 1 above
          int ix = (int)qx, iy = (int)qy, iz = (int)qz;
          if (BND) {
            if (fx && ix == n) ix = n - 1;
            if (fy && iy == n) iy = n - 1;
            if (fz && iz == n) iz = n - 1;
          }
          const int lx = ix - b0x, ly = iy - b0y, lz = iz - b0z;
          // all three offsets in [0, 32): a negative offset sets high bits of the OR
          ok = ok && (((unsigned)lx | (unsigned)ly | (unsigned)lz) < (unsigned)FB);
          const unsigned off = mad_u24((unsigned)lx, 4u * FR_PX,
                                       mad_u24((unsigned)ly, 4u * FR_PY, (unsigned)lz << 2));
          bs += *(const float*)((const char*)sb + (ok ? off : zero_off));
        }
        const float y = bs - kc;
        const float tt = ks + y;
        kc = (tt - ks) - y;
        ks = tt;
      }
      const double tot = wave_sum_dpp((double)ks - (double)kc);
      if (lane == 0 && tot != 0.0) unsafeAtomicAdd(a.acc + (size_t)j * a.ndir + d, tot * delta * delta);
    }
  }
}

template <bool BND, bool BF16>
static int launch_pow2(const wrt_radon_fast_args* a, bool pow2, hipStream_t st) {
  const double n = a->n;
  const double dc = 2.0 / a->n;
  const double inv = 1.0 / dc;
  const double h = n / 2.0;
  if (pow2)
    k_radon_fast<BND, BF16, true><<<a->ntasks, FR_WAVES * 64, 0, st>>>(*a, h, dc, inv);
  else
    k_radon_fast<BND, BF16, false><<<a->ntasks, FR_WAVES * 64, 0, st>>>(*a, h, dc, inv);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

extern "C" int wrtk_radon_fast(const wrt_radon_fast_args* a, int boundary, hipStream_t st) {
  if (a->ntasks <= 0) return WRT_OK;
  const bool pow2 = (a->n & (a->n - 1)) == 0;
  if (boundary) return a->bf16 ? launch_pow2<true, true>(a, pow2, st) : launch_pow2<true, false>(a, pow2, st);
  return a->bf16 ? launch_pow2<false, true>(a, pow2, st) : launch_pow2<false, false>(a, pow2, st);
}

// ------------------------------------------------------------------- brick loader ---
__device__ __forceinline__ unsigned short f32_to_bf16_rne(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__global__ __launch_bounds__(256) void k_brickify(const double* __restrict__ vol, int n, int nb,
                                                  int bf16, void* __restrict__ bricks,
                                                  int* __restrict__ flags,
                                                  unsigned long long* __restrict__ r2max) {
  const size_t total = (size_t)nb * nb * nb * (FB * FB * FB);
  double r2 = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t brick = i / (FB * FB * FB);
    const int li = (int)(i % (FB * FB * FB));
    const int bx = (int)(brick / ((size_t)nb * nb)), by = (int)((brick / nb) % nb), bz = (int)(brick % nb);
    const int x = bx * FB + (li >> 10), y = by * FB + ((li >> 5) & 31), z = bz * FB + (li & 31);
    double v = 0.0;
    if (x < n && y < n && z < n) v = vol[((size_t)x * n + y) * n + z];
    const float f = (float)v;
    if (bf16) ((unsigned short*)bricks)[i] = f32_to_bf16_rne(f);
    else ((float*)bricks)[i] = f;
    if (v != 0.0) {
      flags[brick] = 1;
      const double cx = -1.0 + (2.0 * x + 1.0) / n, cy = -1.0 + (2.0 * y + 1.0) / n,
                   cz = -1.0 + (2.0 * z + 1.0) / n;
      r2 = fmax(r2, cx * cx + cy * cy + cz * cz);
    }
  }
  // non-negative doubles order like their bit patterns
  unsigned long long b = (unsigned long long)__double_as_longlong(r2);
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_down(b, off, 64);
    b = o > b ? o : b;
  }
  if ((threadIdx.x & 63) == 0 && b) atomicMax(r2max, b);
}

extern "C" int wrtk_brickify(const double* vol, int n, int nb, int bf16, void* bricks, int* flags,
                             unsigned long long* r2max, hipStream_t st) {
  const size_t total = (size_t)nb * nb * nb * (FB * FB * FB);
  size_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  k_brickify<<<(unsigned)blocks, 256, 0, st>>>(vol, n, nb, bf16, bricks, flags, r2max);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

__global__ void k_acc_to_f32(const double* __restrict__ acc, float* __restrict__ out, size_t count) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = (float)acc[i];
}

extern "C" int wrtk_acc_to_f32(const double* acc, float* out, size_t count, hipStream_t st) {
  size_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (!blocks) return WRT_OK;
  k_acc_to_f32<<<(unsigned)blocks, 256, 0, st>>>(acc, out, count);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}
