// k_ray_fast.hip -- FP32 form of the layer-by-layer line integrals (ray_grid).
//
// ray_transform (ray_grid/src/integral.c:11-50) sums cube_trilinear_interp
// (ray_grid/src/interp.c:25-60) at the 4n+1 points p_i = (s cos phi, s sin phi, z) +
// (delta dir) i of a ray in the slice z, delta = (2/(n-1))/2.  Every index and fraction
// decision is the reference's, taken in binary64 with its rounding:
//   a = x + 1,  ix = (int)(a / dg),  cx = fmod(a, dg) / dg,  ix == n-1 -> (n-2, cx = 1),
// including the case where a/dg rounds up to an integer k while the exact quotient is below
// it: ix = k and fmod gives nearly dg, so the point takes nearly node k+1 (the reference's
// quirk).  Here the quotient is one corrected-reciprocal division (or an exact power-of-two
// scaling), the exact remainder is a fused multiply-add with the truncated quotient, and the
// z decision is the slice's, made once on the host with the C library's fmod.
//
// Work layout: one workgroup of 512 threads per (slice, 512 rays); a thread walks one ray.
// The points with |x| > 1 or |y| > 1 contribute 0 in the reference; x and y are monotone in
// i, so each ray's range of i with |x|, |y| <= 1 is found exactly at its two ends and only
// that range is walked.  For n <= 181 the workgroup stages the slice's plane interpolated in
// z, P = (1 - cz) V[iz] + cz V[iz+1] (FP32), in LDS; every point is then a bilinear
// interpolation of P from LDS.  Larger volumes read both planes from the FP32 [z][x][y] copy
// in HBM.  Values and weights are FP32 (lerps reordered: z first), the sum along a ray is
// Kahan-compensated in FP32, times delta in FP64 at the end.
#include "radon_brick.hpp"  // div_rn
#include "wrt_dev.h"

// Exact cube test of one axis on a row: x(c) = w + b*c in binary64 (the reference's
// rounding) is monotone in c, so {c : x >= -1} and {c : x <= 1} are a prefix or a suffix of
// the row.  Narrows [rl, rh] to the columns passing the tests that lo / hi ask for.  The
// candidate interval comes from continuous bounds widened by a small slack, so the first
// probe or its neighbour usually settles it; a bisection covers rows almost parallel to the
// face.
__device__ __forceinline__ void cube_ends(double w, double b, bool lo, bool hi, int& rl,
                                          int& rh) {
  if (rl > rh) return;
  for (int side = 0; side < 2; ++side) {
    if (side == 0 ? !lo : !hi) continue;
    // P(c): the test holds at column c; suffix-type when x moves into the cube as c grows
    const bool suffix = side == 0 ? b > 0.0 : b < 0.0;
    auto P = [&](int c) {
      const double x = w + b * (double)c;
      return side == 0 ? x >= -1.0 : x <= 1.0;
    };
    if (b == 0.0) {
      if (!P(rl)) rh = rl - 1;
    } else if (suffix) {
      if (!P(rl)) {
        if (rl < rh && P(rl + 1)) {
          rl = rl + 1;  // the candidate interval's slack added one column
        } else if (!P(rh)) {
          rh = rl - 1;
        } else {
          int l = rl, h = rh;  // P(l) false, P(h) true
          while (h - l > 1) {
            const int m = (l + h) >> 1;
            if (P(m)) h = m; else l = m;
          }
          rl = h;
        }
      }
    } else {
      if (!P(rh)) {
        if (rl < rh && P(rh - 1)) {
          rh = rh - 1;
        } else if (!P(rl)) {
          rh = rl - 1;
        } else {
          int l = rl, h = rh;  // P(l) true, P(h) false
          while (h - l > 1) {
            const int m = (l + h) >> 1;
            if (P(m)) l = m; else h = m;
          }
          rh = l;
        }
      }
    }
    if (rl > rh) return;
  }
}

#define RAY_THREADS 512
#define RAY_LDS_N 181  // largest n whose n x (n | 1) FP32 plane fits the 128 KiB stage

// the LDS planes' row pitch: odd, so that rows start in different banks
__host__ __device__ __forceinline__ int ray_pitch(int n) { return n | 1; }
#define RAY_T_LDS_N 143  // largest n whose n x (n | 1) FP64 plane fits the 160 KiB LDS
#define RAY_T_THREADS 512

// one axis: the reference's index and fraction of coordinate x (|x| <= 1)
template <bool POW2>
__device__ __forceinline__ void ray_axis(double x, int n, double dg, double inv_dg, double h,
                                         int& i, float& f) {
  const double a = x + 1.0;
  // RN(a / dg): h = (n-1)/2 a power of two scales exactly; otherwise the corrected reciprocal
  const double q = POW2 ? a * h : div_rn(a, dg, inv_dg);
  const double qt = trunc(q);
  double r = fma(-qt, dg, a);  // exact: a - qt*dg is representable
  if (r < 0.0) r = fma(-(qt - 1.0), dg, a);  // a/dg rounded up to an integer
  i = (int)qt;
  f = (float)(r * inv_dg);
  if (i == n - 1) {
    i = n - 2;
    f = 1.0f;
  }
}

__device__ __forceinline__ float lerpf(float t, float a, float b) { return fmaf(t, b - a, a); }

// Index and fraction of both axes in fixed point.  Along a ray the reference's quotients
// q = (x + 1)/dg are linear in the point index i, so Q = q + 2^-22 is kept with 32 fraction
// bits and stepped by one 64-bit add per point (D = round(sx/dg * 2^32)).  Against the
// reference's q (x rounded as ray_grid/src/integral.c:41-43 rounds it) Q errs by at most
// (4n + 1) 2^-33 plus roundings below 1e-12: under 2^-22 for n <= RAY_FIX_N.  Wherever the
// fraction of Q is at least 2^-21, q then lies strictly between the integer part k of Q and
// k + 1: the index is k, and the fraction, frac(Q) - 2^-22, is the reference's
// fmod(x + 1, dg)/dg to within 2.4e-7 (an FP32 lerp weight either way).  Points nearer an
// integer, among them every point where the reference's rounding ties, its fmod quirk and
// x = 1, take the exact binary64 path (ray_axis).
#define RAY_FIX_N 500

__device__ __forceinline__ long long ray_fix(double q) {  // 0 <= q < 2^31
  const double fl = floor(q);
  const double fr = fmin(rint((q - fl) * 4294967296.0), 4294967295.0);
  return (long long)(((unsigned long long)(unsigned)(int)fl << 32) | (unsigned)fr) + 1024ll;
}

template <bool POW2>
__device__ __forceinline__ void ray_point(bool fixed, long long Qx, long long Qy, double px,
                                          double py, double sx, double sy, int i, int n,
                                          double dg, double inv_dg, double h, int& ix, int& iy,
                                          float& fx, float& fy) {
  const unsigned lx = (unsigned)Qx, ly = (unsigned)Qy;
  if (fixed && min(lx, ly) >= 2048u) {
    ix = (int)(Qx >> 32);
    iy = (int)(Qy >> 32);
    fx = fmaf((float)lx, 2.3283064365386963e-10f, -2.384185791015625e-7f);
    fy = fmaf((float)ly, 2.3283064365386963e-10f, -2.384185791015625e-7f);
  } else {
    const double fi = (double)i;
    const double x = px + sx * fi, y = py + sy * fi;
    ray_axis<POW2>(x, n, dg, inv_dg, h, ix, fx);
    ray_axis<POW2>(y, n, dg, inv_dg, h, iy, fy);
  }
}

// the fixed-point start (point lo) and step of both axes
__device__ __forceinline__ void ray_fix_start(double px, double py, double sx, double sy, int lo,
                                              int n, long long& Qx, long long& Qy, long long& Dx,
                                              long long& Dy) {
  const double hg = 0.5 * (n - 1);  // 1/dg
  const double flo = (double)lo;
  Qx = ray_fix((px + sx * flo + 1.0) * hg);
  Qy = ray_fix((py + sy * flo + 1.0) * hg);
  Dx = (long long)rint(sx * hg * 4294967296.0);
  Dy = (long long)rint(sy * hg * 4294967296.0);
}

// The exact range [lo, hi] of i in [-2n, 2n] whose point x = px + sx*i, y = py + sy*i (each
// rounded as the reference rounds it) passes |x| <= 1 and |y| <= 1.  x and y are monotone
// in i, so the range is an interval; the continuous bounds, widened by a tolerance far
// above the rounding of x and y, contain it, and cube_ends settles its ends exactly.  The
// tolerance matters for a ray on a face of the cube (s = +-1 along an axis): there sx is
// ~1e-17 and x = s + sx*i rounds to exactly +-1 for dozens of points past the continuous
// bound.  An empty range comes back as lo > hi.
__device__ __forceinline__ void ray_range(double px, double py, double sx, double sy, int n,
                                          int& lo, int& hi) {
  const double tol = 1e-9;
  lo = -2 * n;
  hi = 2 * n;
  const double tl = fmax(fabs(sx) > 0 ? fmin((-1.0 - tol - px) / sx, (1.0 + tol - px) / sx) : -1e300,
                         fabs(sy) > 0 ? fmin((-1.0 - tol - py) / sy, (1.0 + tol - py) / sy) : -1e300);
  const double th = fmin(fabs(sx) > 0 ? fmax((-1.0 - tol - px) / sx, (1.0 + tol - px) / sx) : 1e300,
                         fabs(sy) > 0 ? fmax((-1.0 - tol - py) / sy, (1.0 + tol - py) / sy) : 1e300);
  // clamped to +-4n before the conversion: a ray far outside gives bounds ~1e17
  if (tl > -1e299) lo = max(lo, (int)fmin(fmax(ceil(tl) - 1.0, -4.0 * n), 4.0 * n));
  if (th < 1e299) hi = min(hi, (int)fmax(fmin(floor(th) + 1.0, 4.0 * n), -4.0 * n));
  cube_ends(px, sx, true, true, lo, hi);
  cube_ends(py, sy, true, true, lo, hi);
}

// vol32: [z][x][y] FP32; zi / zc: the slices' index and fraction (reference rounding;
// zi < 0 marks a slice with |z| > 1, whose rays are 0);
// trig: [nphi][4] cos p, sin p, cos(p+pi/2), sin(p+pi/2) (host libm); out [z][s][phi] FP32
template <bool LDS, bool POW2>
__global__ __launch_bounds__(RAY_THREADS) void k_ray_f32(const float* __restrict__ vol32, int n,
                                                         const int* __restrict__ zi,
                                                         const double* __restrict__ zc,
                                                         const double* __restrict__ shift,
                                                         int nshift,
                                                         const double* __restrict__ trig,
                                                         int nphi, float* __restrict__ out) {
  extern __shared__ float plane[];
  const int iz = blockIdx.y;
  const int kz = zi[iz];
  const long long nrays = (long long)nshift * nphi;
  const long long t = (long long)blockIdx.x * RAY_THREADS + threadIdx.x;
  if (kz < 0) {  // |z| > 1: every point fails the cube test
    if (t < nrays) out[((size_t)iz * nshift + (int)(t % nshift)) * nphi + (int)(t / nshift)] = 0.f;
    return;
  }
  const float cz = (float)zc[iz];
  const size_t nn = (size_t)n * n;
  const float* __restrict__ v0 = vol32 + (size_t)kz * nn;
  const float* __restrict__ v1 = v0 + nn;
  // The LDS plane's rows have an odd pitch (n | 1): lanes on neighbouring rays whose points
  // differ in x then fall in different banks (with pitch n = 128 every row starts in bank 0)
  const int P = ray_pitch(n);
  if (LDS) {
    // element k = x n + y of the slice goes to x P + y; (x, y) steps by RAY_THREADS elements
    const int dq = RAY_THREADS / n, dr = RAY_THREADS % n;
    int x = threadIdx.x / n, y = threadIdx.x % n;
    for (int k = threadIdx.x; k < (int)nn; k += RAY_THREADS) {
      plane[x * P + y] = lerpf(cz, v0[k], v1[k]);
      x += dq;
      y += dr;
      if (y >= n) {
        y -= n;
        ++x;
      }
    }
    __syncthreads();
  }
  if (t >= nrays) return;
  const int is = (int)(t % nshift), ip = (int)(t / nshift);
  const double c = trig[4 * ip], sn = trig[4 * ip + 1], dx = trig[4 * ip + 2],
               dy = trig[4 * ip + 3];
  const double s = shift[is];
  const double px = s * c, py = s * sn;
  const double delta = (2.0 / (n - 1)) / (2.0);
  const double dg = 2.0 / (n - 1), inv_dg = 1.0 / dg, h = (n - 1) / 2.0;
  const double sx = delta * dx, sy = delta * dy;
  int lo, hi;
  ray_range(px, py, sx, sy, n, lo, hi);
  const bool fixed = n <= RAY_FIX_N;
  long long Qx = 0, Qy = 0, Dx = 0, Dy = 0;
  if (fixed && lo <= hi) ray_fix_start(px, py, sx, sy, lo, n, Qx, Qy, Dx, Dy);
  float acc = 0.f, comp = 0.f;
  for (int i = lo; i <= hi; ++i, Qx += Dx, Qy += Dy) {
    int ix, iy;
    float fx, fy;
    ray_point<POW2>(fixed, Qx, Qy, px, py, sx, sy, i, n, dg, inv_dg, h, ix, iy, fx, fy);
    float p00, p10, p01, p11;
    if (LDS) {
      const int o = ix * P + iy;
      p00 = plane[o];
      p10 = plane[o + P];
      p01 = plane[o + 1];
      p11 = plane[o + P + 1];
    } else {
      const int o = ix * n + iy;
      p00 = lerpf(cz, v0[o], v1[o]);
      p10 = lerpf(cz, v0[o + n], v1[o + n]);
      p01 = lerpf(cz, v0[o + 1], v1[o + 1]);
      p11 = lerpf(cz, v0[o + n + 1], v1[o + n + 1]);
    }
    const float v = lerpf(fy, lerpf(fx, p00, p10), lerpf(fx, p01, p11));
    const float yk = v - comp;
    const float tk = acc + yk;
    comp = (tk - acc) - yk;
    acc = tk;
  }
  out[((size_t)iz * nshift + is) * nphi + ip] = (float)(((double)acc - (double)comp) * delta);
}

template <bool LDS>
static void launch_ray(const float* vol32, int n, const int* zi, const double* zc, int nz,
                       const double* shift, int nshift, const double* trig, int nphi, float* out,
                       hipStream_t st) {
  const long long nrays = (long long)nshift * nphi;
  const dim3 grid((unsigned)((nrays + RAY_THREADS - 1) / RAY_THREADS), (unsigned)nz);
  const size_t lds = LDS ? (size_t)n * ray_pitch(n) * sizeof(float) : 0;
  const int m = n - 1;
  if ((m & (m - 1)) == 0) {
    if (lds > 65536)
      (void)hipFuncSetAttribute((const void*)k_ray_f32<LDS, true>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    k_ray_f32<LDS, true><<<grid, RAY_THREADS, lds, st>>>(vol32, n, zi, zc, shift, nshift, trig, nphi, out);
  } else {
    if (lds > 65536)
      (void)hipFuncSetAttribute((const void*)k_ray_f32<LDS, false>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    k_ray_f32<LDS, false><<<grid, RAY_THREADS, lds, st>>>(vol32, n, zi, zc, shift, nshift, trig, nphi, out);
  }
}

extern "C" int wrtk_ray_f32(const float* vol32, int n, const int* zi, const double* zc, int nz,
                            const double* shift, int nshift, const double* trig, int nphi,
                            float* out, hipStream_t st) {
  if (n < 2 || nz < 1 || nz > 65535 || nshift < 1 || nphi < 1) return WRT_E_ARG;
  if ((long long)nshift * nphi > 65535LL * RAY_THREADS * 64) return WRT_E_ARG;
  if (n <= RAY_LDS_N)
    launch_ray<true>(vol32, n, zi, zc, nz, shift, nshift, trig, nphi, out, st);
  else
    launch_ray<false>(vol32, n, zi, zc, nz, shift, nshift, trig, nphi, out, st);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

// ------------------------------------------------------ transpose of the line integrals ---
// The exact transpose of k_ray_f32: every point of a ray adds g * w to the four (x, y)
// corners of its cell with the bilinear weights w the forward reads them with, into the
// slice's plane; the plane then goes to the two z-nodes of the slice with weights 1 - cz
// and cz, times delta.  Index and fraction decisions are the forward's (the reference's).
// Work layout: blockIdx.y = slice, blockIdx.x = one of the slice's ray blocks; a thread
// walks rays t = blockIdx.x * RAY_T_THREADS + threadIdx.x + k * gridDim.x * RAY_T_THREADS,
// ray t being shift (r * sstride) mod nshift, r = t mod nshift, of direction t / nshift:
// with sstride coprime to nshift this visits every ray once, and puts neighbouring lanes on
// rays sstride shifts apart, whose cells (and corners) mostly differ, so that fewer lanes
// of one instruction add to one LDS cell (the LDS serialises those).
// For n <= 143 the plane is an LDS FP64 accumulator (ds_add_f64, which needs no returned
// value), added to the FP64 output accumulator once per workgroup; larger volumes add every
// point's weights to the FP64 accumulator directly.  acc: n^3 FP64 [x][y][z], zeroed by the
// caller.
template <bool LDS, bool POW2>
__global__ __launch_bounds__(RAY_T_THREADS) void k_ray_T_f32(const float* __restrict__ g, int n,
                                                           const int* __restrict__ zi,
                                                           const double* __restrict__ zc,
                                                           const double* __restrict__ shift,
                                                           int nshift,
                                                           const double* __restrict__ trig,
                                                           int nphi, int sstride,
                                                           double* __restrict__ acc) {
  extern __shared__ double tplane[];
  const int iz = blockIdx.y;
  const int kz = zi[iz];
  if (kz < 0) return;  // |z| > 1: no point of the slice passes the cube test
  const double czd = zc[iz];
  const double delta = (2.0 / (n - 1)) / (2.0);
  const size_t nn = (size_t)n * n;
  const int P = ray_pitch(n);  // the LDS plane's odd row pitch, as in k_ray_f32
  if (LDS) {
    for (int k = threadIdx.x; k < n * P; k += RAY_T_THREADS) tplane[k] = 0.0;
    __syncthreads();
  }
  const long long nrays = (long long)nshift * nphi;
  const double dg = 2.0 / (n - 1), inv_dg = 1.0 / dg, h = (n - 1) / 2.0;
  const double w0 = (1.0 - czd) * delta, w1 = czd * delta;  // z-node weights (global path)
  for (long long t = (long long)blockIdx.x * RAY_T_THREADS + threadIdx.x; t < nrays;
       t += (long long)gridDim.x * RAY_T_THREADS) {
    const int is = (int)((t % nshift) * sstride % nshift), ip = (int)(t / nshift);
    const float gv = g[((size_t)iz * nshift + is) * nphi + ip];
    if (gv == 0.f) continue;
    const double c = trig[4 * ip], sn = trig[4 * ip + 1], dx = trig[4 * ip + 2],
                 dy = trig[4 * ip + 3];
    const double s = shift[is];
    const double px = s * c, py = s * sn;
    const double sx = delta * dx, sy = delta * dy;
    int lo, hi;
    ray_range(px, py, sx, sy, n, lo, hi);
    const bool fixed = n <= RAY_FIX_N;
    long long Qx = 0, Qy = 0, Dx = 0, Dy = 0;
    if (fixed && lo <= hi) ray_fix_start(px, py, sx, sy, lo, n, Qx, Qy, Dx, Dy);
    for (int i = lo; i <= hi; ++i, Qx += Dx, Qy += Dy) {
      int ix, iy;
      float fx, fy;
      ray_point<POW2>(fixed, Qx, Qy, px, py, sx, sy, i, n, dg, inv_dg, h, ix, iy, fx, fy);
      // the forward's value: lerp(fy, lerp(fx, p00, p10), lerp(fx, p01, p11))
      const float gy0 = gv * (1.f - fy), gy1 = gv * fy;
      const float w00 = gy0 * (1.f - fx), w10 = gy0 * fx, w01 = gy1 * (1.f - fx), w11 = gy1 * fx;
      if (LDS) {
        // LDS FP64 atomic adds: no returned value to wait for (the FP32 LDS atomic runs lane
        // by lane on gfx950)
        const int o = ix * P + iy;
        unsafeAtomicAdd(&tplane[o], (double)w00);
        unsafeAtomicAdd(&tplane[o + P], (double)w10);
        unsafeAtomicAdd(&tplane[o + 1], (double)w01);
        unsafeAtomicAdd(&tplane[o + P + 1], (double)w11);
      } else {
        const int o = ix * n + iy;
        double* p = acc + (size_t)o * n + kz;  // [x][y][z]
        const float wv[4] = {w00, w10, w01, w11};
        const size_t off[4] = {0, nn, (size_t)n, nn + n};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (wv[q] == 0.f) continue;
          unsafeAtomicAdd(p + off[q], (double)wv[q] * w0);
          unsafeAtomicAdd(p + off[q] + 1, (double)wv[q] * w1);
        }
      }
    }
  }
  if (LDS) {
    __syncthreads();
    // element k = x n + y of the plane sits at x P + y; (x, y) steps by RAY_T_THREADS
    const int dq = RAY_T_THREADS / n, dr = RAY_T_THREADS % n;
    int x = threadIdx.x / n, y = threadIdx.x % n;
    for (int k = threadIdx.x; k < (int)nn; k += RAY_T_THREADS) {
      const double v = tplane[x * P + y];
      x += dq;
      y += dr;
      if (y >= n) {
        y -= n;
        ++x;
      }
      if (v == 0.0) continue;
      double* p = acc + (size_t)k * n + kz;  // k = x*n + y
      unsafeAtomicAdd(p, v * w0);
      unsafeAtomicAdd(p + 1, v * w1);
    }
  }
}

template <bool LDS>
static void launch_ray_T(const float* g, int n, const int* zi, const double* zc, int nz,
                         const double* shift, int nshift, const double* trig, int nphi,
                         double* acc, hipStream_t st) {
  const long long nrays = (long long)nshift * nphi;
  // LDS plane: about 512 workgroups in all (two rounds of one per CU), since every workgroup
  // ends by adding its n^2 plane to the volume with global atomics; without: about 2048.
  // At least one ray per thread of each.
  long long per = ((LDS ? 512 : 2048) + nz - 1) / nz;
  const long long most = (nrays + RAY_T_THREADS - 1) / RAY_T_THREADS;
  if (per > most) per = most;
  if (per < 1) per = 1;
  const dim3 grid((unsigned)per, (unsigned)nz);
  const size_t lds = LDS ? (size_t)n * ray_pitch(n) * sizeof(double) : 0;
  // the smallest stride >= 4 coprime to nshift (rays 4 shifts apart: with 257 shifts over
  // 128 cells, 2 cells apart; a stride of 8 measured the same)
  int sstride = 4;
  for (;; ++sstride) {
    int a = sstride, b = nshift;
    while (b) { const int r = a % b; a = b; b = r; }
    if (a == 1) break;
  }
  const int m = n - 1;
  if ((m & (m - 1)) == 0) {
    if (lds > 65536)
      (void)hipFuncSetAttribute((const void*)k_ray_T_f32<LDS, true>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    k_ray_T_f32<LDS, true><<<grid, RAY_T_THREADS, lds, st>>>(g, n, zi, zc, shift, nshift, trig, nphi, sstride, acc);
  } else {
    if (lds > 65536)
      (void)hipFuncSetAttribute((const void*)k_ray_T_f32<LDS, false>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    k_ray_T_f32<LDS, false><<<grid, RAY_T_THREADS, lds, st>>>(g, n, zi, zc, shift, nshift, trig, nphi, sstride, acc);
  }
}

extern "C" int wrtk_ray_T_f32(const float* g, int n, const int* zi, const double* zc, int nz,
                              const double* shift, int nshift, const double* trig, int nphi,
                              double* acc, hipStream_t st) {
  if (n < 2 || nz < 1 || nz > 65535 || nshift < 1 || nphi < 1) return WRT_E_ARG;
  if (n <= RAY_T_LDS_N)
    launch_ray_T<true>(g, n, zi, zc, nz, shift, nshift, trig, nphi, acc, st);
  else
    launch_ray_T<false>(g, n, zi, zc, nz, shift, nshift, trig, nphi, acc, st);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

// FP64 vol[x][y][z] -> FP32 [z][x][y] (each z-plane contiguous)
__global__ __launch_bounds__(256) void k_vol_zxy_f32(const double* __restrict__ vol, int n,
                                                     float* __restrict__ out) {
  const size_t total = (size_t)n * n * n;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (size_t)gridDim.x * blockDim.x) {
    const size_t z = k / ((size_t)n * n), r = k % ((size_t)n * n);  // r = x*n + y
    out[k] = (float)vol[r * n + z];
  }
}

extern "C" int wrtk_vol_zxy_f32(const double* vol, int n, float* out, hipStream_t st) {
  const size_t total = (size_t)n * n * n;
  size_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  k_vol_zxy_f32<<<(unsigned)blocks, 256, 0, st>>>(vol, n, out);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}
