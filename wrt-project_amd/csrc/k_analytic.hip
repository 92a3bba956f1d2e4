// k_analytic.hip -- the analytic programs: plane integrals (radon_analytic) and layer-by-layer
// line integrals (ray_analytic) of a test function given as a sum of terms (wrt_fn_term),
// FP64.
//
// plane_integral (radon_analytic/src/integral.c:25-56) sums value * delta * delta over the
// lattice point = shift*n + (delta*v1)*i1 + (delta*v2)*i2, i1, i2 in [-2g, 2g), delta =
// (2/(g-1))/2; ray_integral (ray_analytic/integral.c:10-47) sums value * delta over
// point = in_point + (delta*dir)*i, i in [-2g, 2g].  Every point is formed with the
// reference's operations in its order (no contraction), and the test function with the
// example files' operations.  Each term vanishes outside a ball of squared radius rb2
// (computed on the host); a point outside every term's ball has value 0, which adds exactly
// nothing to the reference's sum, so only the lattice rows and columns within those balls
// (plus a margin of two points) are visited.
//
// Work layout: planes -- one wavefront per plane, lanes across 64 columns of a row, rows in
// turn, lane sums reduced over the wavefront; rays -- one thread per ray, its points in the
// reference's order (so a ray's sum is the reference's sum whenever the device's exp/sin
// agree with the host's).
#include "wrt_common.hpp"
#include "wrt_dev.h"

#define AN_WAVES 4

// exponential_bump.c:5-16: exp(-1/(r2max - |x-c|^2)) inside |x-c|^2 < r2max - eps (the
// example squares with pow(., 2); x*x is the correctly rounded square)
__device__ __forceinline__ double an_exp_bump(const wrt_fn_term& t, double x, double y, double z) {
  const double xx = x - t.c[0], yy = y - t.c[1], zz = z - t.c[2];
  const double x2 = xx * xx, y2 = yy * yy, z2 = zz * zz;
  if ((x2 + y2 + z2) < (t.p[0] - t.p[1])) return exp((-1) / (t.p[0] - x2 - y2 - z2));
  return 0.0;
}

// sine_bump_template: sinusoidal_bump.c:3-19
__device__ __forceinline__ double an_sine_bump(const wrt_fn_term& t, double x, double y, double z) {
  const double amplitude = t.p[0], rin = t.p[1], rout = t.p[2];
  const double dx = x - t.c[0], dy = y - t.c[1], dz = z - t.c[2];
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  if (r > rout) return 0.0;
  if ((r < rout) && (r > rin)) {
    const double tt = M_PI * (r - rout) / (rout - rin);
    return ((amplitude / (20 * M_PI)) *
            (-sin(6 * tt) / 3 + 3 * sin(4 * tt) - 15 * sin(2 * tt) + 20 * tt));
  }
  if (r < rin) return amplitude;
  return 0.0;
}

// the phantom's solid ellipsoid: density inside |diag(1/a) R^T (x - c)| <= 1
__device__ __forceinline__ double an_ellipsoid(const wrt_fn_term& t, double x, double y, double z) {
  const double dx = x - t.c[0], dy = y - t.c[1], dz = z - t.c[2];
  const double* R = t.p + 4;
  const double u = (R[0] * dx + R[3] * dy + R[6] * dz) / t.p[1];
  const double w = (R[1] * dx + R[4] * dy + R[7] * dz) / t.p[2];
  const double v = (R[2] * dx + R[5] * dy + R[8] * dz) / t.p[3];
  return (u * u + w * w + v * v <= 1.0) ? t.p[0] : 0.0;
}

// the terms summed left to right; a term whose ball the point is outside is 0 without
// evaluating it
__device__ __forceinline__ double an_eval(const wrt_fn_term* __restrict__ t,
                                          const double* __restrict__ rb2, int nt, double x,
                                          double y, double z) {
  double v = 0.0;
  for (int k = 0; k < nt; ++k) {
    const double dx = x - t[k].c[0], dy = y - t[k].c[1], dz = z - t[k].c[2];
    double f = 0.0;
    if (dx * dx + dy * dy + dz * dz < rb2[k]) {
      switch (t[k].kind) {
        case WRT_FN_EXP_BUMP: f = an_exp_bump(t[k], x, y, z); break;
        case WRT_FN_SINE_BUMP: f = an_sine_bump(t[k], x, y, z); break;
        default: f = an_ellipsoid(t[k], x, y, z); break;
      }
    }
    v = k == 0 ? f : v + f;
  }
  return v;
}

// dirs: [ndir][9] = n, delta*v1, delta*v2 (host libm, the reference's rounding)
__global__ __launch_bounds__(AN_WAVES * 64) void k_radon_analytic(
    const wrt_fn_term* __restrict__ terms, const double* __restrict__ rb2, int nt, int ngrid,
    const double* __restrict__ dirs, int ndir, const double* __restrict__ shift,
    long long nplanes, double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const long long plane = (long long)blockIdx.x * AN_WAVES + (threadIdx.x >> 6);
  if (plane >= nplanes) return;
  const int d = (int)(plane % ndir);
  const int is = (int)(plane / ndir);
  const double* D = dirs + 9 * (size_t)d;
  const double s = shift[is];
  const double Sx = s * D[0], Sy = s * D[1], Sz = s * D[2];
  const double ax = D[3], ay = D[4], az = D[5], bx = D[6], by = D[7], bz = D[8];
  const double delta = (2.0 / (ngrid - 1)) / 2.0;
  const double aa = ax * ax + ay * ay + az * az, bb = bx * bx + by * by + bz * bz;
  const int L0 = -2 * ngrid, L1 = 2 * ngrid - 1;
  // rows reached by some term's ball: the ball meets the plane in a disc of radius
  // sqrt(rb2 - dist^2) about the projection of the centre, row (c.a)/|a|^2
  double rlo = 1e300, rhi = -1e300;
  for (int k = 0; k < nt; ++k) {
    const double* c = terms[k].c;
    const double dist = c[0] * D[0] + c[1] * D[1] + c[2] * D[2] - s;
    const double rho2 = rb2[k] - dist * dist;
    if (rho2 <= 0.0) continue;
    const double u1 = (c[0] * ax + c[1] * ay + c[2] * az) / aa;
    const double w = sqrt(rho2 / aa);
    rlo = fmin(rlo, u1 - w);
    rhi = fmax(rhi, u1 + w);
  }
  double sum = 0.0;
  if (rlo <= rhi) {
    // clamped to the lattice before the conversion (a far term gives bounds ~1e9)
    const int R0 = (int)fmin(fmax(floor(rlo) - 2.0, (double)L0), (double)L1 + 1.0);
    const int R1 = (int)fmax(fmin(ceil(rhi) + 2.0, (double)L1), (double)L0 - 1.0);
    for (int i1 = R0; i1 <= R1; ++i1) {
      const double f1 = (double)i1;
      // columns of this row within some term's ball
      double clo = 1e300, chi = -1e300;
      for (int k = 0; k < nt; ++k) {
        const double* c = terms[k].c;
        const double dist = c[0] * D[0] + c[1] * D[1] + c[2] * D[2] - s;
        const double u1 = (c[0] * ax + c[1] * ay + c[2] * az) / aa;
        const double dr = (f1 - u1) * (f1 - u1) * aa;
        const double rho2 = rb2[k] - dist * dist - dr;
        if (rho2 <= -4.0 * aa) continue;  // farther than two rows from the disc
        const double u2 = (c[0] * bx + c[1] * by + c[2] * bz) / bb;
        const double w = sqrt(fmax(rho2, 0.0) / bb);
        clo = fmin(clo, u2 - w);
        chi = fmax(chi, u2 + w);
      }
      if (clo > chi) continue;
      const int C0 = (int)fmin(fmax(floor(clo) - 2.0, (double)L0), (double)L1 + 1.0);
      const int C1 = (int)fmax(fmin(ceil(chi) + 2.0, (double)L1), (double)L0 - 1.0);
      const double rx = Sx + ax * f1, ry = Sy + ay * f1, rz = Sz + az * f1;
      for (int i2 = C0 + lane; i2 <= C1; i2 += 64) {
        const double f2 = (double)i2;
        const double x = rx + bx * f2, y = ry + by * f2, z = rz + bz * f2;
        const double value = an_eval(terms, rb2, nt, x, y, z);
        sum += value * delta * delta;
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
  if (lane == 0) out[plane] = sum;
}

extern "C" int wrtk_radon_analytic(const wrt_fn_term* terms, const double* rb2, int nt, int ngrid,
                                   const double* dirs, int ndir, const double* shift, int nshift,
                                   double* out, hipStream_t st) {
  const long long nplanes = (long long)ndir * nshift;
  if (nplanes <= 0) return WRT_OK;
  const long long blocks = (nplanes + AN_WAVES - 1) / AN_WAVES;
  if (blocks >= (1LL << 31)) return WRT_E_ARG;
  k_radon_analytic<<<(unsigned)blocks, AN_WAVES * 64, 0, st>>>(terms, rb2, nt, ngrid, dirs, ndir,
                                                               shift, nplanes, out);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

// trig: [nphi][4] = cos p, sin p, cos(p + pi/2), sin(p + pi/2) (host libm); out[z][s][phi]
__global__ __launch_bounds__(256) void k_ray_analytic(
    const wrt_fn_term* __restrict__ terms, const double* __restrict__ rb2, int nt, int ngrid,
    const double* __restrict__ zs, int nz, const double* __restrict__ shift, int nshift,
    const double* __restrict__ trig, int nphi, double* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)nz * nshift * nphi;
  if (t >= total) return;
  const int ip = (int)(t % nphi);
  const int is = (int)((t / nphi) % nshift);
  const int iz = (int)(t / ((long long)nphi * nshift));
  const double c0 = trig[4 * ip], s0 = trig[4 * ip + 1], dx = trig[4 * ip + 2], dy = trig[4 * ip + 3];
  const double sv = shift[is];
  const double px = sv * c0, py = sv * s0, pz = zs[iz];
  const double delta = (2.0 / (ngrid - 1)) / 2.0;
  const double sx = delta * dx, sy = delta * dy, sz = delta * 0.0;
  // points reached by some term's ball (the direction is a unit vector in the slice)
  double ilo = 1e300, ihi = -1e300;
  for (int k = 0; k < nt; ++k) {
    const double* c = terms[k].c;
    const double wx = c[0] - px, wy = c[1] - py, wz = c[2] - pz;
    const double along = wx * dx + wy * dy;
    const double perp2 = wx * wx + wy * wy + wz * wz - along * along;
    const double h2 = rb2[k] - perp2;
    if (h2 <= 0.0) continue;
    const double h = sqrt(h2);
    ilo = fmin(ilo, (along - h) / delta);
    ihi = fmax(ihi, (along + h) / delta);
  }
  double sum = 0.0;
  if (ilo <= ihi) {
    const int lo = (int)fmin(fmax(floor(ilo) - 2.0, -2.0 * ngrid), 2.0 * ngrid + 1.0);
    const int hi = (int)fmax(fmin(ceil(ihi) + 2.0, 2.0 * ngrid), -2.0 * ngrid - 1.0);
    for (int i = lo; i <= hi; ++i) {
      const double fi = (double)i;
      const double x = px + sx * fi, y = py + sy * fi, z = pz + sz * fi;
      const double value = an_eval(terms, rb2, nt, x, y, z);
      sum += value * delta;
    }
  }
  out[t] = sum;
}

extern "C" int wrtk_ray_analytic(const wrt_fn_term* terms, const double* rb2, int nt, int ngrid,
                                 const double* z, int nz, const double* shift, int nshift,
                                 const double* trig, int nphi, double* out, hipStream_t st) {
  const long long total = (long long)nz * nshift * nphi;
  if (total <= 0) return WRT_OK;
  const long long blocks = (total + 255) / 256;
  if (blocks >= (1LL << 31)) return WRT_E_ARG;
  k_ray_analytic<<<(unsigned)blocks, 256, 0, st>>>(terms, rb2, nt, ngrid, z, nz, shift, nshift,
                                                   trig, nphi, out);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}
