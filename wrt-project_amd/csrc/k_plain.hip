// k_plain.hip -- the plain FP64 forms of the four operators, the Gauss-Legendre rule,
// the natural-spline row solver, and the FP64 exact transposes.
//
// Plain means: the reference's arithmetic, operation for operation, in IEEE binary64 on the
// GPU (-ffp-contract=off, correctly rounded division), reading the volume and data from HBM
// with no staging, no clipping and no reuse.  ray/reduction/adjoint sum in the reference's
// sequential order (one thread per output, bit-identical to the oracle); radon sums one
// plane with a 64-lane wavefront and a shuffle tree (rounding-level difference only).
#include "wrt_common.hpp"
#include "wrt_dev.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

// ---------------------------------------------------------------- radon_grid (plain) ---
// plane_integral (radon_grid/src/integral.c:26-59): 4n x 4n lattice, i1,i2 in [-2n, 2n),
// point = ((s*n) + (a*i1)) + (b*i2) with a = delta*v1, b = delta*v2 (host-rounded),
// term (v*delta)*delta.  One wavefront per plane; lanes stride the i2 columns.
__global__ __launch_bounds__(256) void k_radon_plain(const double* __restrict__ vol, int n,
                                                     const double* __restrict__ dirs, int ndir,
                                                     const double* __restrict__ shift,
                                                     long long nplanes,
                                                     double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const long long plane = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (plane >= nplanes) return;
  const int d = (int)(plane % ndir);
  const int is = (int)(plane / ndir);
  const double* D = dirs + 9 * (size_t)d;
  const double s = shift[is];
  const double Sx = s * D[0], Sy = s * D[1], Sz = s * D[2];
  const double ax = D[3], ay = D[4], az = D[5], bx = D[6], by = D[7], bz = D[8];
  const double delta = (1.0 / (n));
  const double dc = 2.0 / n;
  const int half = 2 * n;
  double acc = 0.0;
  for (int i1 = -half; i1 < half; ++i1) {
    const double f1 = (double)i1;
    const double rx = Sx + ax * f1, ry = Sy + ay * f1, rz = Sz + az * f1;
    for (int i2 = -half + lane; i2 < half; i2 += 64) {
      const double f2 = (double)i2;
      const double x = rx + bx * f2, y = ry + by * f2, z = rz + bz * f2;
      const double v = wrt_zero_interp(vol, n, dc, x, y, z);
      acc += v * delta * delta;
    }
  }
  acc = wrt_wave_sum(acc);
  if (lane == 0) out[plane] = acc;
}

extern "C" int wrtk_radon_plain(const double* vol, int n, const double* dirs, int ndir,
                                const double* shift, int nshift, double* out, hipStream_t st) {
  const long long nplanes = (long long)ndir * nshift;
  if (nplanes <= 0) return WRT_OK;
  const long long blocks = (nplanes + 3) / 4;
  if (blocks * 256 >= (1LL << 32)) return WRT_E_ARG;  // a dispatch holds < 2^32 work-items
  k_radon_plain<<<dim3((unsigned)blocks), dim3(256), 0, st>>>(vol, n, dirs, ndir, shift,
                                                               nplanes, out);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

// ------------------------------------------------------------------ ray_grid (plain) ---
// ray_transform (ray_grid/src/integral.c:11-50): 4n+1 samples, i in [-2n, 2n],
// point = in_point + (delta*dir)*i, term trilinear*delta, summed in order.  One thread per
// ray; lanes take consecutive shifts of one (z, phi) so that neighbouring lanes sample
// neighbouring voxels.
__global__ __launch_bounds__(256) void k_ray_plain(const double* __restrict__ vol, int n,
                                                   const double* __restrict__ zs, int nz,
                                                   const double* __restrict__ shift, int nshift,
                                                   const double* __restrict__ trig, int nphi,
                                                   double* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)nz * nphi * nshift;
  if (t >= total) return;
  const int is = (int)(t % nshift);
  const int ip = (int)((t / nshift) % nphi);
  const int iz = (int)(t / ((long long)nshift * nphi));
  const double c = trig[4 * ip], sn = trig[4 * ip + 1], dx = trig[4 * ip + 2],
               dy = trig[4 * ip + 3];
  const double s = shift[is];
  const double px = s * c, py = s * sn, pz = zs[iz];
  const double delta = (2.0 / (n - 1)) / (2.0);
  const double dg = 2.0 / (n - 1);
  const double sx = delta * dx, sy = delta * dy, sz = delta * 0.0;
  const int half = 2 * n;
  double acc = 0.0;
  for (int i = -half; i < half + 1; ++i) {
    const double fi = (double)i;
    const double x = px + sx * fi, y = py + sy * fi, z = pz + sz * fi;
    acc += wrt_trilinear(vol, n, dg, x, y, z) * delta;
  }
  out[((size_t)iz * nshift + is) * nphi + ip] = acc;
}

extern "C" int wrtk_ray_plain(const double* vol, int n, const double* z, int nz,
                              const double* shift, int nshift, const double* trig, int nphi,
                              double* out, hipStream_t st) {
  const long long total = (long long)nz * nphi * nshift;
  if (total <= 0) return WRT_OK;
  k_ray_plain<<<dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st>>>(
      vol, n, z, nz, shift, nshift, trig, nphi, out);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

// ---------------------------------------------------------------- Gauss-Legendre rule ---
// Double-double Newton iteration on P_n; the rounded node and weight are the correctly
// rounded values of the exact rule (what GSL's tabulated rules hold).
struct ddv {
  double hi, lo;
};
__device__ __forceinline__ ddv dd_qsum(double a, double b) {
  const double s = a + b;
  return ddv{s, b - (s - a)};
}
__device__ __forceinline__ ddv dd_tsum(double a, double b) {
  const double s = a + b, bb = s - a;
  return ddv{s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ ddv dd_tprod(double a, double b) {
  const double p = a * b;
  return ddv{p, fma(a, b, -p)};
}
__device__ __forceinline__ ddv dd_add(ddv a, ddv b) {
  const ddv s = dd_tsum(a.hi, b.hi);
  return dd_qsum(s.hi, s.lo + (a.lo + b.lo));
}
__device__ __forceinline__ ddv dd_sub(ddv a, ddv b) { return dd_add(a, ddv{-b.hi, -b.lo}); }
__device__ __forceinline__ ddv dd_mul(ddv a, ddv b) {
  const ddv p = dd_tprod(a.hi, b.hi);
  return dd_qsum(p.hi, p.lo + (a.hi * b.lo + a.lo * b.hi));
}
__device__ __forceinline__ ddv dd_muld(ddv a, double b) {
  const ddv p = dd_tprod(a.hi, b);
  return dd_qsum(p.hi, p.lo + a.lo * b);
}
__device__ __forceinline__ ddv dd_div(ddv a, ddv b) {
  const double q1 = a.hi / b.hi;
  ddv r = dd_sub(a, dd_muld(b, q1));
  const double q2 = r.hi / b.hi;
  r = dd_sub(r, dd_muld(b, q2));
  const double q3 = r.hi / b.hi;
  return dd_add(dd_qsum(q1, q2), ddv{q3, 0.0});
}
__device__ void dd_legendre(int n, ddv x, ddv* pn, ddv* pm) {
  ddv p0{1.0, 0.0}, p1 = x;
  for (int k = 1; k < n; ++k) {
    const ddv t = dd_muld(dd_mul(x, p1), (double)(2 * k + 1));
    const ddv u = dd_muld(p0, (double)k);
    const ddv p2 = dd_div(dd_sub(t, u), ddv{(double)(k + 1), 0.0});
    p0 = p1;
    p1 = p2;
  }
  *pn = p1;
  *pm = p0;
}

__global__ void k_gl(int n, double* __restrict__ xo, double* __restrict__ wo) {
  const int m = n / 2;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > m) return;
  if (k == m) {  // middle node of an odd rule
    if (n & 1) {
      ddv pn, pm;
      dd_legendre(n, ddv{0.0, 0.0}, &pn, &pm);
      const ddv npm = dd_muld(pm, (double)n);
      const ddv w = dd_div(ddv{2.0, 0.0}, dd_mul(npm, npm));
      xo[m] = 0.0;
      wo[m] = w.hi;
    }
    return;
  }
  double x = cos(M_PI * (k + 0.75) / (n + 0.5));
  for (int it = 0; it < 100; ++it) {
    double p0 = 1.0, p1 = x;
    for (int j = 1; j < n; ++j) {
      const double p2 = ((2 * j + 1) * x * p1 - j * p0) / (j + 1);
      p0 = p1;
      p1 = p2;
    }
    const double dp = n * (x * p1 - p0) / (x * x - 1.0);
    const double dx = p1 / dp;
    x -= dx;
    if (fabs(dx) <= 1e-16) break;
  }
  ddv X{x, 0.0};
  for (int it = 0; it < 3; ++it) {
    ddv pn, pm;
    dd_legendre(n, X, &pn, &pm);
    const ddv x2m1 = dd_sub(dd_mul(X, X), ddv{1.0, 0.0});
    const ddv dp = dd_div(dd_muld(dd_sub(dd_mul(X, pn), pm), (double)n), x2m1);
    X = dd_sub(X, dd_div(pn, dp));
  }
  ddv pn, pm;
  dd_legendre(n, X, &pn, &pm);
  const ddv npm = dd_muld(pm, (double)n);
  const ddv num = dd_muld(dd_sub(ddv{1.0, 0.0}, dd_mul(X, X)), 2.0);
  const ddv w = dd_div(num, dd_mul(npm, npm));
  xo[n - 1 - k] = X.hi;
  wo[n - 1 - k] = w.hi;
  xo[k] = -X.hi;
  wo[k] = w.hi;
}

extern "C" int wrtk_gl(int n, double* x, double* w, hipStream_t st) {
  if (n < 1) return WRT_E_ARG;
  const int m = n / 2 + 1;
  k_gl<<<dim3((m + 63) / 64), dim3(64), 0, st>>>(n, x, w);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

// ------------------------------------------------------------- natural cubic spline ---
// The tridiagonal matrix of the natural spline depends only on the abscissae: factor it
// once (L D L^T, alpha = pivots, gamma = multipliers), then every row is a forward sweep,
// a diagonal solve and a back substitution.  Same operation order as the oracle.
__global__ void k_spline_factor(const double* __restrict__ x, int ns, double* __restrict__ fac) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int m = ns - 2;
  double* alp = fac;
  double* gam = fac + ns;
  for (int k = 0; k < m; ++k) {
    const double hl = x[k + 1] - x[k], hr = x[k + 2] - x[k + 1];
    const double dg = 2.0 * (hl + hr);
    alp[k] = (k == 0) ? dg : dg - hl * gam[k - 1];
    if (k < m - 1) gam[k] = hr / alp[k];
  }
}

__global__ __launch_bounds__(256) void k_spline_rows(const double* __restrict__ x, int ns,
                                                     const double* __restrict__ fac,
                                                     const double* __restrict__ y,
                                                     double* __restrict__ tab, int n_outer,
                                                     int n_inner, long long outer_stride,
                                                     long long es) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (long long)n_outer * n_inner) return;
  const long long o = r / n_inner, i = r % n_inner;
  const double* yr = y + o * outer_stride + i;
  double* t = tab + r * (long long)ns * 2;
  const double* alp = fac;
  const double* gam = fac + ns;
  const int m = ns - 2;
  for (int k = 0; k < ns; ++k) t[2 * k] = yr[k * es];
  t[1] = 0.0;
  t[2 * (ns - 1) + 1] = 0.0;
  for (int k = 0; k < m; ++k) {
    const double hl = x[k + 1] - x[k], hr = x[k + 2] - x[k + 1];
    const double y0 = t[2 * k], y1 = t[2 * k + 2], y2 = t[2 * k + 4];
    const double rhs = 3.0 * ((y2 - y1) / hr - (y1 - y0) / hl);
    t[2 * (k + 1) + 1] = (k == 0) ? rhs : rhs - gam[k - 1] * t[2 * k + 1];
  }
  for (int k = 0; k < m; ++k) t[2 * (k + 1) + 1] = t[2 * (k + 1) + 1] / alp[k];
  for (int k = m - 2; k >= 0; --k) t[2 * (k + 1) + 1] = t[2 * (k + 1) + 1] - gam[k] * t[2 * (k + 2) + 1];
}

extern "C" int wrtk_spline_factor(const double* x, int ns, double* fac, hipStream_t st) {
  if (ns < 3) return WRT_E_ARG;
  k_spline_factor<<<1, 1, 0, st>>>(x, ns, fac);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

extern "C" int wrtk_spline_rows(const double* x, int ns, const double* fac, const double* y,
                                double* tab, int n_outer, int n_inner, long long outer_stride,
                                long long es, hipStream_t st) {
  const long long rows = (long long)n_outer * n_inner;
  if (rows <= 0) return WRT_OK;
  k_spline_rows<<<dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, st>>>(
      x, ns, fac, y, tab, n_outer, n_inner, outer_stride, es);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

__device__ __forceinline__ double tab_eval(const double* __restrict__ x,
                                           const double* __restrict__ row, int i, double v) {
  const double xl = x[i], xh = x[i + 1];
  const double yl = row[2 * i], cl = row[2 * i + 1], yh = row[2 * i + 2], ch = row[2 * i + 3];
  const double h = xh - xl;
  const double t = v - xl;
  const double b = (yh - yl) / h - h * (ch + 2.0 * cl) / 3.0;
  const double d = (ch - cl) / (3.0 * h);
  return yl + t * (b + t * (cl + t * d));
}

// ------------------------------------------------------------ radon_reduction (FP64) ---
// radon_reduction + ray_interpolate (radon_reduction/slicing.c:5-119), one thread per
// plane, tau-steps in order.  Spline rows (z, phi) over the shift grid come from the
// packed table; the quadratic in z is the 3-point divided-difference polynomial.  The
// planes' offsets pshift[np] and the spline nodes shift[ns] are separate arrays; the
// reference's form passes the same grid for both.
__global__ __launch_bounds__(256) void k_reduction(const double* __restrict__ tab, int nz,
                                                   const double* __restrict__ shift, int ns,
                                                   const double* __restrict__ pshift, int np,
                                                   int nphi, const double* __restrict__ ttrig,
                                                   int ntheta, double* __restrict__ out) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)np * nphi * ntheta;
  if (k >= total) return;
  const int it = (int)(k % ntheta);
  const int ip = (int)((k / ntheta) % nphi);
  const int is = (int)(k / ((long long)ntheta * nphi));
  const double p = pshift[is];
  const double ct = ttrig[2 * it], st = ttrig[2 * it + 1];
  const double dtau = (2.0 / (nz - 1)) / (2.0);
  const double dz = 2.0 / (nz - 1);
  const double x0 = shift[0], inv_h = 1.0 / (shift[1] - shift[0]);
  const int half = 2 * nz;
  const long long rowlen = 2LL * ns;
  double acc = 0.0;
  for (int i = -half; i < half + 1; ++i) {
    const double tau = (double)i * dtau;
    const double zz = p * ct + tau * st;
    const double ss = p * st - tau * ct;
    double add;
    if ((fabs(tau * tau + p * p) > 1.0) || (fabs(zz) > 1.0) || (fabs(ss) > 1.0)) {
      add = 0.0;
    } else {
      int iz = (int)((zz + 1.0) / dz);
      if (iz == (nz - 1)) iz -= 1;
      const double za0 = (-1.0) + dz * iz;
      const double za1 = (-1.0) + dz * (iz + 1);
      const double za2 = (-1.0) + dz * (iz + 2);
      const int si = wrt_interval(shift, ns, x0, inv_h, ss);
      const double* r0 = tab + ((long long)iz * nphi + ip) * rowlen;
      const double* r1 = tab + ((long long)(iz + 1) * nphi + ip) * rowlen;
      const double v0 = tab_eval(shift, r0, si, ss);
      const double v1 = tab_eval(shift, r1, si, ss);
      double v2 = 0.0;
      if (!(iz >= (nz - 2))) {
        const double* r2 = tab + ((long long)(iz + 2) * nphi + ip) * rowlen;
        v2 = tab_eval(shift, r2, si, ss);
      }
      add = wrt_poly3(za0, za1, za2, v0, v1, v2, zz);
    }
    acc += (add * dtau);
  }
  out[k] = acc;
}

extern "C" int wrtk_reduction(const double* tab, int nz, const double* shift, int nshift,
                              const double* pshift, int np, int nphi,
                              const double* theta_trig, int ntheta, double* out,
                              hipStream_t st) {
  const long long total = (long long)np * nphi * ntheta;
  if (total <= 0) return WRT_OK;
  k_reduction<<<dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st>>>(
      tab, nz, shift, nshift, pshift, np, nphi, theta_trig, ntheta, out);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

// Ray rows at the planes' azimuths (the reduction on independent ray and plane grids):
// out[z][j][s] = sum_m w[j][m] * rays[z][s_m][k[j][m]] with s_m = ns-1-s where the stencil
// row is taken at the opposite azimuth (P(z, s, phi + pi) = P(z, -s, phi) on a shift grid
// symmetric about 0), else s.  stencil[j] = {w0..w3, k0..k3, flip0..flip3} as doubles; the
// four terms are summed in stencil order.
__global__ __launch_bounds__(256) void k_ray_phi_resample(const double* __restrict__ rays,
                                                          int nz, int ns, int nrphi,
                                                          const double* __restrict__ stencil,
                                                          int nphi, double* __restrict__ out) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)nz * nphi * ns;
  if (k >= total) return;
  const int is = (int)(k % ns);
  const int j = (int)((k / ns) % nphi);
  const long long iz = k / ((long long)ns * nphi);
  const double* sj = stencil + 12 * j;
  const double* zr = rays + iz * ns * (long long)nrphi;
  double acc = 0.0;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int kr = (int)sj[4 + m];
    const int sr = sj[8 + m] != 0.0 ? ns - 1 - is : is;
    acc += sj[m] * zr[(long long)sr * nrphi + kr];
  }
  out[k] = acc;
}

extern "C" int wrtk_ray_phi_resample(const double* rays, int nz, int nshift, int nrphi,
                                     const double* stencil, int nphi, double* out,
                                     hipStream_t st) {
  const long long total = (long long)nz * nphi * nshift;
  if (total <= 0) return WRT_OK;
  k_ray_phi_resample<<<dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st>>>(
      rays, nz, nshift, nrphi, stencil, nphi, out);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

// ----------------------------------------------------------- radon3d_adjoint (FP64) ---
// radon3d_adjoint (radon_inversion/backprojection/C/src/integral.c:6-24): one thread per
// voxel centre, directions phi-outer theta-inner in order,
// s = ((x*sin t)*cos p + (y*sin t)*sin p) + z*cos t, term ((2pi/nphi)*w_t)*S(s).
__global__ __launch_bounds__(256) void k_adjoint(const double* __restrict__ tab,
                                                 const double* __restrict__ shift, int ns,
                                                 const double* __restrict__ ptrig, int nphi,
                                                 const double* __restrict__ ttrig,
                                                 const double* __restrict__ wgt, int ntheta,
                                                 int n, double* __restrict__ out) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)n * n * n;
  if (k >= total) return;
  const int iy = (int)(k % n);
  const int ix = (int)((k / n) % n);
  const int iz = (int)(k / ((long long)n * n));
  const double x = (-1.0) + (1.0 / n) + ix * (2.0 / n);
  const double y = (-1.0) + (1.0 / n) + iy * (2.0 / n);
  const double z = (-1.0) + (1.0 / n) + iz * (2.0 / n);
  const double x0 = shift[0], inv_h = 1.0 / (shift[1] - shift[0]);
  const long long rowlen = 2LL * ns;
  double acc = 0.0;
  for (int ip = 0; ip < nphi; ++ip) {
    const double cp = ptrig[2 * ip], sp = ptrig[2 * ip + 1];
    const double* rows = tab + (long long)ip * ntheta * rowlen;
    for (int it = 0; it < ntheta; ++it) {
      const double ct = ttrig[2 * it], st = ttrig[2 * it + 1];
      const double s = x * st * cp + y * st * sp + z * ct;
      double v = 0.0;
      if (!(fabs(s) > 1.0)) {
        const int si = wrt_interval(shift, ns, x0, inv_h, s);
        v = tab_eval(shift, rows + (long long)it * rowlen, si, s);
      }
      acc += wgt[it] * v;
    }
  }
  out[k] = acc;
}

extern "C" int wrtk_adjoint(const double* tab, const double* shift, int nshift,
                            const double* phi_trig, int nphi, const double* theta_trig,
                            const double* wgt, int ntheta, int n, double* out, hipStream_t st) {
  const long long total = (long long)n * n * n;
  if (total <= 0) return WRT_OK;
  k_adjoint<<<dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st>>>(
      tab, shift, nshift, phi_trig, nphi, theta_trig, wgt, ntheta, n, out);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

// ------------------------------------------------------------ exact transposes (FP64) ---
// <R f, g> = sum_p g_p sum_samples f[cell] delta^2: scatter (g_p*delta)*delta into the cell
// of every in-cube sample, with the forward's exact index decisions.  FP64 atomics, so the
// summation order (not the value) varies run to run.
__global__ __launch_bounds__(256) void k_radon_T_plain(const double* __restrict__ g, int n,
                                                       const double* __restrict__ dirs, int ndir,
                                                       const double* __restrict__ shift,
                                                       long long nplanes,
                                                       double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const long long plane = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (plane >= nplanes) return;
  const double gv = g[plane];
  if (gv == 0.0) return;
  const int d = (int)(plane % ndir);
  const int is = (int)(plane / ndir);
  const double* D = dirs + 9 * (size_t)d;
  const double s = shift[is];
  const double Sx = s * D[0], Sy = s * D[1], Sz = s * D[2];
  const double ax = D[3], ay = D[4], az = D[5], bx = D[6], by = D[7], bz = D[8];
  const double delta = (1.0 / (n));
  const double dc = 2.0 / n;
  const double wgt = gv * delta * delta;
  const int half = 2 * n;
  for (int i1 = -half; i1 < half; ++i1) {
    const double f1 = (double)i1;
    const double rx = Sx + ax * f1, ry = Sy + ay * f1, rz = Sz + az * f1;
    for (int i2 = -half + lane; i2 < half; i2 += 64) {
      const double f2 = (double)i2;
      const double x = rx + bx * f2, y = ry + by * f2, z = rz + bz * f2;
      if ((fabs(x) > 1.0) || (fabs(y) > 1.0) || (fabs(z) > 1.0)) continue;
      int ix = (int)((x + 1.0) / dc), iy = (int)((y + 1.0) / dc), iz = (int)((z + 1.0) / dc);
      if (ix == n) ix -= 1;
      if (iy == n) iy -= 1;
      if (iz == n) iz -= 1;
      unsafeAtomicAdd(out + ((size_t)ix * n + iy) * n + iz, wgt);
    }
  }
}

extern "C" int wrtk_radon_T_plain(const double* g, int n, const double* dirs, int ndir,
                                  const double* shift, int nshift, double* out, hipStream_t st) {
  const long long nplanes = (long long)ndir * nshift;
  if (nplanes <= 0) return WRT_OK;
  if (((nplanes + 3) / 4) * 256 >= (1LL << 32)) return WRT_E_ARG;
  k_radon_T_plain<<<dim3((unsigned)((nplanes + 3) / 4)), dim3(256), 0, st>>>(
      g, n, dirs, ndir, shift, nplanes, out);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

__global__ __launch_bounds__(256) void k_ray_T_plain(const double* __restrict__ g, int n,
                                                     const double* __restrict__ zs, int nz,
                                                     const double* __restrict__ shift,
                                                     int nshift,
                                                     const double* __restrict__ trig, int nphi,
                                                     double* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)nz * nphi * nshift;
  if (t >= total) return;
  const int is = (int)(t % nshift);
  const int ip = (int)((t / nshift) % nphi);
  const int iz = (int)(t / ((long long)nshift * nphi));
  const double delta = (2.0 / (n - 1)) / (2.0);
  const double gv = g[((size_t)iz * nshift + is) * nphi + ip] * delta;
  if (gv == 0.0) return;
  const double c = trig[4 * ip], sn = trig[4 * ip + 1], dx = trig[4 * ip + 2],
               dy = trig[4 * ip + 3];
  const double s = shift[is];
  const double px = s * c, py = s * sn, pz = zs[iz];
  const double dg = 2.0 / (n - 1);
  const double sx = delta * dx, sy = delta * dy, sz = delta * 0.0;
  const int half = 2 * n;
  const size_t nn = (size_t)n * n;
  for (int i = -half; i < half + 1; ++i) {
    const double fi = (double)i;
    const double x = px + sx * fi, y = py + sy * fi, z = pz + sz * fi;
    if ((fabs(x) > 1.0) || (fabs(y) > 1.0) || (fabs(z) > 1.0)) continue;
    const double ax = x + 1.0, ay = y + 1.0, az = z + 1.0;
    int ix = (int)(ax / dg), iy = (int)(ay / dg), iz2 = (int)(az / dg);
    double cx = wrt_fmod_pos(ax, dg) / dg, cy = wrt_fmod_pos(ay, dg) / dg,
           cz = wrt_fmod_pos(az, dg) / dg;
    if (ix == n - 1) { ix -= 1; cx = 1.0; }
    if (iy == n - 1) { iy -= 1; cy = 1.0; }
    if (iz2 == n - 1) { iz2 -= 1; cz = 1.0; }
    const double wx[2] = {1.0 - cx, cx}, wy[2] = {1.0 - cy, cy}, wz[2] = {1.0 - cz, cz};
    double* p = out + (size_t)ix * nn + (size_t)iy * n + iz2;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int cc = 0; cc < 2; ++cc)
          unsafeAtomicAdd(p + a * nn + b * n + cc, gv * wx[a] * wy[b] * wz[cc]);
  }
}

extern "C" int wrtk_ray_T_plain(const double* g, int n, const double* z, int nz,
                                const double* shift, int nshift, const double* trig, int nphi,
                                double* out, hipStream_t st) {
  const long long total = (long long)nz * nphi * nshift;
  if (total <= 0) return WRT_OK;
  k_ray_T_plain<<<dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st>>>(
      g, n, z, nz, shift, nshift, trig, nphi, out);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

// ----------------------------------------------------------------------- utilities ---
__global__ void k_fill_f64(double* __restrict__ p, double v, size_t count) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_f64_to_f32(const double* __restrict__ s, float* __restrict__ d, size_t count) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    d[i] = (float)s[i];
}
__global__ void k_f32_to_f64(const float* __restrict__ s, double* __restrict__ d, size_t count) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    d[i] = (double)s[i];
}
static unsigned grid_for(size_t count) {
  size_t b = (count + 255) / 256;
  return (unsigned)(b > 65536 ? 65536 : (b ? b : 1));
}
extern "C" int wrtk_fill_f64(double* p, double v, size_t count, hipStream_t st) {
  k_fill_f64<<<grid_for(count), 256, 0, st>>>(p, v, count);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}
extern "C" int wrtk_f64_to_f32(const double* src, float* dst, size_t count, hipStream_t st) {
  k_f64_to_f32<<<grid_for(count), 256, 0, st>>>(src, dst, count);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}
extern "C" int wrtk_f32_to_f64(const float* src, double* dst, size_t count, hipStream_t st) {
  k_f32_to_f64<<<grid_for(count), 256, 0, st>>>(src, dst, count);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

// -------------------------------------------------------------- in-cube sample count ---
// Number of lattice samples of each plane that pass zero_interp's cube test, with the
// reference's arithmetic.  Along a lattice row every coordinate x(i2) = RN(R + RN(b*i2)) is
// monotone in i2, so each row's in-cube set is an interval; its ends are found by bisection
// on the exact predicate.  One thread per (plane, row).
__device__ __forceinline__ bool in_axis(double R, double b, int i2) {
  const double x = R + b * (double)i2;
  return !(fabs(x) > 1.0);
}

__global__ __launch_bounds__(256) void k_radon_incube(int n, const double* __restrict__ dirs,
                                                      int ndir, const double* __restrict__ shift,
                                                      long long nplanes,
                                                      unsigned long long* __restrict__ total) {
  const int rows = 4 * n;
  const long long work = nplanes * rows;
  unsigned long long cnt = 0;
  // grid-stride: a dispatch holds fewer than 2^32 work-items
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < work;
       t += (long long)gridDim.x * blockDim.x) {
    const long long plane = t / rows;
    const int i1 = (int)(t % rows) - 2 * n;
    const int d = (int)(plane % ndir), is = (int)(plane / ndir);
    const double* D = dirs + 9 * (size_t)d;
    const double s = shift[is];
    int lo = -2 * n, hi = 2 * n - 1;
    for (int c = 0; c < 3 && lo <= hi; ++c) {
      const double R = s * D[c] + D[3 + c] * (double)i1;
      const double b = D[6 + c];
      if (b == 0.0) {
        if (!in_axis(R, b, 0)) lo = hi + 1;
        continue;
      }
      // x(i2) is non-decreasing for b > 0, non-increasing for b < 0; "x <= 1" and "x >= -1"
      // are each monotone: bisect for the first and last index satisfying both
      auto ok_hi = [&](int i) { const double x = R + b * (double)i; return !(x > 1.0); };
      auto ok_lo = [&](int i) { const double x = R + b * (double)i; return !(x < -1.0); };
      int a0, a1;
      if (b > 0) {
        int l = lo, r = hi + 1;
        while (l < r) { const int m = l + ((r - l) >> 1); if (ok_lo(m)) r = m; else l = m + 1; }
        a0 = l;
        l = lo - 1; r = hi;
        while (l < r) { const int m = l + ((r - l + 1) >> 1); if (ok_hi(m)) l = m; else r = m - 1; }
        a1 = l;
      } else {
        int l = lo, r = hi + 1;
        while (l < r) { const int m = l + ((r - l) >> 1); if (ok_hi(m)) r = m; else l = m + 1; }
        a0 = l;
        l = lo - 1; r = hi;
        while (l < r) { const int m = l + ((r - l + 1) >> 1); if (ok_lo(m)) l = m; else r = m - 1; }
        a1 = l;
      }
      lo = max(lo, a0);
      hi = min(hi, a1);
    }
    if (hi >= lo) cnt += (unsigned long long)(hi - lo + 1);
  }
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off, 64);
  if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(total, cnt);
}

extern "C" int wrtk_radon_incube(int n, const double* dirs, int ndir, const double* shift,
                                 int nshift, unsigned long long* total, hipStream_t st) {
  const long long nplanes = (long long)ndir * nshift;
  const long long threads = nplanes * 4 * n;
  if (threads <= 0) return WRT_OK;
  long long blocks = (threads + 255) / 256;
  if (blocks > (1LL << 20)) blocks = 1LL << 20;
  k_radon_incube<<<dim3((unsigned)blocks), 256, 0, st>>>(n, dirs, ndir, shift, nplanes, total);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}
