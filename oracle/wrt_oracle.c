/*
 * wrt_oracle.c -- CPU FP64 restatement of the reference's 3D projection hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library.  The product path in
 * wrt-project_amd/ never calls it.
 *
 * Each function restates one reference function (cited file:line) with the same
 * operation order, so that in IEEE binary64 without fused multiply-add it yields the
 * reference's bits.  Where the reference calls GSL (absent from this image; an external,
 * unpinned ">= 2.5" dependency), the function restates GSL's documented algorithm:
 * glfixed = the n-point Gauss-Legendre rule in ascending node order; cspline = the natural
 * cubic spline; polynomial = Newton divided differences.  Those functions agree with GSL to
 * rounding (about 1e-15), not to the bit.
 *
 * Parity pins (tests/test_oracle_*.py): golden vectors from the reference's own GSL-free
 * sources built under oracle/_ref (plane_integral, zero_interp, ray_transform,
 * cube_trilinear_interp: bit-exact), numpy leggauss and the README theta values (GL),
 * scipy CubicSpline(bc_type='natural') (spline), exact quadratics (polynomial),
 * closed-form ellipsoid plane/line integrals and constant-data adjoint (4*pi*c), the
 * backprojection's orientation, sign and [z][x][y] order (data s * n_k gives (4 pi / 3) x_k
 * on each axis k), the reduction's z stencil (constant rays: the top cell's closed form with
 * its zeroed third slice; cubic-in-z rays: the sign of the forward stencil's remainder), each
 * with planted mistakes that the check catches, and the dot-product test for the
 * transposes.  The reduction on independent grids (no
 * reference counterpart) is pinned to the reference form on identical grids (bit for bit),
 * to reflected full-turn rays, to cubics in phi and to an off-centre Gaussian's closed-form
 * plane integrals, with planted stencil mistakes that fail the last check.
 */
#include "wrt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------------------ */
/* grids                                                                                */
/* ------------------------------------------------------------------------------------ */

/* phi_k = ((2*pi)/nphi)*k : radon_grid/src/init.c:167-170, ray_grid/src/init.c:173-176 */
void orc_phi_grid(int nphi, double* phi) {
  const double dphi = (2 * M_PI) / nphi;
  for (int k = 0; k < nphi; ++k) phi[k] = dphi * k;
}

/* s_j = -1 + j*(2/(nshift-1)) : radon_grid/src/init.c:173-176 */
void orc_shift_grid(int nshift, double* shift) {
  const double ds = 2.0 / (nshift - 1);
  for (int j = 0; j < nshift; ++j) shift[j] = (-1.0) + j * ds;
}

/* z_l = -1 + l*(2/(nslices-1)) : ray_grid/src/ray_grid.c:240-247 */
void orc_zslice_grid(int nslices, double* z) {
  const double dz = 2.0 / (nslices - 1);
  for (int l = 0; l < nslices; ++l) z[l] = (-1.0) + l * dz;
}

/* Double-double arithmetic for the Gauss-Legendre Newton iteration.  The rounded nodes and
 * weights are the correctly rounded values of the exact rule, which is what GSL's tabulated
 * rules hold (decimal literals of 25+ digits). */
typedef struct { double hi, lo; } ddv;

static ddv dd_qsum(double a, double b) { double s = a + b; ddv r = {s, b - (s - a)}; return r; }
static ddv dd_tsum(double a, double b) {
  double s = a + b, bb = s - a;
  ddv r = {s, (a - (s - bb)) + (b - bb)};
  return r;
}
static ddv dd_tprod(double a, double b) { double p = a * b; ddv r = {p, fma(a, b, -p)}; return r; }
static ddv dd_add(ddv a, ddv b) { ddv s = dd_tsum(a.hi, b.hi); return dd_qsum(s.hi, s.lo + (a.lo + b.lo)); }
static ddv dd_neg(ddv a) { ddv r = {-a.hi, -a.lo}; return r; }
static ddv dd_sub(ddv a, ddv b) { return dd_add(a, dd_neg(b)); }
static ddv dd_mul(ddv a, ddv b) {
  ddv p = dd_tprod(a.hi, b.hi);
  return dd_qsum(p.hi, p.lo + (a.hi * b.lo + a.lo * b.hi));
}
static ddv dd_muld(ddv a, double b) { ddv p = dd_tprod(a.hi, b); return dd_qsum(p.hi, p.lo + a.lo * b); }
static ddv dd_div(ddv a, ddv b) {
  double q1 = a.hi / b.hi;
  ddv r = dd_sub(a, dd_muld(b, q1));
  double q2 = r.hi / b.hi;
  r = dd_sub(r, dd_muld(b, q2));
  double q3 = r.hi / b.hi;
  ddv q = dd_qsum(q1, q2);
  ddv t = {q3, 0.0};
  return dd_add(q, t);
}
static ddv dd_d(double a) { ddv r = {a, 0.0}; return r; }

/* P_n(x) and P_{n-1}(x) by the three-term recurrence, double-double */
static void dd_legendre(int n, ddv x, ddv* pn, ddv* pnm1) {
  ddv p0 = dd_d(1.0), p1 = x;
  for (int k = 1; k < n; ++k) {
    ddv t = dd_muld(dd_mul(x, p1), (double)(2 * k + 1));
    ddv u = dd_muld(p0, (double)k);
    ddv p2 = dd_div(dd_sub(t, u), dd_d((double)(k + 1)));
    p0 = p1;
    p1 = p2;
  }
  *pn = p1;
  *pnm1 = p0;
}

static void d_legendre(int n, double x, double* pn, double* pnm1) {
  double p0 = 1.0, p1 = x;
  for (int k = 1; k < n; ++k) {
    double p2 = ((2 * k + 1) * x * p1 - k * p0) / (k + 1);
    p0 = p1;
    p1 = p2;
  }
  *pn = p1;
  *pnm1 = p0;
}

/* One positive root (k-th largest) and its weight, rounded from double-double. */
static void gl_root(int n, int k, double* xr, double* wr) {
  double x = cos(M_PI * (k + 0.75) / (n + 0.5));
  for (int it = 0; it < 100; ++it) {
    double pn, pm;
    d_legendre(n, x, &pn, &pm);
    double dp = n * (x * pn - pm) / (x * x - 1.0);
    double dx = pn / dp;
    x -= dx;
    if (fabs(dx) <= 1e-16) break;
  }
  ddv X = dd_d(x);
  for (int it = 0; it < 3; ++it) {
    ddv pn, pm;
    dd_legendre(n, X, &pn, &pm);
    ddv x2m1 = dd_sub(dd_mul(X, X), dd_d(1.0));
    ddv dp = dd_div(dd_muld(dd_sub(dd_mul(X, pn), pm), (double)n), x2m1);
    X = dd_sub(X, dd_div(pn, dp));
  }
  ddv pn, pm;
  dd_legendre(n, X, &pn, &pm);
  ddv npm = dd_muld(pm, (double)n);
  ddv num = dd_muld(dd_sub(dd_d(1.0), dd_mul(X, X)), 2.0);
  ddv w = dd_div(num, dd_mul(npm, npm));
  *xr = X.hi;
  *wr = w.hi;
}

/* gsl_integration_glfixed_point(-1,1,i,...) semantics: ascending nodes, exact mirror
 * symmetry x_i = -x_{n-1-i} (call sites radon_grid/src/init.c:180-187,
 * radon_inversion/backprojection/C/src/init.c:186-194). */
int orc_gl_nodes(int n, double* x, double* w) {
  if (n < 1) return -1;
  int m = n / 2;
  for (int k = 0; k < m; ++k) {
    double r, wk;
    gl_root(n, k, &r, &wk);
    x[n - 1 - k] = r;
    w[n - 1 - k] = wk;
    x[k] = -r;
    w[k] = wk;
  }
  if (n & 1) {
    ddv pn, pm;
    dd_legendre(n, dd_d(0.0), &pn, &pm);
    ddv npm = dd_muld(pm, (double)n);
    ddv wv = dd_div(dd_d(2.0), dd_mul(npm, npm));
    x[m] = 0.0;
    w[m] = wv.hi;
  }
  return 0;
}

/* theta_i = acos(x_i) : radon_grid/src/init.c:184 */
int orc_theta_grid(int ntheta, double* theta, double* w) {
  double* x = (double*)malloc(sizeof(double) * ntheta);
  double* ww = w ? w : (double*)malloc(sizeof(double) * ntheta);
  int rc = orc_gl_nodes(ntheta, x, ww);
  for (int i = 0; i < ntheta; ++i) theta[i] = acos(x[i]);
  free(x);
  if (!w) free(ww);
  return rc;
}

/* ------------------------------------------------------------------------------------ */
/* radon_grid                                                                           */
/* ------------------------------------------------------------------------------------ */

/* get_sph_direction + vec_product : radon_grid/src/integral.c:10-24, used :29-36 */
void orc_plane_basis(double phi, double theta, double nrm[3], double v1[3], double v2[3]) {
  nrm[0] = sin(theta) * cos(phi);
  nrm[1] = sin(theta) * sin(phi);
  nrm[2] = cos(theta);
  const double t2 = theta + M_PI / 2;
  v1[0] = sin(t2) * cos(phi);
  v1[1] = sin(t2) * sin(phi);
  v1[2] = cos(t2);
  v2[0] = nrm[1] * v1[2] - nrm[2] * v1[1];
  v2[1] = -(nrm[0] * v1[2] - nrm[2] * v1[0]);
  v2[2] = nrm[0] * v1[1] - nrm[1] * v1[0];
}

/* zero_interp : radon_grid/src/interp.c:63-88 (cell-centred, cube test, truncation) */
double orc_zero_interp(const double* vol, int n, double x, double y, double z) {
  if ((fabs(x) > 1) || (fabs(y) > 1) || (fabs(z) > 1)) return 0;
  const double dc = 2.0 / n;
  int ix = (int)((x + 1.0) / dc), iy = (int)((y + 1.0) / dc), iz = (int)((z + 1.0) / dc);
  if (ix == n) ix -= 1;
  if (iy == n) iy -= 1;
  if (iz == n) iz -= 1;
  return vol[((size_t)ix * n + iy) * n + iz];
}

/* plane_integral : radon_grid/src/integral.c:26-59.  4n x 4n lattice, i in [-2n, 2n),
 * point = shift*normal + (delta*v1)*i1 + (delta*v2)*i2, left to right. */
double orc_plane_integral(const double* vol, int n, double phi, double theta, double shift) {
  double nr[3], v1[3], v2[3];
  orc_plane_basis(phi, theta, nr, v1, v2);
  const double delta = (1.0 / (n));
  const int nsize = 4 * n;
  double integral = 0;
  for (int i1 = -(nsize / 2); i1 < (nsize / 2); ++i1) {
    for (int i2 = -(nsize / 2); i2 < (nsize / 2); ++i2) {
      const double x = shift * nr[0] + delta * v1[0] * i1 + delta * v2[0] * i2;
      const double y = shift * nr[1] + delta * v1[1] * i1 + delta * v2[1] * i2;
      const double z = shift * nr[2] + delta * v1[2] * i1 + delta * v2[2] * i2;
      const double value = orc_zero_interp(vol, n, x, y, z);
      integral += value * delta * delta;
    }
  }
  return integral;
}

long long orc_plane_incube(int n, double phi, double theta, double shift) {
  double nr[3], v1[3], v2[3];
  orc_plane_basis(phi, theta, nr, v1, v2);
  const double delta = (1.0 / (n));
  const int nsize = 4 * n;
  long long cnt = 0;
  for (int i1 = -(nsize / 2); i1 < (nsize / 2); ++i1)
    for (int i2 = -(nsize / 2); i2 < (nsize / 2); ++i2) {
      const double x = shift * nr[0] + delta * v1[0] * i1 + delta * v2[0] * i2;
      const double y = shift * nr[1] + delta * v1[1] * i1 + delta * v2[1] * i2;
      const double z = shift * nr[2] + delta * v1[2] * i1 + delta * v2[2] * i2;
      if (!((fabs(x) > 1) || (fabs(y) > 1) || (fabs(z) > 1))) ++cnt;
    }
  return cnt;
}

/* driver loop order s -> phi -> theta : radon_grid/src/radon_grid.c:248-258 */
void orc_radon_grid(const double* vol, int n, const double* phi, int nphi, const double* theta,
                    int ntheta, const double* shift, int nshift, double* out) {
  const long long total = (long long)nshift * nphi * ntheta;
#pragma omp parallel for schedule(dynamic, 16)
  for (long long k = 0; k < total; ++k) {
    const int it = (int)(k % ntheta);
    const int ip = (int)((k / ntheta) % nphi);
    const int is = (int)(k / ((long long)ntheta * nphi));
    out[k] = orc_plane_integral(vol, n, phi[ip], theta[it], shift[is]);
  }
}

void orc_radon_planes(const double* vol, int n, int count, const double* phi,
                      const double* theta, const double* shift, double* out) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int k = 0; k < count; ++k) out[k] = orc_plane_integral(vol, n, phi[k], theta[k], shift[k]);
}

/* ------------------------------------------------------------------------------------ */
/* ray_grid                                                                             */
/* ------------------------------------------------------------------------------------ */

/* cell_{linear,bilinear,trilinear}_interp : ray_grid/src/interp.c:3-21 */
static double lerp1(double t, double a, double b) { return ((1 - t) * a + t * b); }
static double lerp2(double x, double y, double v1, double v2, double v3, double v4) {
  double c1 = lerp1(x, v1, v2), c2 = lerp1(x, v4, v3);
  return lerp1(y, c1, c2);
}
static double lerp3(double x, double y, double z, double v1, double v2, double v3, double v4,
                    double v5, double v6, double v7, double v8) {
  double c1 = lerp2(x, y, v1, v2, v3, v4), c2 = lerp2(x, y, v5, v6, v7, v8);
  return lerp1(z, c1, c2);
}

/* cube_trilinear_interp : ray_grid/src/interp.c:25-60 (node-centred; the cell index is a
 * truncated quotient while the fraction is fmod/delta, so a quotient that rounds up to an
 * integer k reads with fraction ~1 from cell k, i.e. near node k+1). */
double orc_trilinear(const double* vol, int n, double x, double y, double z) {
  if ((fabs(x) > 1) || (fabs(y) > 1) || (fabs(z) > 1)) return 0;
  const double delta = 2.0 / (n - 1);
  int ix = (int)((x + 1.0) / delta), iy = (int)((y + 1.0) / delta), iz = (int)((z + 1.0) / delta);
  double cx = fmod((x + 1.0), delta) / delta, cy = fmod((y + 1.0), delta) / delta,
         cz = fmod((z + 1.0), delta) / delta;
  if (ix == (n - 1)) { ix -= 1; cx = 1.0; }
  if (iy == (n - 1)) { iy -= 1; cy = 1.0; }
  if (iz == (n - 1)) { iz -= 1; cz = 1.0; }
#define V(a, b, c) vol[((size_t)(a) * n + (b)) * n + (c)]
  return lerp3(cx, cy, cz, V(ix, iy, iz), V(ix + 1, iy, iz), V(ix + 1, iy + 1, iz),
               V(ix, iy + 1, iz), V(ix, iy, iz + 1), V(ix + 1, iy, iz + 1),
               V(ix + 1, iy + 1, iz + 1), V(ix, iy + 1, iz + 1));
#undef V
}

/* ray_transform : ray_grid/src/integral.c:11-50 (4n+1 samples, i in [-2n, 2n]) */
double orc_ray_transform(const double* vol, int n, double zslice, double shift, double phi) {
  const double nx = cos(phi), ny = sin(phi), nz = 0;
  const double dx = cos(phi + M_PI / 2), dy = sin(phi + M_PI / 2), dz = 0;
  const double px = shift * nx, py = shift * ny, pz = zslice;
  (void)nz;
  const double delta = (2.0 / (n - 1)) / (2.0);
  const int nsize = 4 * n;
  double integral = 0;
  for (int i = -(nsize / 2); i < (nsize / 2) + 1; ++i) {
    const double x = px + delta * dx * i;
    const double y = py + delta * dy * i;
    const double z = pz + delta * dz * i;
    integral += orc_trilinear(vol, n, x, y, z) * delta;
  }
  return integral;
}

/* driver loop order z -> s -> phi : ray_grid/src/ray_grid.c:242-255 */
void orc_ray_grid(const double* vol, int n, const double* z, int nz, const double* shift,
                  int nshift, const double* phi, int nphi, double* out) {
  const long long total = (long long)nz * nshift * nphi;
#pragma omp parallel for schedule(dynamic, 64)
  for (long long k = 0; k < total; ++k) {
    const int ip = (int)(k % nphi);
    const int is = (int)((k / nphi) % nshift);
    const int iz = (int)(k / ((long long)nphi * nshift));
    out[k] = orc_ray_transform(vol, n, z[iz], shift[is], phi[ip]);
  }
}

void orc_rays(const double* vol, int n, int count, const double* z, const double* shift,
              const double* phi, double* out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int k = 0; k < count; ++k) out[k] = orc_ray_transform(vol, n, z[k], shift[k], phi[k]);
}

/* ------------------------------------------------------------------------------------ */
/* natural cubic spline and 3-point polynomial                                          */
/* ------------------------------------------------------------------------------------ */

/* Natural cubic spline (GSL "cspline" semantics, second derivative zero at both ends).
 * c[i] is half the second derivative at node i.  Interior equations
 *   h_{i-1} c_{i-1} + 2 (h_{i-1} + h_i) c_i + h_i c_{i+1}
 *       = 3 ((y_{i+1} - y_i)/h_i - (y_i - y_{i-1})/h_{i-1})
 * are solved by the symmetric tridiagonal L D L^T factorisation.
 * Called per evaluation by the reference: radon_reduction/slicing.c:45-61,
 * radon_inversion/backprojection/C/src/interp.c:17-25. */
void orc_cspline_init(const double* x, const double* y, int n, double* c) {
  c[0] = 0.0;
  c[n - 1] = 0.0;
  const int m = n - 2;
  if (m <= 0) return;
  double* gam = (double*)malloc(sizeof(double) * (size_t)m);
  double* alp = (double*)malloc(sizeof(double) * (size_t)m);
  for (int k = 0; k < m; ++k) {
    const double hl = x[k + 1] - x[k], hr = x[k + 2] - x[k + 1];
    const double dg = 2.0 * (hl + hr);
    const double rhs = 3.0 * ((y[k + 2] - y[k + 1]) / hr - (y[k + 1] - y[k]) / hl);
    if (k == 0) {
      alp[0] = dg;
      c[1] = rhs;
    } else {
      const double off = hl; /* coupling of unknowns k-1 and k */
      alp[k] = dg - off * gam[k - 1];
      c[k + 1] = rhs - gam[k - 1] * c[k];
    }
    if (k < m - 1) gam[k] = hr / alp[k];
  }
  for (int k = 0; k < m; ++k) c[k + 1] = c[k + 1] / alp[k];
  for (int k = m - 2; k >= 0; --k) c[k + 1] = c[k + 1] - gam[k] * c[k + 2];
  free(gam);
  free(alp);
}

/* interval i with x[i] <= xv < x[i+1] (last interval for xv == x[n-1]); outside the
 * table the end interval is used (the reference's GSL call aborts there instead) */
static int spline_interval(const double* x, int n, double xv) {
  int lo = 0, hi = n - 1;
  while (hi > lo + 1) {
    int mid = (lo + hi) / 2;
    if (x[mid] > xv) hi = mid; else lo = mid;
  }
  return lo;
}

double orc_cspline_eval(const double* x, const double* y, const double* c, int n, double xv) {
  const int i = spline_interval(x, n, xv);
  const double h = x[i + 1] - x[i];
  const double t = xv - x[i];
  const double b = (y[i + 1] - y[i]) / h - h * (c[i + 1] + 2.0 * c[i]) / 3.0;
  const double d = (c[i + 1] - c[i]) / (3.0 * h);
  return y[i] + t * (b + t * (c[i] + t * d));
}

/* Newton divided-difference quadratic through three points (GSL "polynomial", size 3):
 * radon_reduction/slicing.c:67-74 */
double orc_poly3_eval(const double xa[3], const double ya[3], double x) {
  const double d1 = (ya[1] - ya[0]) / (xa[1] - xa[0]);
  const double d2a = (ya[2] - ya[1]) / (xa[2] - xa[1]);
  const double d2 = (d2a - d1) / (xa[2] - xa[0]);
  double r = d2;
  r = d1 + (x - xa[1]) * r;
  r = ya[0] + (x - xa[0]) * r;
  return r;
}

/* ------------------------------------------------------------------------------------ */
/* radon_reduction                                                                      */
/* ------------------------------------------------------------------------------------ */

/* rows[z][phi][s] and their spline coefficients, transposed from file order [z][s][phi]
 * as radon_reduction/init.c:134-143 does */
typedef struct {
  int nz, nshift, nphi;
  double* rows;
  double* coef;
} ray_table;

static void ray_table_build(ray_table* t, const double* rays, int nz, const double* shift,
                            int nshift, int nphi) {
  t->nz = nz;
  t->nshift = nshift;
  t->nphi = nphi;
  const size_t tot = (size_t)nz * nshift * nphi;
  t->rows = (double*)malloc(sizeof(double) * tot);
  t->coef = (double*)malloc(sizeof(double) * tot);
  for (int iz = 0; iz < nz; ++iz)
    for (int is = 0; is < nshift; ++is)
      for (int ip = 0; ip < nphi; ++ip)
        t->rows[((size_t)iz * nphi + ip) * nshift + is] = rays[((size_t)iz * nshift + is) * nphi + ip];
  for (size_t r = 0; r < (size_t)nz * nphi; ++r)
    orc_cspline_init(shift, t->rows + r * nshift, nshift, t->coef + r * nshift);
}

static void ray_table_free(ray_table* t) {
  free(t->rows);
  free(t->coef);
}

/* ray_interpolate : radon_reduction/slicing.c:5-77 */
static double ray_interpolate(const ray_table* t, const double* shift, double slice_z,
                              double slice_shift, int i_phi) {
  if ((fabs(slice_z) > 1.0) || (fabs(slice_shift) > 1.0)) return 0;
  const int ngrid = t->nz, nshift = t->nshift;
  const double delta_z = 2.0 / (ngrid - 1);
  int iz = (int)((slice_z + 1.0) / delta_z);
  if (iz == (ngrid - 1)) iz -= 1;
  double za[3], va[3];
  za[0] = (-1.0) + delta_z * iz;
  za[1] = (-1.0) + delta_z * (iz + 1);
  za[2] = (-1.0) + delta_z * (iz + 2);
  for (int j = 0; j < 3; ++j) {
    if (j == 2 && iz >= (ngrid - 2)) {
      va[2] = 0;
      break;
    }
    const size_t r = ((size_t)(iz + j) * t->nphi + i_phi) * nshift;
    va[j] = orc_cspline_eval(shift, t->rows + r, t->coef + r, nshift, slice_shift);
  }
  return orc_poly3_eval(za, va, slice_z);
}

/* radon_reduction : radon_reduction/slicing.c:80-119 */
static double radon_reduction_one(const ray_table* t, const double* shift, const double* phi,
                                  const double* theta, int i_shift, int i_phi, int i_theta) {
  (void)phi;
  const int ngrid = t->nz;
  const double dtau = (2.0 / (ngrid - 1)) / (2.0);
  const int nsize = 4 * ngrid;
  const double p_shift = shift[i_shift];
  const double p_theta = theta[i_theta];
  double integral = 0, add = 0;
  for (int i_tau = -(nsize / 2); i_tau < (nsize / 2) + 1; ++i_tau) {
    const double tau = (i_tau) * dtau;
    const double slice_z = p_shift * cos(p_theta) + tau * sin(p_theta);
    const double slice_shift = p_shift * sin(p_theta) - tau * cos(p_theta);
    if ((fabs(tau * tau + p_shift * p_shift) > 1.0) || (fabs(slice_z) > 1.0) ||
        (fabs(slice_shift) > 1.0)) {
      add = 0;
    } else {
      add = ray_interpolate(t, shift, slice_z, slice_shift, i_phi);
    }
    integral += (add * dtau);
  }
  return integral;
}

/* driver loop order s -> phi -> theta : radon_reduction/reduction.c:213-229 */
void orc_radon_reduction_grid(const double* rays, int nz, const double* shift, int nshift,
                              const double* phi, int nphi, const double* theta, int ntheta,
                              double* out) {
  ray_table t;
  ray_table_build(&t, rays, nz, shift, nshift, nphi);
  const long long total = (long long)nshift * nphi * ntheta;
#pragma omp parallel for schedule(dynamic, 16)
  for (long long k = 0; k < total; ++k) {
    const int it = (int)(k % ntheta);
    const int ip = (int)((k / ntheta) % nphi);
    const int is = (int)(k / ((long long)ntheta * nphi));
    out[k] = radon_reduction_one(&t, shift, phi, theta, is, ip, it);
  }
  ray_table_free(&t);
}

double orc_radon_reduction_one(const double* rays, int nz, const double* shift, int nshift,
                               const double* phi, int nphi, const double* theta, int ntheta,
                               int i_shift, int i_phi, int i_theta) {
  (void)ntheta;
  ray_table t;
  ray_table_build(&t, rays, nz, shift, nshift, nphi);
  double r = radon_reduction_one(&t, shift, phi, theta, i_shift, i_phi, i_theta);
  ray_table_free(&t);
  return r;
}

/* ------------------------------------------------------------------------------------ */
/* the reduction on independent ray and plane grids (no reference counterpart)          */
/* ------------------------------------------------------------------------------------ */

/* The ray row at azimuth phi from rays on nrphi angles k*period/nrphi (period pi or 2 pi):
 * P(z, s, phi + pi) = P(z, -s, phi) (the same line, sampled by the symmetric parameter set
 * of ray_grid/src/integral.c:11-50) maps phi into the period, and cubic Lagrange
 * interpolation over the four nearest angles, each row's spline read at s or -s, gives the
 * value.  An azimuth within 1e-12 of a node spacing of a ray angle takes that row alone. */
typedef struct {
  int k[4], flip[4];
  double w[4];
} phi_stencil;

static void phi_stencil_make(double phi, int nrphi, int half_turn, phi_stencil* ps) {
  const double period = half_turn ? M_PI : 2.0 * M_PI;
  double a = fmod(phi, 2.0 * M_PI);
  if (a < 0) a += 2.0 * M_PI;
  int flip = 0;
  if (half_turn && a >= M_PI) {
    a -= M_PI;
    flip = 1;
  }
  const double u = a / (period / nrphi);
  int k = (int)floor(u);
  double fr = u - k;
  if (fr < 1e-12) fr = 0.0;
  if (fr > 1.0 - 1e-12) {
    fr = 0.0;
    k += 1;
  }
  const double node[4] = {-1.0, 0.0, 1.0, 2.0};
  for (int m = 0; m < 4; ++m) {
    double w = 1.0;
    for (int i = 0; i < 4; ++i)
      if (i != m) w *= (fr - node[i]) / (node[m] - node[i]);
    ps->w[m] = w;
    /* index k-1+m on the unrolled grid; every period crossed is a half turn when
     * half_turn, i.e. one more s -> -s */
    int idx = k - 1 + m, f = flip;
    const int q = (int)floor((double)idx / nrphi);
    idx -= q * nrphi;
    if (half_turn && (q & 1)) f ^= 1;
    ps->k[m] = idx;
    ps->flip[m] = f;
  }
}

/* ray_interpolate (radon_reduction/slicing.c:5-77) with each z-row formed by the stencil */
static double ray_interpolate_stencil(const ray_table* t, const double* rshift, double slice_z,
                                      double slice_shift, const phi_stencil* ps) {
  if ((fabs(slice_z) > 1.0) || (fabs(slice_shift) > 1.0)) return 0;
  const int ngrid = t->nz, nshift = t->nshift;
  const double delta_z = 2.0 / (ngrid - 1);
  int iz = (int)((slice_z + 1.0) / delta_z);
  if (iz == (ngrid - 1)) iz -= 1;
  double za[3], va[3];
  za[0] = (-1.0) + delta_z * iz;
  za[1] = (-1.0) + delta_z * (iz + 1);
  za[2] = (-1.0) + delta_z * (iz + 2);
  for (int j = 0; j < 3; ++j) {
    if (j == 2 && iz >= (ngrid - 2)) {
      va[2] = 0;
      break;
    }
    double v = 0.0;
    for (int m = 0; m < 4; ++m) {
      const size_t r = ((size_t)(iz + j) * t->nphi + ps->k[m]) * nshift;
      const double sv = ps->flip[m] ? -slice_shift : slice_shift;
      v += ps->w[m] * orc_cspline_eval(rshift, t->rows + r, t->coef + r, nshift, sv);
    }
    va[j] = v;
  }
  return orc_poly3_eval(za, va, slice_z);
}

/* radon_reduction's tau loop (radon_reduction/slicing.c:80-119) at the plane (p, ., th) */
static double reduction_rays_one(const ray_table* t, const double* rshift, double p_shift,
                                 double p_theta, const phi_stencil* ps) {
  const int ngrid = t->nz;
  const double dtau = (2.0 / (ngrid - 1)) / (2.0);
  const int nsize = 4 * ngrid;
  double integral = 0, add = 0;
  for (int i_tau = -(nsize / 2); i_tau < (nsize / 2) + 1; ++i_tau) {
    const double tau = (i_tau) * dtau;
    const double slice_z = p_shift * cos(p_theta) + tau * sin(p_theta);
    const double slice_shift = p_shift * sin(p_theta) - tau * cos(p_theta);
    if ((fabs(tau * tau + p_shift * p_shift) > 1.0) || (fabs(slice_z) > 1.0) ||
        (fabs(slice_shift) > 1.0)) {
      add = 0;
    } else {
      add = ray_interpolate_stencil(t, rshift, slice_z, slice_shift, ps);
    }
    integral += (add * dtau);
  }
  return integral;
}

/* out[s][phi][theta] at the planes (shift, phi, theta) from rays[z][rs][k] */
void orc_radon_reduction_rays(const double* rays, int nz, const double* rshift, int nrshift,
                              int nrphi, int half_turn, const double* shift, int nshift,
                              const double* phi, int nphi, const double* theta, int ntheta,
                              double* out) {
  ray_table t;
  ray_table_build(&t, rays, nz, rshift, nrshift, nrphi);
  phi_stencil* ps = (phi_stencil*)malloc(sizeof(phi_stencil) * (size_t)nphi);
  for (int j = 0; j < nphi; ++j) phi_stencil_make(phi[j], nrphi, half_turn, ps + j);
  const long long total = (long long)nshift * nphi * ntheta;
#pragma omp parallel for schedule(dynamic, 16)
  for (long long k = 0; k < total; ++k) {
    const int it = (int)(k % ntheta);
    const int ip = (int)((k / ntheta) % nphi);
    const int is = (int)(k / ((long long)ntheta * nphi));
    out[k] = reduction_rays_one(&t, rshift, shift[is], theta[it], ps + ip);
  }
  free(ps);
  ray_table_free(&t);
}

/* one plane (i_shift, i_phi, i_theta) of orc_radon_reduction_rays */
double orc_radon_reduction_rays_one(const double* rays, int nz, const double* rshift,
                                    int nrshift, int nrphi, int half_turn, double p_shift,
                                    double p_phi, double p_theta) {
  ray_table t;
  ray_table_build(&t, rays, nz, rshift, nrshift, nrphi);
  phi_stencil ps;
  phi_stencil_make(p_phi, nrphi, half_turn, &ps);
  const double r = reduction_rays_one(&t, rshift, p_shift, p_theta, &ps);
  ray_table_free(&t);
  return r;
}

/* ------------------------------------------------------------------------------------ */
/* radon3d_adjoint                                                                      */
/* ------------------------------------------------------------------------------------ */

typedef struct {
  int nphi, ntheta, nshift;
  double* rows; /* [phi][theta][s] */
  double* coef;
} radon_table;

/* radon_values_init: file order [s][phi][theta] -> [phi][theta][s]
 * (radon_inversion/backprojection/C/src/init.c:130-136) */
static void radon_table_build(radon_table* t, const double* radon, int nphi, int ntheta,
                              const double* shift, int nshift) {
  t->nphi = nphi;
  t->ntheta = ntheta;
  t->nshift = nshift;
  const size_t tot = (size_t)nphi * ntheta * nshift;
  t->rows = (double*)malloc(sizeof(double) * tot);
  t->coef = (double*)malloc(sizeof(double) * tot);
  for (int is = 0; is < nshift; ++is)
    for (int ip = 0; ip < nphi; ++ip)
      for (int it = 0; it < ntheta; ++it)
        t->rows[((size_t)ip * ntheta + it) * nshift + is] = radon[((size_t)is * nphi + ip) * ntheta + it];
  for (size_t r = 0; r < (size_t)nphi * ntheta; ++r)
    orc_cspline_init(shift, t->rows + r * nshift, nshift, t->coef + r * nshift);
}

/* radon3d_adjoint + radon_interpolate_s :
 * radon_inversion/backprojection/C/src/integral.c:6-24, interp.c:5-28 */
static double adjoint_point(const radon_table* t, const double* phi, const double* theta,
                            const double* w, const double* shift, double x, double y, double z) {
  const int nphi = t->nphi, ntheta = t->ntheta, nshift = t->nshift;
  double acc = 0;
  const double dphi = 2 * M_PI / nphi;
  for (int ip = 0; ip < nphi; ++ip) {
    for (int it = 0; it < ntheta; ++it) {
      const double s = x * sin(theta[it]) * cos(phi[ip]) + y * sin(theta[it]) * sin(phi[ip]) +
                       z * cos(theta[it]);
      double v;
      if (fabs(s) > 1) {
        v = 0.0;
      } else {
        const size_t r = ((size_t)ip * ntheta + it) * nshift;
        v = orc_cspline_eval(shift, t->rows + r, t->coef + r, nshift, s);
      }
      acc += dphi * w[it] * v;
    }
  }
  return acc;
}

/* driver loop order z -> x -> y, voxel centres -1 + 1/n + i*2/n :
 * radon_inversion/backprojection/C/src/radon3d_ajoint.c:240-254 */
void orc_radon_adjoint_grid(const double* radon, const double* phi, int nphi,
                            const double* theta, const double* theta_w, int ntheta,
                            const double* shift, int nshift, int n, double* out) {
  radon_table t;
  radon_table_build(&t, radon, nphi, ntheta, shift, nshift);
  const long long total = (long long)n * n * n;
#pragma omp parallel for schedule(dynamic, 64)
  for (long long k = 0; k < total; ++k) {
    const int iy = (int)(k % n);
    const int ix = (int)((k / n) % n);
    const int iz = (int)(k / ((long long)n * n));
    const double x = (-1.0) + (1.0 / n) + ix * (2.0 / n), y = (-1.0) + (1.0 / n) + iy * (2.0 / n),
                 z = (-1.0) + (1.0 / n) + iz * (2.0 / n);
    out[k] = adjoint_point(&t, phi, theta, theta_w, shift, x, y, z);
  }
  free(t.rows);
  free(t.coef);
}

double orc_radon_adjoint_point(const double* radon, const double* phi, int nphi,
                               const double* theta, const double* theta_w, int ntheta,
                               const double* shift, int nshift, double x, double y, double z) {
  radon_table t;
  radon_table_build(&t, radon, nphi, ntheta, shift, nshift);
  double r = adjoint_point(&t, phi, theta, theta_w, shift, x, y, z);
  free(t.rows);
  free(t.coef);
  return r;
}

/* ------------------------------------------------------------------------------------ */
/* exact transposes (no reference counterpart)                                          */
/* ------------------------------------------------------------------------------------ */

/* <R f, g> = sum_planes g_p sum_samples f[cell]*delta*delta, so R^T g scatters
 * g_p*delta*delta into the cell of every in-cube sample of plane p. */
void orc_radon_grid_T(const double* g, int n, const double* phi, int nphi, const double* theta,
                      int ntheta, const double* shift, int nshift, double* out) {
  memset(out, 0, sizeof(double) * (size_t)n * n * n);
  const double delta = (1.0 / (n));
  const int nsize = 4 * n;
  const double dc = 2.0 / n;
  for (int is = 0; is < nshift; ++is)
    for (int ip = 0; ip < nphi; ++ip)
      for (int it = 0; it < ntheta; ++it) {
        const double gv = g[((size_t)is * nphi + ip) * ntheta + it];
        double nr[3], v1[3], v2[3];
        orc_plane_basis(phi[ip], theta[it], nr, v1, v2);
        const double shift_v = shift[is];
        const double wgt = gv * delta * delta;
        for (int i1 = -(nsize / 2); i1 < (nsize / 2); ++i1)
          for (int i2 = -(nsize / 2); i2 < (nsize / 2); ++i2) {
            const double x = shift_v * nr[0] + delta * v1[0] * i1 + delta * v2[0] * i2;
            const double y = shift_v * nr[1] + delta * v1[1] * i1 + delta * v2[1] * i2;
            const double z = shift_v * nr[2] + delta * v1[2] * i1 + delta * v2[2] * i2;
            if ((fabs(x) > 1) || (fabs(y) > 1) || (fabs(z) > 1)) continue;
            int ix = (int)((x + 1.0) / dc), iy = (int)((y + 1.0) / dc), iz = (int)((z + 1.0) / dc);
            if (ix == n) ix -= 1;
            if (iy == n) iy -= 1;
            if (iz == n) iz -= 1;
            out[((size_t)ix * n + iy) * n + iz] += wgt;
          }
      }
}

void orc_ray_grid_T(const double* g, int n, const double* zs, int nz, const double* shift,
                    int nshift, const double* phi, int nphi, double* out) {
  memset(out, 0, sizeof(double) * (size_t)n * n * n);
  const double delta = (2.0 / (n - 1)) / (2.0);
  const int nsize = 4 * n;
  const double dg = 2.0 / (n - 1);
  for (int l = 0; l < nz; ++l)
    for (int is = 0; is < nshift; ++is)
      for (int ip = 0; ip < nphi; ++ip) {
        const double gv = g[((size_t)l * nshift + is) * nphi + ip] * delta;
        const double nx = cos(phi[ip]), ny = sin(phi[ip]);
        const double dx = cos(phi[ip] + M_PI / 2), dy = sin(phi[ip] + M_PI / 2), dz = 0;
        const double px = shift[is] * nx, py = shift[is] * ny, pz = zs[l];
        for (int i = -(nsize / 2); i < (nsize / 2) + 1; ++i) {
          const double x = px + delta * dx * i, y = py + delta * dy * i, z = pz + delta * dz * i;
          if ((fabs(x) > 1) || (fabs(y) > 1) || (fabs(z) > 1)) continue;
          int ix = (int)((x + 1.0) / dg), iy = (int)((y + 1.0) / dg), iz = (int)((z + 1.0) / dg);
          double cx = fmod((x + 1.0), dg) / dg, cy = fmod((y + 1.0), dg) / dg,
                 cz = fmod((z + 1.0), dg) / dg;
          if (ix == (n - 1)) { ix -= 1; cx = 1.0; }
          if (iy == (n - 1)) { iy -= 1; cy = 1.0; }
          if (iz == (n - 1)) { iz -= 1; cz = 1.0; }
          const double wx[2] = {1 - cx, cx}, wy[2] = {1 - cy, cy}, wz[2] = {1 - cz, cz};
          for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
              for (int c = 0; c < 2; ++c)
                out[((size_t)(ix + a) * n + (iy + b)) * n + (iz + c)] += gv * wx[a] * wy[b] * wz[c];
        }
      }
}

/* ------------------------------------------------------------------------------------ */
/* closed forms                                                                         */
/* ------------------------------------------------------------------------------------ */

/* Solid ellipsoid {c + A u : |u| <= 1}, A = R diag(axes), density rho.  The plane
 * {x : x.n = p} cuts it in an ellipse of area pi (1 - t^2) det(A) / |A^T n|, with
 * t = (p - c.n) / |A^T n|. */
double orc_ellipsoid_plane(const double c[3], const double axes[3], const double rot[9],
                           double rho, const double nrm[3], double p) {
  double m[3];
  for (int j = 0; j < 3; ++j) {
    double s = 0;
    for (int i = 0; i < 3; ++i) s += rot[i * 3 + j] * nrm[i];
    m[j] = s * axes[j];
  }
  const double mn = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  const double t = (p - (c[0] * nrm[0] + c[1] * nrm[1] + c[2] * nrm[2])) / mn;
  if (fabs(t) >= 1) return 0;
  return rho * M_PI * (1 - t * t) * axes[0] * axes[1] * axes[2] / mn;
}

/* chord length of the line p0 + t dir (|dir| = 1) through the same ellipsoid, times rho */
double orc_ellipsoid_line(const double c[3], const double axes[3], const double rot[9],
                          double rho, const double p0[3], const double dir[3]) {
  double q[3], r[3];
  for (int j = 0; j < 3; ++j) {
    double sq = 0, sr = 0;
    for (int i = 0; i < 3; ++i) {
      sq += rot[i * 3 + j] * (p0[i] - c[i]);
      sr += rot[i * 3 + j] * dir[i];
    }
    q[j] = sq / axes[j];
    r[j] = sr / axes[j];
  }
  const double a = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  const double b = q[0] * r[0] + q[1] * r[1] + q[2] * r[2];
  const double cc = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] - 1;
  const double disc = b * b - a * cc;
  if (disc <= 0) return 0;
  return rho * 2 * sqrt(disc) / a;
}

/* ------------------------------------------------------------------------------------ */
/* analytic programs                                                                    */
/* ------------------------------------------------------------------------------------ */

/* The example functions square with pow(., 2).  The reference builds them at -O0, where that
 * is a call of the C library's pow, which is not correctly rounded everywhere; at -O2 gcc
 * would fold it to x*x.  Calling through a volatile pointer keeps the library call. */
static double (*volatile lib_pow)(double, double) = pow;
#define pow2(v) lib_pow((v), 2)

/* exponential_bump.c:5-16 with the centre and the radius terms as parameters */
static double fn_exp_bump(const orc_fn_term* t, double x, double y, double z) {
  const double xx = x - t->c[0], yy = y - t->c[1], zz = z - t->c[2];
  double value;
  if ((pow2(xx) + pow2(yy) + pow2(zz)) < (t->p[0] - t->p[1])) {
    value = exp((-1) / (t->p[0] - pow2(xx) - pow2(yy) - pow2(zz)));
  } else {
    value = 0;
  }
  return value;
}

/* sine_bump_template : sinusoidal_bump.c:3-19 */
static double fn_sine_bump(const orc_fn_term* t, double x, double y, double z) {
  const double amplitude = t->p[0], rin = t->p[1], rout = t->p[2];
  double r = sqrt(pow2(x - t->c[0]) + pow2(y - t->c[1]) + pow2(z - t->c[2]));
  if (r > rout) return 0;
  if ((r < rout) && (r > rin)) {
    const double tt = M_PI * (r - rout) / (rout - rin);
    return ((amplitude / (20 * M_PI)) * (-sin(6 * tt) / 3 + 3 * sin(4 * tt) - 15 * sin(2 * tt) + 20 * tt));
  }
  if (r < rin) return amplitude;
  return 0;
}

/* the phantom's solid ellipsoid (build-owned): inside when |diag(1/a) R^T (x - c)| <= 1 */
static double fn_ellipsoid(const orc_fn_term* t, double x, double y, double z) {
  const double dx = x - t->c[0], dy = y - t->c[1], dz = z - t->c[2];
  const double* R = t->p + 4;
  const double u = (R[0] * dx + R[3] * dy + R[6] * dz) / t->p[1];
  const double w = (R[1] * dx + R[4] * dy + R[7] * dz) / t->p[2];
  const double v = (R[2] * dx + R[5] * dy + R[8] * dz) / t->p[3];
  return (u * u + w * w + v * v <= 1.0) ? t->p[0] : 0.0;
}

/* the terms summed left to right, as sinusoidal_bump.c:32-34 sums its three bumps */
double orc_fn_eval(const orc_fn_term* t, int nt, double x, double y, double z) {
  double v = 0;
  for (int k = 0; k < nt; ++k) {
    double f = 0;
    switch (t[k].kind) {
      case 0: f = fn_exp_bump(t + k, x, y, z); break;
      case 1: f = fn_sine_bump(t + k, x, y, z); break;
      case 2: f = fn_ellipsoid(t + k, x, y, z); break;
    }
    v = k == 0 ? f : v + f;
  }
  return v;
}

/* plane_integral : radon_analytic/src/integral.c:25-56 (basis as radon_grid's, :9-22) */
double orc_analytic_plane(const orc_fn_term* t, int nt, int ngrid, double phi, double theta,
                          double shift) {
  double nr[3], v1[3], v2[3];
  orc_plane_basis(phi, theta, nr, v1, v2);
  const double delta = (2.0 / (ngrid - 1)) / 2.0;
  const int nsize = 4 * ngrid;
  double integral = 0;
  for (int i1 = -(nsize / 2); i1 < (nsize / 2); ++i1) {
    for (int i2 = -(nsize / 2); i2 < (nsize / 2); ++i2) {
      const double x = shift * nr[0] + delta * v1[0] * i1 + delta * v2[0] * i2;
      const double y = shift * nr[1] + delta * v1[1] * i1 + delta * v2[1] * i2;
      const double z = shift * nr[2] + delta * v1[2] * i1 + delta * v2[2] * i2;
      const double value = orc_fn_eval(t, nt, x, y, z);
      integral += value * delta * delta;
    }
  }
  return integral;
}

/* ray_integral : ray_analytic/integral.c:10-47 */
double orc_analytic_ray(const orc_fn_term* t, int nt, double zslice, double shift, double phi,
                        int ngrid) {
  const double nx = cos(phi), ny = sin(phi);
  const double dx = cos(phi + M_PI / 2), dy = sin(phi + M_PI / 2), dz = 0;
  const double px = shift * nx, py = shift * ny, pz = zslice;
  const double delta = (2.0 / (ngrid - 1)) / 2.0;
  const int nsize = 4 * ngrid;
  double integral = 0;
  for (int i = -(nsize / 2); i < (nsize / 2) + 1; ++i) {
    const double x = px + delta * dx * i;
    const double y = py + delta * dy * i;
    const double z = pz + delta * dz * i;
    const double value = orc_fn_eval(t, nt, x, y, z);
    integral += value * delta;
  }
  return integral;
}

void orc_analytic_radon_grid(const orc_fn_term* t, int nt, int ngrid, const double* phi, int nphi,
                             const double* theta, int ntheta, const double* shift, int nshift,
                             double* out) {
  const long long total = (long long)nshift * nphi * ntheta;
#pragma omp parallel for schedule(dynamic, 4)
  for (long long k = 0; k < total; ++k) {
    const int it = (int)(k % ntheta), ip = (int)((k / ntheta) % nphi), is = (int)(k / ((long long)ntheta * nphi));
    out[k] = orc_analytic_plane(t, nt, ngrid, phi[ip], theta[it], shift[is]);
  }
}

void orc_analytic_ray_grid(const orc_fn_term* t, int nt, int ngrid, const double* z, int nz,
                           const double* shift, int nshift, const double* phi, int nphi,
                           double* out) {
  const long long total = (long long)nz * nshift * nphi;
#pragma omp parallel for schedule(dynamic, 16)
  for (long long k = 0; k < total; ++k) {
    const int ip = (int)(k % nphi), is = (int)((k / nphi) % nshift), iz = (int)(k / ((long long)nphi * nshift));
    out[k] = orc_analytic_ray(t, nt, z[iz], shift[is], phi[ip], ngrid);
  }
}

void orc_analytic_zslices(int ngrid, double* z) {
  for (int i = 0; i < ngrid; ++i) z[i] = (-1.0) + i * 2.0 / (ngrid - 1);
}

/* ------------------------------------------------------------------------------------ */
/* Fourier-slice inversion                                                              */
/* ------------------------------------------------------------------------------------ */

/* the adjoint NFFT of nfft_reconstruct_3d.m:19-53 as its exact sum */
void orc_fourier_adjoint(const double* coef, const double* rho, int n, const double* dirs,
                         int ndir, int ngrid, double* out) {
  const double dx = 2.0 / ngrid;
  const long long total = (long long)ngrid * ngrid * ngrid;
#pragma omp parallel for schedule(dynamic, 64)
  for (long long v = 0; v < total; ++v) {
    const int iz = (int)(v % ngrid), iy = (int)((v / ngrid) % ngrid), ix = (int)(v / ((long long)ngrid * ngrid));
    const double px = (ix - ngrid / 2) * dx, py = (iy - ngrid / 2) * dx, pz = (iz - ngrid / 2) * dx;
    double acc = 0.0;
    for (int d = 0; d < ndir; ++d) {
      const double t = px * dirs[3 * d] + py * dirs[3 * d + 1] + pz * dirs[3 * d + 2];
      const double* c = coef + 2 * (size_t)d * n;
      for (int m = 0; m < n; ++m) {
        const double a = 2 * M_PI * rho[m] * t;
        acc += c[2 * m] * cos(a) - c[2 * m + 1] * sin(a);
      }
    }
    out[v] = acc;
  }
}

/* ------------------------------------------------------------------------------------ */
/* arithmetic property used by the fast kernels                                          */
/* ------------------------------------------------------------------------------------ */

/* Number of u in [0, 2] (random, and within a few ulps of every multiple k*d) for which the
 * corrected reciprocal q = fma(fma(-y, d, u), inv, y), y = u*inv, inv = RN(1/d), differs
 * from the IEEE quotient u/d, for the cell width d = RN(2/n) of zero_interp.  u ranges over
 * the values x + 1 can take for a double x >= -1: zero or at least 2^-53 (subnormal u, where
 * the reciprocal form loses its guarantee, cannot occur). */
long long orc_div_rn_mismatches(int n, long long count, unsigned long long seed) {
  const double d = 2.0 / n, inv = 1.0 / d;
  unsigned long long s = seed ? seed : 1;
  long long bad = 0;
  for (long long i = 0; i < count; ++i) {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    double u = (double)(s >> 11) * (2.0 / 9007199254740992.0);
    if (u != 0.0 && u < 0x1p-53) continue;
    const double y = u * inv, r = fma(-y, d, u), q = fma(r, inv, y);
    bad += q != u / d;
  }
  for (int k = 0; k <= n; ++k) {
    double u = k * d;
    for (int t = 0; t < 8; ++t) u = nextafter(u, -1.0);
    for (int t = 0; t < 17; ++t, u = nextafter(u, 3.0)) {
      if (u < 0 || u > 2.0 || (u != 0.0 && u < 0x1p-53)) continue;
      const double y = u * inv, r = fma(-y, d, u), q = fma(r, inv, y);
      bad += q != u / d;
    }
  }
  return bad;
}
