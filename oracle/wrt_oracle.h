/*
 * wrt_oracle.h -- CPU FP64 restatement of the reference's 3D projection hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP kernels in
 * wrt-project_amd/.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it, and only to check or to time the CPU baseline; the product path
 * never calls it and fails loudly when its own HIP library is missing.
 *
 * Layouts are flat row-major arrays with the reference's index order:
 *   volume  vol[x][y][z]   (radon_grid/src/init.c:120-124, ray_grid/src/init.c:118-122)
 *   radon   out[s][phi][theta]  (radon_grid/src/radon_grid.c:248-258)
 *   rays    out[z][s][phi]      (ray_grid/src/ray_grid.c:244-255)
 *   adjoint out[z][x][y]        (radon_inversion/backprojection/C/src/radon3d_ajoint.c:240-254)
 *
 * Arithmetic is IEEE binary64 with the reference's operation order; build with
 * -ffp-contract=off (and no -mfma) so that no multiply-add is fused.
 */
#ifndef WRT_ORACLE_H
#define WRT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- grids (R1, R2) ---- */
void orc_phi_grid(int nphi, double* phi);
void orc_shift_grid(int nshift, double* shift);
void orc_zslice_grid(int nslices, double* z);
int  orc_gl_nodes(int n, double* x, double* w);            /* ascending nodes on [-1,1] */
int  orc_theta_grid(int ntheta, double* theta, double* w);  /* theta_i = acos(x_i) */

/* ---- radon_grid (R3, R4, R5) ---- */
void   orc_plane_basis(double phi, double theta, double normal[3], double v1[3], double v2[3]);
double orc_zero_interp(const double* vol, int n, double x, double y, double z);
double orc_plane_integral(const double* vol, int n, double phi, double theta, double shift);
void   orc_radon_grid(const double* vol, int n, const double* phi, int nphi,
                      const double* theta, int ntheta, const double* shift, int nshift,
                      double* out);
void   orc_radon_planes(const double* vol, int n, int count, const double* phi,
                        const double* theta, const double* shift, double* out);
/* number of samples of one plane that pass the cube test of zero_interp */
long long orc_plane_incube(int n, double phi, double theta, double shift);

/* ---- ray_grid (R6, R7) ---- */
double orc_trilinear(const double* vol, int n, double x, double y, double z);
double orc_ray_transform(const double* vol, int n, double zslice, double shift, double phi);
void   orc_ray_grid(const double* vol, int n, const double* z, int nz, const double* shift,
                    int nshift, const double* phi, int nphi, double* out);
void   orc_rays(const double* vol, int n, int count, const double* z, const double* shift,
                const double* phi, double* out);

/* ---- natural cubic spline and 3-point polynomial (GSL cspline / polynomial semantics) ---- */
void   orc_cspline_init(const double* x, const double* y, int n, double* c);
double orc_cspline_eval(const double* x, const double* y, const double* c, int n, double xv);
double orc_poly3_eval(const double xa[3], const double ya[3], double x);

/* ---- radon_reduction (R8, R9) ---- */
/* rays in file order [nz][nshift][nphi]; out[s][phi][theta] */
void   orc_radon_reduction_grid(const double* rays, int nz, const double* shift, int nshift,
                                const double* phi, int nphi, const double* theta, int ntheta,
                                double* out);
double orc_radon_reduction_one(const double* rays, int nz, const double* shift, int nshift,
                               const double* phi, int nphi, const double* theta, int ntheta,
                               int i_shift, int i_phi, int i_theta);
/* the same shift integral on independent grids (no reference counterpart): rays[z][rs][k]
 * on nrphi angles k*pi/nrphi (half_turn) or k*2pi/nrphi, rows at each plane azimuth by
 * P(z,s,phi+pi) = P(z,-s,phi) and cubic Lagrange interpolation in phi */
void   orc_radon_reduction_rays(const double* rays, int nz, const double* rshift, int nrshift,
                                int nrphi, int half_turn, const double* shift, int nshift,
                                const double* phi, int nphi, const double* theta, int ntheta,
                                double* out);
double orc_radon_reduction_rays_one(const double* rays, int nz, const double* rshift,
                                    int nrshift, int nrphi, int half_turn, double p_shift,
                                    double p_phi, double p_theta);

/* ---- radon3d_adjoint (R10, R11) ---- */
/* radon in file order [nshift][nphi][ntheta]; out[z][x][y] with n^3 voxel centres */
void   orc_radon_adjoint_grid(const double* radon, const double* phi, int nphi,
                              const double* theta, const double* theta_w, int ntheta,
                              const double* shift, int nshift, int n, double* out);
double orc_radon_adjoint_point(const double* radon, const double* phi, int nphi,
                               const double* theta, const double* theta_w, int ntheta,
                               const double* shift, int nshift, double x, double y, double z);

/* ---- exact transposes (R15, no reference counterpart) ---- */
void orc_radon_grid_T(const double* g, int n, const double* phi, int nphi,
                      const double* theta, int ntheta, const double* shift, int nshift,
                      double* out /* vol[x][y][z] */);
void orc_ray_grid_T(const double* g, int n, const double* z, int nz, const double* shift,
                    int nshift, const double* phi, int nphi, double* out /* vol[x][y][z] */);

/* ---- closed forms used to pin the restatement ---- */
/* plane integral of the solid ellipsoid {x : |M^T (x-c)|^2 <= 1}, M = R diag(1/a) */
double orc_ellipsoid_plane(const double c[3], const double axes[3], const double rot[9],
                           double rho, const double nrm[3], double p);
double orc_ellipsoid_line(const double c[3], const double axes[3], const double rot[9],
                          double rho, const double p0[3], const double dir[3]);

/* ---- analytic programs (radon_analytic, ray_analytic) ---- */
/* A test function as a sum of terms; the layout equals the product's wrt_fn_term.
 *   kind 0, exponential bump  (radon_analytic/examples-test-functions/exponential_bump.c:5-16):
 *           p[0] = r2max, p[1] = eps: exp(-1/(r2max - |x-c|^2)) where |x-c|^2 < r2max - eps
 *   kind 1, sine bump         (…/sinusoidal_bump.c:3-19 sine_bump_template):
 *           p[0] = amplitude, p[1] = rin, p[2] = rout
 *   kind 2, solid ellipsoid   (build-owned, the phantom's ellipsoid):
 *           p[0] = density, p[1..3] = semi-axes, p[4..12] = rotation, row-major */
typedef struct {
  int kind;
  double c[3];
  double p[13];
} orc_fn_term;
double orc_fn_eval(const orc_fn_term* t, int nt, double x, double y, double z);
/* plane_integral (radon_analytic/src/integral.c:25-56): delta = (2/(ngrid-1))/2, lattice
 * i in [-2 ngrid, 2 ngrid), every sample evaluated, serial sum */
double orc_analytic_plane(const orc_fn_term* t, int nt, int ngrid, double phi, double theta,
                          double shift);
/* ray_integral (ray_analytic/integral.c:10-47): i in [-2 ngrid, 2 ngrid] */
double orc_analytic_ray(const orc_fn_term* t, int nt, double zslice, double shift, double phi,
                        int ngrid);
/* whole grids, out[s][phi][theta] and out[z][s][phi] */
void orc_analytic_radon_grid(const orc_fn_term* t, int nt, int ngrid, const double* phi, int nphi,
                             const double* theta, int ntheta, const double* shift, int nshift,
                             double* out);
void orc_analytic_ray_grid(const orc_fn_term* t, int nt, int ngrid, const double* z, int nz,
                           const double* shift, int nshift, const double* phi, int nphi,
                           double* out);
/* z_i = (-1.0) + i*2.0/(ngrid-1) (ray_analytic/ray_analytic.c:255) */
void orc_analytic_zslices(int ngrid, double* z);

/* ---- Fourier-slice inversion (radon_inversion/nfft_reconstruct_3d.m:19-53) ---- */
/* out[ix][iy][iz] = Re sum_d sum_m coef[d][m] exp(2 pi i rho_m (p . u_d)) at
 * p = ((ix, iy, iz) - ngrid/2) * 2/ngrid: the adjoint NFFT's sum, evaluated exactly with one
 * cosine and sine per term.  coef [ndir][n] complex (re, im), dirs [ndir][3]. */
void orc_fourier_adjoint(const double* coef, const double* rho, int n, const double* dirs,
                         int ndir, int ngrid, double* out);

/* ---- arithmetic property of the fast kernels' division ---- */
long long orc_div_rn_mismatches(int n, long long count, unsigned long long seed);

#ifdef __cplusplus
}
#endif
#endif
