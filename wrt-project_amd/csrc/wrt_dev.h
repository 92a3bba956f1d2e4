/*
 * wrt_dev.h -- internal launch interface between the C host layer (host/ sources) and the HIP
 * kernels (csrc/ sources).  Every pointer is a device pointer; every launcher enqueues on the
 * given stream, returns 0 or a negative WRT_E_* code, and never synchronises.
 *
 * Per-direction tables are computed on the host with the C library's sin/cos (the
 * reference calls libm per element; its bits decide index decisions), then uploaded:
 *   radon dirs   [ndir][9]  = n.x n.y n.z  a.x a.y a.z  b.x b.y b.z
 *                with a = delta*v1, b = delta*v2 (each product rounded once, as in
 *                radon_grid/src/integral.c:47-49), ndir = nphi*ntheta, index phi*ntheta+theta
 *   ray trig     [nphi][4]  = cos(phi) sin(phi) cos(phi+pi/2) sin(phi+pi/2)
 *   theta trig   [ntheta][2]= cos(theta) sin(theta)
 */
#ifndef WRT_DEV_H
#define WRT_DEV_H

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* GL rule (double-double Newton), ascending nodes */
int wrtk_gl(int n, double* x, double* w, hipStream_t st);

/* plain FP64 forms: the reference's arithmetic, one wave per plane (radon), one thread per
 * ray (ray), per plane (reduction) and per voxel (adjoint) */
int wrtk_radon_plain(const double* vol, int n, const double* dirs, int ndir,
                     const double* shift, int nshift, double* out, hipStream_t st);
int wrtk_ray_plain(const double* vol, int n, const double* z, int nz, const double* shift,
                   int nshift, const double* trig, int nphi, double* out, hipStream_t st);

/* natural cubic spline: factor the (shift-grid only) tridiagonal matrix once, then solve
 * every data row.  Row r = o*n_inner + i reads y[o*outer_stride + i + k*es], k < ns, and
 * writes the packed table tab[r][k][2] = (y_k, c_k) so that one interval is 32
 * contiguous bytes. */
int wrtk_spline_factor(const double* x, int ns, double* fac /* 2*ns */, hipStream_t st);
int wrtk_spline_rows(const double* x, int ns, const double* fac, const double* y, double* tab,
                     int n_outer, int n_inner, long long outer_stride, long long es,
                     hipStream_t st);

/* tab rows (z, phi): r = iz*nphi + iphi, splines over shift[nshift]; planes at the offsets
 * pshift[np] (the same array as shift in the reference's form) */
int wrtk_reduction(const double* tab, int nz, const double* shift, int nshift,
                   const double* pshift, int np, int nphi,
                   const double* theta_trig /* [ntheta][2] cos sin */, int ntheta, double* out,
                   hipStream_t st);
/* rays[z][s][k] (nrphi angles) -> out[z][j][s] at nphi plane azimuths; stencil[j][12] =
 * 4 weights, 4 ray-angle indices, 4 flags (1: row taken at -s) */
int wrtk_ray_phi_resample(const double* rays, int nz, int nshift, int nrphi,
                          const double* stencil, int nphi, double* out, hipStream_t st);
/* tab rows (phi, theta): r = iphi*ntheta + itheta;
 * wgt[theta] = (2pi/nphi) * w_theta, rounded once (integral.c:20) */
int wrtk_adjoint(const double* tab, const double* shift, int nshift,
                 const double* phi_trig /* [nphi][2] cos sin */, int nphi,
                 const double* theta_trig /* [ntheta][2] cos sin */, const double* wgt,
                 int ntheta, int n, double* out, hipStream_t st);

/* exact transposes, FP64 scatter */
int wrtk_radon_T_plain(const double* g, int n, const double* dirs, int ndir,
                       const double* shift, int nshift, double* out, hipStream_t st);
int wrtk_ray_T_plain(const double* g, int n, const double* z, int nz, const double* shift,
                     int nshift, const double* trig, int nphi, double* out, hipStream_t st);

/* FP32 backprojection: per-interval FP32 cubics from the FP64 packed spline table, then one
 * thread per voxel over float4 directions (n.x, n.y, n.z, weight); uniform shift grid.
 * trig [ndir][4] = sin(theta) cos(phi) sin(phi) cos(theta) in FP64 (cut decisions); order
 * the voxel tiles in launch order (wrt_adj_tile_order) */
int wrtk_adjoint_cubics(const double* x, int ns, const double* tab, void* cub, long long rows,
                        hipStream_t st);
int wrtk_adjoint_f32(const void* dirs, const void* trig, int ndir, const void* cub, int ns,
                     double s0, double ds, int n, const int* order, float* out, hipStream_t st);

/* FP32 rays: vol32 is the FP32 volume in [z][x][y] order; zi / zc the slices' index and
 * fraction as the reference decides them (zi < 0: |z| > 1); out [z][s][phi] FP32 */
int wrtk_vol_zxy_f32(const double* vol, int n, float* out, hipStream_t st);
int wrtk_ray_f32(const float* vol32, int n, const int* zi, const double* zc, int nz,
                 const double* shift, int nshift, const double* trig, int nphi, float* out,
                 hipStream_t st);
/* exact transpose of wrtk_ray_f32: g [z][s][phi] FP32 -> acc n^3 FP64 [x][y][z] (added to;
 * zeroed by the caller), with the forward's index and fraction decisions */
int wrtk_ray_T_f32(const float* g, int n, const int* zi, const double* zc, int nz,
                   const double* shift, int nshift, const double* trig, int nphi, double* acc,
                   hipStream_t st);

/* number of lattice samples passing zero_interp's cube test, summed over all planes */
int wrtk_radon_incube(int n, const double* dirs, int ndir, const double* shift, int nshift,
                      unsigned long long* total, hipStream_t st);

/* utilities */
int wrtk_fill_f64(double* p, double v, size_t count, hipStream_t st);
int wrtk_f64_to_f32(const double* src, float* dst, size_t count, hipStream_t st);
int wrtk_f32_to_f64(const float* src, double* dst, size_t count, hipStream_t st);

#ifdef __cplusplus
}
#endif
#endif
