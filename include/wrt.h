/*
 * wrt.h -- C-ABI of the MI355X-native 3D projection hot path.
 *
 * Drop-in for the numerical core of the reference programs radon_grid, ray_grid,
 * radon_reduction and radon3d_ajoint (devhliu/wrt-project).  The reference computes one
 * output element per call of a pure C function over a jagged double*** array; this ABI
 * takes whole grids as flat row-major arrays (same index order as the reference's files)
 * and runs them as HIP kernels on one AMD Instinct MI355X (gfx950).
 *
 * Layouts (x-major volumes, z fastest, as read by tfunction_values_init):
 *   volume   vol[x][y][z]            n^3       radon_grid/src/init.c:120-124
 *   radon    out[s][phi][theta]                radon_grid/src/radon_grid.c:248-258
 *   rays     out[z][s][phi]                    ray_grid/src/ray_grid.c:244-255
 *   adjoint  out[z][x][y]                      radon_inversion/backprojection/C/src/radon3d_ajoint.c:240-254
 *
 * Conventions:
 *   - every function returns WRT_OK (0) or a negative WRT_E_* code, never aborts;
 *   - the caller owns every host buffer; a wrt_ctx owns all device memory it allocates;
 *   - one submitting host thread per wrt_ctx;
 *   - precision WRT_F64 is the plain form: the reference's arithmetic and index decisions
 *     in IEEE binary64 (relative error vs the reference <= 1e-10);
 *     WRT_F32 / WRT_BF16 are the fast forms (FP32 or BF16 values, FP32 compensated sums);
 *     their error against WRT_F64 is reported, not assumed.
 */
#ifndef WRT_H
#define WRT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ error codes --- */
#define WRT_OK 0
#define WRT_E_ARG (-1)     /* invalid size, grid or pointer                           */
#define WRT_E_NODEV (-2)   /* no usable HIP device                                    */
#define WRT_E_HIP (-3)     /* HIP runtime error                                       */
#define WRT_E_NOMEM (-4)   /* device or host allocation failed                        */
#define WRT_E_ADMIT (-5)   /* serve loop: the job does not fit the free-HBM budget    */
#define WRT_E_IO (-6)      /* file open / parse error (config, volume, data files)    */
#define WRT_E_STATE (-7)   /* call out of order (e.g. run before load)                */
#define WRT_E_SCRATCH (-8) /* wrt_radon_grid_noalloc: the call needs more device memory
                              than the library's fixed scratch holds                  */

typedef enum { WRT_F64 = 0, WRT_F32 = 1, WRT_BF16 = 2 } wrt_prec;

typedef enum {
  WRT_OP_RADON = 0,     /* plane integrals       (radon_grid)                      */
  WRT_OP_RAY = 1,       /* layer-by-layer rays   (ray_grid)                        */
  WRT_OP_REDUCTION = 2, /* rays -> Radon         (radon_reduction)                 */
  WRT_OP_ADJOINT = 3,   /* backprojection        (radon3d_ajoint)                  */
  WRT_OP_RADON_T = 4,   /* exact transpose of WRT_OP_RADON (no reference counterpart) */
  WRT_OP_RAY_T = 5      /* exact transpose of WRT_OP_RAY   (no reference counterpart) */
} wrt_op;

typedef struct wrt_ctx wrt_ctx;

/* Timing and work accounting of one call.  seconds: device time from HIP events on the
 * context stream around the whole device part of the call (no host transfers);
 * kernel_seconds: the dominant kernel only.  logical_samples counts the reference's
 * samples (plane: 16 n^2 per plane, ray: 4n+1 per ray, reduction: 4n+1 per plane,
 * adjoint: nphi*ntheta per voxel); alg_bytes is the algorithmic byte count of the
 * dominant kernel. */
typedef struct wrt_stats {
  double seconds;
  double kernel_seconds;
  double logical_samples;
  double alg_bytes;
  long long outputs;
} wrt_stats;

/* -------------------------------------------------------------------- context --- */
int wrt_init(wrt_ctx** ctx, int device);
void wrt_free(wrt_ctx* ctx);
const char* wrt_strerror(int code);
/* free / total device memory as reported by the HIP runtime (admission input) */
int wrt_mem_info(wrt_ctx* ctx, size_t* free_bytes, size_t* total_bytes);

/* ---------------------------------------------------------------------- grids --- */
/* n-point Gauss-Legendre rule on [-1,1], ascending nodes, computed on the GPU in
 * double-double and rounded.  Replaces gsl_integration_glfixed_table_alloc/point at
 * radon_grid/src/init.c:180-187, radon_reduction/init.c:213-220,
 * radon_inversion/backprojection/C/src/init.c:186-194. */
int wrt_gl_nodes(wrt_ctx* ctx, int n, double* x, double* w);

/* phi_k = (2pi/nphi)k, theta_i = acos(x_i), theta_w[i] = w_i (may be NULL),
 * s_j = -1 + j*2/(nshift-1).  Replaces radon_grid_init (radon_grid/src/init.c:165-190)
 * and its weighted copy (radon_inversion/backprojection/C/src/init.c:170-197). */
int wrt_radon_grid_init(wrt_ctx* ctx, int nphi, int ntheta, int nshift, double* phi,
                        double* theta, double* theta_w, double* shift);
/* s_j as above, phi_k = (2pi/nphi)k.  Replaces ray_grid_init (ray_grid/src/init.c:163-177). */
int wrt_ray_grid_init(int nshift, int nphi, double* shift, double* phi);
/* z_l = -1 + l*2/(nslices-1) (ray_grid/src/ray_grid.c:240-247). */
int wrt_zslice_grid(int nslices, double* z);

/* ------------------------------------------------- host-buffer entry points --- */
/* Plane integrals over every (shift, phi, theta): replaces plane_integral
 * (radon_grid/src/integral.c:26-59) called from radon_grid/src/radon_grid.c:248-258.
 * out[s][phi][theta].  The FP32 and BF16 forms need every volume value finite in FP32
 * (WRT_E_ARG otherwise); the plain FP64 form propagates non-finite values as the reference
 * does. */
int wrt_radon_grid(wrt_ctx* ctx, const double* vol, int n, const double* phi, int nphi,
                   const double* theta, int ntheta, const double* shift, int nshift,
                   wrt_prec prec, double* out, wrt_stats* st);

/* wrt_radon_grid in its FP32 form, allocating nothing: no device allocation (hipMalloc,
 * hipFree or any library routine that allocates) and no host heap.  No counterpart in the
 * reference: it is for callers that cannot allocate, such as tight serving loops and
 * harnesses that forbid allocation.  Every device buffer is carved from the library's fixed
 * scratch of 64 MiB per GPU, held from wrt_init to the end of the process; host tables use
 * static storage.  Limits: nphi * ntheta <= 65536, nshift <= 8192, and the call's bricks,
 * one 32-plane FP64 slab of the volume, its tables and 64-bit plane sums within the scratch:
 * about 6.7e6 planes at n = 128, 3.5e6 at n = 192, and no n above 224 (the 32^3 FP32 bricks
 * of a 256^3 volume alone fill it); beyond them it returns WRT_E_SCRATCH before any device
 * work and writes nothing.  Same arguments, output
 * layout and values as wrt_radon_grid(..., WRT_F32, ...), bit for bit when that call walks
 * the same 8 x 8 bands (WRT_BAND_ROWS=8).  Synchronous; calls that overlap on one GPU, or
 * from several host threads, share the scratch and race. */
int wrt_radon_grid_noalloc(wrt_ctx* ctx, const double* vol, int n, const double* phi,
                           int nphi, const double* theta, int ntheta, const double* shift,
                           int nshift, double* out, wrt_stats* st);

/* Line integrals in z-slices: replaces ray_transform (ray_grid/src/integral.c:11-50)
 * called from ray_grid/src/ray_grid.c:242-255.  out[z][s][phi]. */
int wrt_ray_grid(wrt_ctx* ctx, const double* vol, int n, const double* z, int nz,
                 const double* shift, int nshift, const double* phi, int nphi, wrt_prec prec,
                 double* out, wrt_stats* st);

/* Ray data -> Radon data: replaces radon_reduction + ray_interpolate
 * (radon_reduction/slicing.c:5-119) called from radon_reduction/reduction.c:213-229.
 * rays[z][s][phi] share the output's shift and phi grids; nz is the number of z-slices
 * (the reference's "ngrid").  out[s][phi][theta].  FP64 only. */
int wrt_radon_reduction(wrt_ctx* ctx, const double* rays, int nz, const double* shift,
                        int nshift, const double* phi, int nphi, const double* theta,
                        int ntheta, double* out, wrt_stats* st);

/* The reduction on independent ray and plane grids (no reference counterpart: the
 * reference indexes its rays with the output's phi and shift grids,
 * radon_reduction/slicing.c:50-60,114).  rays[z][s][k] on nz slices
 * z_l = -1 + l*2/(nz-1), the increasing ray shifts rshift[nrshift] and nrphi >= 4 angles
 * phi_k = k*pi/nrphi (half_turn = 1; rshift must then be symmetric about 0) or
 * k*2pi/nrphi (half_turn = 0).  The row at each plane azimuth phi[j] (any real value) comes
 * from P(z, s, phi + pi) = P(z, -s, phi) and 4-point Lagrange interpolation over the
 * periodic angle grid; an azimuth on a ray angle takes that row alone, so with
 * half_turn = 0 and identical grids the result equals wrt_radon_reduction's bit for bit.
 * Splines over rshift, planes at the offsets shift[nshift]; out[s][phi][theta].  FP64. */
int wrt_radon_reduction_rays(wrt_ctx* ctx, const double* rays, int nz, const double* rshift,
                             int nrshift, int nrphi, int half_turn, const double* shift,
                             int nshift, const double* phi, int nphi, const double* theta,
                             int ntheta, double* out, wrt_stats* st);

/* Backprojection at the n^3 voxel centres: replaces radon3d_adjoint + radon_interpolate_s
 * (radon_inversion/backprojection/C/src/integral.c:6-24, interp.c:5-28) called from
 * radon3d_ajoint.c:240-254.  radon[s][phi][theta]; out[z][x][y]. */
int wrt_radon_adjoint(wrt_ctx* ctx, const double* radon, const double* phi, int nphi,
                      const double* theta, const double* theta_w, int ntheta,
                      const double* shift, int nshift, int n, wrt_prec prec, double* out,
                      wrt_stats* st);

/* Exact transposes of wrt_radon_grid and wrt_ray_grid (no reference counterpart): the
 * volume v with <A f, g> = <f, v> for every f.  g in the forward's output layout;
 * out[x][y][z]. */
int wrt_radon_grid_T(wrt_ctx* ctx, const double* g, int n, const double* phi, int nphi,
                     const double* theta, int ntheta, const double* shift, int nshift,
                     wrt_prec prec, double* out, wrt_stats* st);
int wrt_ray_grid_T(wrt_ctx* ctx, const double* g, int n, const double* z, int nz,
                   const double* shift, int nshift, const double* phi, int nphi, wrt_prec prec,
                   double* out, wrt_stats* st);

/* Work accounting (no reference counterpart): the number of lattice samples of the whole
 * plane grid that pass zero_interp's cube test (radon_grid/src/interp.c:65-66), i.e. that
 * read a voxel; computed exactly on the GPU with the reference's arithmetic. */
int wrt_radon_count_incube(wrt_ctx* ctx, int n, const double* phi, int nphi,
                           const double* theta, int ntheta, const double* shift, int nshift,
                           double* count);

/* --------------------------------------------------------------- file formats --- */
/* The reference's text files, read and written in parallel with the same values and bytes
 * as its fscanf / fprintf loops, and a binary format for chaining programs without text.
 * wrt_read_text: count records "%lf\n" (cols = 1) or "%lf, %lf, %lf, %lf\n" keeping the
 * last value (cols = 4), as tfunction_values_init (radon_grid/src/init.c:107-139),
 * ray_values_init (radon_reduction/init.c:116-164) and radon_values_init
 * (radon_inversion/backprojection/C/src/init.c:117-153) read them; cols = 2 or 3 reads the
 * shorter comma-separated records the same way (the 2D Radon data "s, phi, value" that
 * rtft2d.m:16-17 takes the last column of); WRT_E_IO on a short or
 * malformed file (with the reference programs' messages on stderr). */
int wrt_read_text(const char* path, size_t count, int cols, double* out);
/* rows "%lf, %lf, %lf, %lf\n" = a[i], b[j], c[k], values[(i*nb + j)*nc + k] in the order
 * i -> j -> k, as the drivers write them (radon_grid/src/radon_grid.c:248-258,
 * ray_grid/src/ray_grid.c:244-255, the analytic programs) */
int wrt_write_text_grid(const char* path, const double* a, int na, const double* b, int nb,
                        const double* c, int nc, const double* values);
/* rows "%lf\n" (radon_reduction/reduction.c:213-229, radon3d_ajoint.c:240-254) */
int wrt_write_text_values(const char* path, const double* values, size_t count);
/* binary (no reference counterpart): magic "WRTF64\0\1", uint64 count, little-endian
 * binary64 values in the text files' element order (the value column only) */
int wrt_write_f64(const char* path, const double* values, size_t count);
int wrt_read_f64(const char* path, size_t count, double* out);

/* --------------------------------------------------------- spectral Laplacians --- */
/* The step after the backprojection (the inversion is the Laplacian of the adjoint in 3D):
 * output = ifftn(U .* fftn(u)), U = -(2 pi / period)^2 * |k|^2 with the signed integer
 * frequencies k (index i <= n/2 -> i, else i - n).  FP64; u and out are n^3 (n^2) arrays in
 * any index order (the operator is the same along every axis); the real part is returned.
 * Replace laplace3dfft (radon_inversion/backprojection/laplace3dfft.m:1-25) and laplace2dfft
 * (radon_inversion/backprojection/laplace2dfft.m:1-23). */
int wrt_laplace3d(wrt_ctx* ctx, const double* u, int n, double period, double* out, wrt_stats* st);
int wrt_laplace2d(wrt_ctx* ctx, const double* u, int n, double period, double* out, wrt_stats* st);

/* ------------------------------------------------- Fourier-slice inversion --- */
/* Chang-type inversion through the Fourier slice theorem (radon_inversion/chang3d.m:3-30):
 * rtft3d (radon_inversion/rtft3d.m:3-137) Fourier transforms every direction's data along s,
 * zero-padded to ntotal = padding * floor(1/dshift) points (dshift = 2 rsupp/(nshift-1)),
 * with the shift correction exp(2 pi i rsupp k dfreq) and the jacobian weights
 * 0.5 dshift dfreq dphi w_theta rho^2; nfft_reconstruct_3d (radon_inversion/
 * nfft_reconstruct_3d.m:19-53) sums the weighted values back over the nodes rho * u(phi,
 * theta) at the points p = k * 2/ngrid, k in [-ngrid/2, ngrid/2)^3 (the adjoint NFFT; here
 * evaluated exactly), real part, divided by w0.  radon[s][phi][theta] on the grids given
 * (uniform phi and s, Gauss-Legendre theta with weights theta_w); ngrid even.
 * out[ix][iy][iz] = f((ix, iy, iz) * 2/ngrid - 1), i.e. the reference's
 * test_function(iy, ix, iz).  FP64.  ntotal must be an even integer (WRT_E_ARG otherwise). */
int wrt_fourier_inversion(wrt_ctx* ctx, const double* radon, const double* phi, int nphi,
                          const double* theta, const double* theta_w, int ntheta, int nshift,
                          double rsupp, double padding, double w0, int ngrid, double* out,
                          wrt_stats* st);

/* The same in 2D (radon_inversion/chang2d.m:3-26 = rtft2d.m:3-62 then
 * nfft_reconstruct_2d.m:19-45): radon[s][phi] on uniform phi and s, directions
 * (cos phi, sin phi), weights 0.5 dshift dfreq dphi |rho|; out[ix][iy] = f((ix, iy) * 2/ngrid
 * - 1), i.e. the reference's test_function(iy, ix).  ngrid even.  FP64. */
int wrt_fourier_inversion2d(wrt_ctx* ctx, const double* radon, const double* phi, int nphi,
                            int nshift, double rsupp, double padding, double w0, int ngrid,
                            double* out, wrt_stats* st);

/* ------------------------------------------------------ analytic test functions --- */
/* The analytic programs (radon_analytic, ray_analytic) integrate a test function that the
 * reference compiles in from test_function.c.  Here the test function is a sum of terms
 * evaluated on the GPU; every term vanishes outside a known ball, so the lattice points
 * where the sum is 0 are skipped (they add exactly nothing to the reference's sums). */
typedef enum {
  /* exp(-1/(r2max - |x-c|^2)) where |x-c|^2 < r2max - eps, else 0; p[0] = r2max, p[1] = eps
   * (radon_analytic/examples-test-functions/exponential_bump.c:5-16) */
  WRT_FN_EXP_BUMP = 0,
  /* sine_bump_template (radon_analytic/examples-test-functions/sinusoidal_bump.c:3-19);
   * p[0] = amplitude, p[1] = rin, p[2] = rout */
  WRT_FN_SINE_BUMP = 1,
  /* solid ellipsoid (no reference counterpart; the phantom's ellipsoid): p[0] = density,
   * p[1..3] = semi-axes, p[4..12] = rotation R row-major (columns are the principal axes);
   * inside when |diag(1/a) R^T (x - c)| <= 1 */
  WRT_FN_ELLIPSOID = 2
} wrt_fn_kind;

typedef struct wrt_fn_term {
  int kind;    /* wrt_fn_kind */
  double c[3]; /* centre */
  double p[13];
} wrt_fn_term;

#define WRT_FN_MAX_TERMS 256

/* The reference's example test functions as terms: "exponential_bump" (1 term) and
 * "sinusoidal_bump" (3 terms).  Returns the number of terms (<= max_terms) or WRT_E_ARG. */
int wrt_fn_example(const char* name, wrt_fn_term* terms, int max_terms);
/* the ellipsoids of wrt_phantom(n, ., nell, seed, 0) as nell WRT_FN_ELLIPSOID terms */
int wrt_fn_phantom(int nell, unsigned long long seed, wrt_fn_term* terms);
/* z_i = (-1.0) + i*2.0/(ngrid-1): the analytic ray program's slices
 * (ray_analytic/ray_analytic.c:255; rounds differently from wrt_zslice_grid) */
int wrt_analytic_zslice_grid(int ngrid, double* z);

/* Plane integrals of a test function over every (shift, phi, theta), FP64: replaces
 * plane_integral (radon_analytic/src/integral.c:25-56) called from
 * radon_analytic/src/radon_analytic.c:253-266.  delta = (2/(ngrid-1))/2 and the lattice
 * i1, i2 in [-2 ngrid, 2 ngrid).  out[s][phi][theta]. */
int wrt_radon_analytic(wrt_ctx* ctx, const wrt_fn_term* terms, int nterms, int ngrid,
                       const double* phi, int nphi, const double* theta, int ntheta,
                       const double* shift, int nshift, double* out, wrt_stats* st);
/* Line integrals of a test function in z-slices, FP64: replaces ray_integral
 * (ray_analytic/integral.c:10-47) called from ray_analytic/ray_analytic.c:253-272.
 * i in [-2 ngrid, 2 ngrid].  out[z][s][phi]. */
int wrt_ray_analytic(wrt_ctx* ctx, const wrt_fn_term* terms, int nterms, int ngrid,
                     const double* z, int nz, const double* shift, int nshift,
                     const double* phi, int nphi, double* out, wrt_stats* st);

/* ---------------------------------------------- device-resident jobs (batch/serve) --- */
/* A job owns its device input, output and tables; load/run/fetch may repeat.  The grids
 * passed at creation are copied.  Unused grid fields may be NULL/0 for an op:
 *   RADON, RADON_T : n, phi, theta, shift        RAY, RAY_T : n, z, shift, phi
 *   REDUCTION      : nz, shift, phi, theta       ADJOINT    : n, phi, theta, theta_w, shift */
typedef struct wrt_job_desc {
  wrt_op op;
  wrt_prec prec;
  int n;
  const double* phi;
  int nphi;
  const double* theta;
  const double* theta_w;
  int ntheta;
  const double* shift;
  int nshift;
  const double* z;
  int nz;
} wrt_job_desc;

typedef struct wrt_job wrt_job;

/* device bytes a job of this description holds (admission control) */
size_t wrt_job_bytes(const wrt_job_desc* d);
size_t wrt_job_input_count(const wrt_job_desc* d);
size_t wrt_job_output_count(const wrt_job_desc* d);
int wrt_job_create(wrt_ctx* ctx, const wrt_job_desc* d, wrt_job** job);
/* host -> device input (converted to the job's precision/layout); a RADON job in FP32 or
 * BF16 refuses a volume with a value that is not finite in FP32 (WRT_E_ARG) */
int wrt_job_load(wrt_job* job, const double* in);
/* synthetic input from the seeded generators below: an ellipsoid phantom for RADON (cell
 * centred) and RAY (node centred), Gaussian data for the other ops */
int wrt_job_load_synthetic(wrt_job* job, unsigned long long seed, int n_ellipsoids,
                           double texture_sigma);
/* enqueue on the job's own stream and wait; fills st when non-NULL */
int wrt_job_run(wrt_job* job, wrt_stats* st);
/* asynchronous form: wrt_job_launch enqueues the job's device work on its own stream and
 * returns; launches of different jobs overlap on the device.  wrt_job_poll returns 1 while
 * the work runs and WRT_OK (stats filled) once it is done; wrt_job_wait blocks.  A launched
 * job accepts no load/launch/fetch until polled done or waited for (WRT_E_STATE). */
int wrt_job_launch(wrt_job* job);
int wrt_job_poll(wrt_job* job, wrt_stats* st);
int wrt_job_wait(wrt_job* job, wrt_stats* st);
int wrt_job_fetch(wrt_job* job, double* out);
void wrt_job_destroy(wrt_job* job);
/* The tile shape an FP32/BF16 plane-integral (RADON) job's brick kernel runs with (no
 * reference counterpart): rows per band of the section walk (4, 8 or 16; 64 / rows columns
 * per step) and directions per task.  The first launch after a load picks them: the fastest
 * of the band shapes at the load's direction group, then of half and twice that group,
 * timed once per problem shape (edge, directions, shifts, precision, non-empty bricks) and
 * kept in a tuning cache file (WRT_TUNE_CACHE, else $HOME/.cache/wrt_tune.txt);
 * WRT_AUTOTUNE=0 keeps 8 rows and the load's group, WRT_BAND_ROWS and WRT_RADON_DGROUP fix
 * them.  WRT_E_STATE before that first launch, WRT_E_ARG for other jobs. */
int wrt_job_tuning(const wrt_job* job, int* band_rows, int* dgroup);

/* Run many jobs (forward and adjoint, uneven sizes) as one submission (no reference
 * counterpart).  The FP32/BF16 plane-integral jobs and the FP32 backprojection jobs run as
 * one persistent launch captured in a HIP graph: their tasks (brick x direction group,
 * blocks of 16 backprojection voxel tiles) are ordered longest first and taken by the CUs
 * from a device counter.
 * The direction-group size is autotuned on the first call for a list shape (cached in
 * memory and in the file WRT_TUNE_CACHE or $HOME/.cache/wrt_tune.txt; WRT_AUTOTUNE=0
 * disables tuning, WRT_BATCH_DGROUP=G >= 16 fixes the size).  The graph is replayed while the
 * list and the jobs' loads stay the same; any other job (plain FP64 forms, reductions) runs
 * on its own stream beside the graph.  Each job's output then equals its wrt_job_run output,
 * bit for bit when that run walks the persistent kernel's 8 x 8 bands (WRT_BAND_ROWS=8;
 * plane sums are exact 64-bit fixed-point sums in any order).  st: kernel_seconds = the
 * graph's device time, seconds = the longest of it and the other jobs' device times. */
int wrt_batch_run(wrt_ctx* ctx, wrt_job* const* jobs, int njobs, wrt_stats* st);
/* direction-group size of the context's current batch plan (0 if none) */
int wrt_batch_dgroup(wrt_ctx* ctx);

/* ------------------------------------------------------------------- serve loop --- */
/* A request stream served from one context (no reference counterpart: the reference runs
 * one program per input file).  Requests arrive as a seeded Poisson process; each names an
 * operation, a cube edge e and one of `ndatasets` seeded datasets of that (op, e), on the
 * grids (e/2) phi x (e/4) GL theta x (e+1) shifts (REDUCTION: rays on e z-slices with the
 * same phi and shift grids).  RADON and ADJOINT run in FP32, REDUCTION in FP64.
 * Admission: a request's dataset becomes resident (job created and loaded) when the free
 * device memory hipMemGetInfo reports at that moment holds its bytes plus 1 GiB, and the
 * optional budget allows, evicting idle resident datasets least recently used first; a
 * resident dataset serves later requests without reloading.  The fast requests (RADON and
 * ADJOINT in FP32) are batched continuously: their tasks (bricks x direction groups, voxel
 * blocks) go into rounds of a few milliseconds of one persistent kernel, small requests
 * whole and large ones in chunks, so a request arriving while a large one runs starts
 * within a round or two; rounds alternate between two streams so that one fills the CUs the
 * other frees.  The others (REDUCTION, and every request in plain mode) run on their
 * dataset's own stream, up to max_inflight at once.  A request whose dataset serves an
 * earlier request waits for it.  Latency = completion - arrival, on the host's monotonic
 * clock. */
typedef struct wrt_serve_cfg {
  double qps;          /* mean arrival rate (1/s); <= 0: every request arrives at t = 0 */
  int nreq;            /* number of requests */
  unsigned long long seed;
  double mix[3];       /* relative weights of RADON, ADJOINT, REDUCTION */
  const int* edges;    /* cube edges drawn uniformly (each even, >= 8) */
  int nedges;
  int ndatasets;       /* distinct datasets per (op, edge), >= 1 */
  size_t hbm_budget;   /* device bytes for resident datasets; 0: 90% of the free HBM */
  int max_inflight;    /* concurrent requests on their own streams (not batched), >= 1 */
  int plain;           /* 1: every request runs the plain FP64 form, one at a time (the
                          per-request baseline the served form is compared with) */
} wrt_serve_cfg;

typedef struct wrt_serve_req {
  wrt_op op;
  wrt_prec prec;
  int edge;
  int dataset;
  unsigned long long data_seed; /* seed of the dataset's synthetic input */
  double arrival_s;             /* arrival time after the start of the stream */
} wrt_serve_req;

typedef struct wrt_serve_report {
  int completed;
  double wall_s;                 /* start of the stream to the last completion */
  double p50_ms, p99_ms, mean_ms, max_ms;
  double samples, samples_per_s; /* logical samples of all requests, per wall second */
  int loads, evictions, peak_inflight;
  size_t peak_bytes;             /* largest device bytes held by resident datasets */
  int batched;                   /* requests served through the rounds */
  int rounds;                    /* rounds launched */
} wrt_serve_report;

/* request i of the stream cfg describes (deterministic in cfg) */
int wrt_serve_describe(const wrt_serve_cfg* cfg, int i, wrt_serve_req* req);
/* the grids of an edge: phi, theta, theta_w hold e/2, e/4, e/4 values, shift e+1, z e */
int wrt_serve_grids(wrt_ctx* ctx, int edge, double* phi, double* theta, double* theta_w,
                    double* shift, double* z);
/* a request's job description on grids from wrt_serve_grids (pointers kept, not copied) */
int wrt_serve_desc(const wrt_serve_req* req, const double* phi, const double* theta,
                   const double* theta_w, const double* shift, const double* z,
                   wrt_job_desc* d);
/* the input of a request's dataset: wrt_job_input_count(desc) doubles */
int wrt_serve_input(const wrt_serve_req* req, double* in);
/* Serve the stream.  latency_ms (nreq entries) and outputs (nreq pointers, NULL entries
 * skipped; each non-NULL one receives wrt_job_output_count doubles) are optional. */
int wrt_serve_run(wrt_ctx* ctx, const wrt_serve_cfg* cfg, wrt_serve_report* rep,
                  double* latency_ms, double* const* outputs);

/* ----------------------------------------------------------- synthetic inputs --- */
/* Seeded ellipsoid phantom on [-1,1]^3 (no reference counterpart): nell ellipsoids with
 * centres uniform in |c| < 0.5, semi-axes U(0.05,0.4), uniform random rotations and
 * densities U(-0.5,1.0), summed and cut to the unit ball, plus N(0, sigma^2) texture inside
 * the ball when sigma > 0.  cell_centred = 1 samples x_i = -1 + (2i+1)/n (the cells of
 * radon_grid's zero_interp); 0 samples x_i = -1 + 2i/(n-1) (the nodes of ray_grid's
 * trilinear interpolation).  vol[x][y][z].  Identical bits on every host. */
int wrt_phantom(int n, int cell_centred, int nell, unsigned long long seed, double sigma,
                double* vol);
/* the ellipsoids wrt_phantom draws: centre[3], axes[3], rot[9] (row-major, columns are
 * the principal axes), density -- 16 doubles per ellipsoid */
int wrt_phantom_params(int nell, unsigned long long seed, double* params);
/* count standard normal deviates from a seeded counter-based generator */
int wrt_gaussian(size_t count, unsigned long long seed, double* out);

#ifdef __cplusplus
}
#endif
#endif /* WRT_H */
