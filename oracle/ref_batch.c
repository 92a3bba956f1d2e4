/*
 * ref_batch.c -- OpenMP batch drivers around the reference's own per-element functions,
 * linked with the reference's sources by oracle/Makefile into oracle/_ref (test and
 * benchmark infrastructure only: the CPU baseline of bench.py).
 *
 * The reference's programs call plane_integral / ray_transform once per output element from
 * an OpenMP loop over the outer index (radon_grid/src/radon_grid.c:229-264,
 * ray_grid/src/ray_grid.c:224-262); these drivers do the same over a caller's list of
 * seeded elements, one thread per core, so that a bounded sample times the reference's own
 * code without any per-call interpreter overhead.  Build with -DREF_PLANE (radon_grid's
 * sources) or -DREF_RAY (ray_grid's sources).
 */
#include <omp.h>

#ifdef REF_PLANE
/* radon_grid/src/integral.h:4-5 */
double plane_integral(double*** values, int ngrid, double phi, double theta, double shift);

void ref_plane_batch(double*** values, int ngrid, const double* phi, const double* theta,
                     const double* shift, long count, int threads, double* out) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
  for (long k = 0; k < count; ++k) out[k] = plane_integral(values, ngrid, phi[k], theta[k], shift[k]);
}
#endif

#ifdef REF_RAY
/* ray_grid/src/integral.h:4-5 */
double ray_transform(double*** values, int ngrid, double zslice, double shift, double phi);

void ref_ray_batch(double*** values, int ngrid, const double* z, const double* shift,
                   const double* phi, long count, int threads, double* out) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
  for (long k = 0; k < count; ++k) out[k] = ray_transform(values, ngrid, z[k], shift[k], phi[k]);
}
#endif
