/*
 * radon_reduction -- drop-in for the reference program radon_reduction
 * (radon_reduction/reduction.c).  Config order: nshift, nphi, ntheta, ngrid = number of
 * z-slices (reduction.c:145-148).  Input: rays "%lf, %lf, %lf, %lf" (z, s, phi, P) in the
 * order z -> s -> phi on the output's own shift and phi grids, or one value per line with -s
 * (the reference's -s dereferences a missing argument; here it is a plain flag).
 * Output: "%lf\n" = Rf in the order s -> phi -> theta (reduction.c:213-229).
 */
#include "cli_common.h"

int main(int argc, char** argv) {
  cli_opts o;
  cli_parse(argc, argv, &o, 1);
  int p[4];
  printf("  Reading parameters from %s...", o.params);
  if (cli_params(o.params, p) != 0) { perror("Aborting : parameters_init.\n"); exit(EXIT_FAILURE); }
  printf("Done.\n");
  const int nshift = p[0], nphi = p[1], ntheta = p[2], nz = p[3];
  if (nshift < 3 || nphi < 1 || ntheta < 1 || nz < 3) { fprintf(stderr, "Aborting : invalid grid sizes\n"); exit(EXIT_FAILURE); }
  printf("  Initializing data for ray transforms...");
  const size_t nin = (size_t)nz * nshift * nphi;
  double* rays = (double*)cli_alloc(sizeof(double) * nin);
  if (cli_read_input(&o, nin, o.small ? 1 : 4, rays) != 0) { perror("Aborting : ray_values_init.\n"); exit(EXIT_FAILURE); }
  printf("Done.\n");
  wrt_ctx* ctx = NULL;
  int rc = wrt_init(&ctx, o.device);
  if (rc) cli_die("wrt_init", rc);
  double* phi = (double*)cli_alloc(sizeof(double) * nphi);
  double* theta = (double*)cli_alloc(sizeof(double) * ntheta);
  double* shift = (double*)cli_alloc(sizeof(double) * nshift);
  printf("  Initializing the grid in Radon space...");
  rc = wrt_radon_grid_init(ctx, nphi, ntheta, nshift, phi, theta, NULL, shift);
  if (rc) cli_die("wrt_radon_grid_init", rc);
  printf("Done.\n\n");
  const size_t tot = (size_t)nshift * nphi * ntheta;
  double* out = (double*)cli_alloc(sizeof(double) * tot);
  const double t0 = cli_now();
  wrt_stats st;
  rc = wrt_radon_reduction(ctx, rays, nz, shift, nshift, phi, nphi, theta, ntheta, out, &st);
  if (rc) cli_die("wrt_radon_reduction", rc);
  cli_write_output(o.bin_out, o.output, NULL, 0, NULL, 0, NULL, 0, out, tot);
  printf("\n  Number of threads: %d, elapsed time: %3.1f sec (GPU kernel %.3f sec)\n", o.nthreads,
         cli_now() - t0, st.kernel_seconds);
  const char* names[] = {"number of angles of [phi]  ", "number of angles of [theta]", "number of shifts of [s]    "};
  const int vals[] = {nphi, ntheta, nshift};
  cli_readme(o.output, "Radon transforms in 3D reduced from ray transforms", "[Rf] (s -> phi -> theta)", names, vals, 3);
  wrt_free(ctx);
  free(rays); free(phi); free(theta); free(shift); free(out);
  return EXIT_SUCCESS;
}
