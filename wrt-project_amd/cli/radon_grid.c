/*
 * radon_grid -- drop-in for the reference program radon_grid (radon_grid/src/radon_grid.c).
 * Config order: nphi, ntheta, nshift, ngrid (radon_grid.c:182-185).  Input: ngrid^3 lines
 * "%lf" (x-major, z fastest).  Output: "%lf, %lf, %lf, %lf\n" = s, phi, theta, Rf in the
 * order s -> phi -> theta (radon_grid.c:248-258), plus readme_<output>.
 */
#include "cli_common.h"

int main(int argc, char** argv) {
  cli_opts o;
  cli_parse(argc, argv, &o, 0);
  int p[4];
  printf("  Reading parameters from %s...", o.params);
  if (cli_params(o.params, p) != 0) { perror("Aborting : parameters_init failed.\n"); exit(EXIT_FAILURE); }
  printf("Done.\n");
  const int nphi = p[0], ntheta = p[1], nshift = p[2], n = p[3];
  if (nphi < 1 || ntheta < 1 || nshift < 2 || n < 1) { fprintf(stderr, "Aborting : invalid grid sizes\n"); exit(EXIT_FAILURE); }
  printf("  Initializing data for the test-function...");
  double* vol = (double*)cli_alloc(sizeof(double) * (size_t)n * n * n);
  if (cli_read_input(&o, (size_t)n * n * n, 1, vol) != 0) { perror("Aborting : tfunction_values_init\n"); exit(EXIT_FAILURE); }
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
  rc = wrt_radon_grid(ctx, vol, n, phi, nphi, theta, ntheta, shift, nshift, o.prec, out, &st);
  if (rc) cli_die("wrt_radon_grid", rc);
  cli_write_output(o.bin_out, o.output, shift, nshift, phi, nphi, theta, ntheta, out, tot);
  printf("\n  Number of threads: %d, elapsed time: %3.1f sec (GPU kernel %.3f sec)\n", o.nthreads,
         cli_now() - t0, st.kernel_seconds);
  const char* names[] = {"number of grid ponts [ngrid]", "number of angles of [phi]  ",
                         "number of angles of [theta]", "number of shifts of [s]    "};
  const int vals[] = {n, nphi, ntheta, nshift};
  printf("  Generating readme...");
  cli_readme(o.output, "Radon transforms in 3D for test-function", "[s], [phi], [theta], [Rf]", names, vals, 4);
  printf("  Done.\n");
  wrt_free(ctx);
  free(vol); free(phi); free(theta); free(shift); free(out);
  return EXIT_SUCCESS;
}
