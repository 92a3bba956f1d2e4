/*
 * radon_analytic -- drop-in for the reference program radon_analytic
 * (radon_analytic/src/radon_analytic.c).  Config order: nphi, ntheta, nshift, ngrid
 * (radon_analytic.c:213-216).  The test function, compiled into the reference from
 * test_function.c, is chosen here with -F (the reference's example functions) or -f (a term
 * file).  Output: "%lf, %lf, %lf, %lf\n" = s, phi, theta, Rf in the order s -> phi -> theta
 * (radon_analytic.c:253-266), plus readme_<output>.
 */
#include "cli_common.h"

int main(int argc, char** argv) {
  cli_an_opts o;
  cli_an_parse(argc, argv, &o);
  wrt_fn_term terms[WRT_FN_MAX_TERMS];
  const int nt = cli_an_terms(&o, terms, WRT_FN_MAX_TERMS);
  int p[4];
  printf("  Reading parameters from %s...", o.params);
  if (cli_params_n(o.params, p, 4) != 0) { perror("Aborting : parameters_init.\n"); exit(EXIT_FAILURE); }
  printf("Done.\n");
  const int nphi = p[0], ntheta = p[1], nshift = p[2], ngrid = p[3];
  if (nphi < 1 || ntheta < 1 || nshift < 2 || ngrid < 2) { fprintf(stderr, "Aborting : invalid grid sizes\n"); exit(EXIT_FAILURE); }
  wrt_ctx* ctx = NULL;
  int rc = wrt_init(&ctx, o.device);
  if (rc) cli_die("wrt_init", rc);
  double* phi = (double*)cli_alloc(sizeof(double) * nphi);
  double* theta = (double*)cli_alloc(sizeof(double) * ntheta);
  double* shift = (double*)cli_alloc(sizeof(double) * nshift);
  printf("  Initializing the grid in Radon space...");
  rc = wrt_radon_grid_init(ctx, nphi, ntheta, nshift, phi, theta, NULL, shift);
  if (rc) cli_die("radon_grid_init", rc);
  printf("Done.\n\n");
  const size_t tot = (size_t)nshift * nphi * ntheta;
  double* out = (double*)cli_alloc(sizeof(double) * tot);
  const double t0 = cli_now();
  wrt_stats st;
  rc = wrt_radon_analytic(ctx, terms, nt, ngrid, phi, nphi, theta, ntheta, shift, nshift, out, &st);
  if (rc) cli_die("wrt_radon_analytic", rc);
  cli_write_output(o.bin_out, o.output, shift, nshift, phi, nphi, theta, ntheta, out, tot);
  printf("\n  Number of threads: %d, elapsed time: %3.1f sec (GPU kernel %.3f sec)\n", o.nthreads,
         cli_now() - t0, st.kernel_seconds);
  const char* names[] = {"number of angles of [phi]  ", "number of angles of [theta]",
                         "number of shifts of [s]    "};
  const int vals[] = {nphi, ntheta, nshift};
  cli_readme(o.output, "Radon transforms in 3D of an analytic test-function", "[s], [phi], [theta], [Rf]",
             names, vals, 3);
  wrt_free(ctx);
  free(phi); free(theta); free(shift); free(out);
  return EXIT_SUCCESS;
}
