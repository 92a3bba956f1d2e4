/*
 * ray_analytic -- drop-in for the reference program ray_analytic (ray_analytic/ray_analytic.c).
 * Config order: nshift, nphi, ngrid (= number of z-slices) (ray_analytic.c:201-203).  The test
 * function is chosen with -F (the reference's example functions) or -f (a term file).
 * Output: "%lf, %lf, %lf, %lf\n" = z, s, phi, Pf in the order z -> s -> phi
 * (ray_analytic.c:253-272) with z_i = -1 + i*2.0/(ngrid-1) (:255), plus readme_<output>.
 */
#include "cli_common.h"

int main(int argc, char** argv) {
  cli_an_opts o;
  cli_an_parse(argc, argv, &o);
  wrt_fn_term terms[WRT_FN_MAX_TERMS];
  const int nt = cli_an_terms(&o, terms, WRT_FN_MAX_TERMS);
  int p[3];
  printf("  Reading parameters from %s...", o.params);
  if (cli_params_n(o.params, p, 3) != 0) { perror("Aborting : parameters_init.\n"); exit(EXIT_FAILURE); }
  printf("Done.\n");
  const int nshift = p[0], nphi = p[1], ngrid = p[2];
  if (nshift < 2 || nphi < 1 || ngrid < 2) { fprintf(stderr, "Aborting : invalid grid sizes\n"); exit(EXIT_FAILURE); }
  double* phi = (double*)cli_alloc(sizeof(double) * nphi);
  double* shift = (double*)cli_alloc(sizeof(double) * nshift);
  double* z = (double*)cli_alloc(sizeof(double) * ngrid);
  printf("  Initializing the grid of rays...");
  wrt_ray_grid_init(nshift, nphi, shift, phi);
  wrt_analytic_zslice_grid(ngrid, z);
  printf("Done.\n\n");
  wrt_ctx* ctx = NULL;
  int rc = wrt_init(&ctx, o.device);
  if (rc) cli_die("wrt_init", rc);
  const size_t tot = (size_t)ngrid * nshift * nphi;
  double* out = (double*)cli_alloc(sizeof(double) * tot);
  const double t0 = cli_now();
  wrt_stats st;
  rc = wrt_ray_analytic(ctx, terms, nt, ngrid, z, ngrid, shift, nshift, phi, nphi, out, &st);
  if (rc) cli_die("wrt_ray_analytic", rc);
  cli_write_output(o.bin_out, o.output, z, ngrid, shift, nshift, phi, nphi, out, tot);
  printf("\n  Number of threads: %d, elapsed time: %3.1f sec (GPU kernel %.3f sec)\n", o.nthreads,
         cli_now() - t0, st.kernel_seconds);
  const char* names[] = {"number of z-slices [nz]", "number of shifts   [s] ", "number of angles   [phi]"};
  const int vals[] = {ngrid, nshift, nphi};
  cli_readme(o.output, "Ray transforms in 3D of an analytic test-function", "[z], [s], [phi], [Pf]",
             names, vals, 3);
  wrt_free(ctx);
  free(phi); free(shift); free(z); free(out);
  return EXIT_SUCCESS;
}
