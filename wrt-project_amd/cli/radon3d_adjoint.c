/*
 * radon3d_adjoint -- drop-in for the reference program radon3d_ajoint
 * (radon_inversion/backprojection/C/src/radon3d_ajoint.c).  Config order: nphi, ntheta,
 * nshift, ngrid (radon3d_ajoint.c:181-184).  Input: Radon data "%lf, %lf, %lf, %lf"
 * (s, phi, theta, Rf) in the order s -> phi -> theta.  Output: "%lf\n" at the ngrid^3 voxel
 * centres in the order z -> x -> y (radon3d_ajoint.c:240-254).
 */
#include "cli_common.h"

int main(int argc, char** argv) {
  cli_opts o;
  cli_parse(argc, argv, &o, 0);
  int p[4];
  printf("  Reading parameters from %s...", o.params);
  if (cli_params(o.params, p) != 0) { perror("Aborting : init_parameters.\n"); exit(EXIT_FAILURE); }
  printf("Done.\n");
  const int nphi = p[0], ntheta = p[1], nshift = p[2], n = p[3];
  if (nphi < 1 || ntheta < 1 || nshift < 3 || n < 1) { fprintf(stderr, "Aborting : invalid grid sizes\n"); exit(EXIT_FAILURE); }
  wrt_ctx* ctx = NULL;
  int rc = wrt_init(&ctx, o.device);
  if (rc) cli_die("wrt_init", rc);
  double* phi = (double*)cli_alloc(sizeof(double) * nphi);
  double* theta = (double*)cli_alloc(sizeof(double) * ntheta);
  double* w = (double*)cli_alloc(sizeof(double) * ntheta);
  double* shift = (double*)cli_alloc(sizeof(double) * nshift);
  printf("  Reading Radon transforms...");
  rc = wrt_radon_grid_init(ctx, nphi, ntheta, nshift, phi, theta, w, shift);
  if (rc) cli_die("wrt_radon_grid_init", rc);
  const size_t nin = (size_t)nshift * nphi * ntheta;
  double* radon = (double*)cli_alloc(sizeof(double) * nin);
  if (cli_read_input(&o, nin, 4, radon) != 0) { perror("Aborting : radon_values_init.\n"); exit(EXIT_FAILURE); }
  printf("Done.\n\n");
  const size_t tot = (size_t)n * n * n;
  double* out = (double*)cli_alloc(sizeof(double) * tot);
  const double t0 = cli_now();
  wrt_stats st;
  rc = wrt_radon_adjoint(ctx, radon, phi, nphi, theta, w, ntheta, shift, nshift, n, o.prec, out, &st);
  if (rc) cli_die("wrt_radon_adjoint", rc);
  cli_write_output(o.bin_out, o.output, NULL, 0, NULL, 0, NULL, 0, out, tot);
  printf("\n  Number of threads: %d, elapsed time: %3.1f sec (GPU kernel %.3f sec)\n", o.nthreads,
         cli_now() - t0, st.kernel_seconds);
  const char* names[] = {"number of grid points [ngrid]", "number of angles of [phi]    ",
                         "number of angles of [theta]  ", "number of shifts of [s]      "};
  const int vals[] = {n, nphi, ntheta, nshift};
  cli_readme(o.output, "Adjoint Radon transform in 3D", "[f] (z -> x -> y)", names, vals, 4);
  wrt_free(ctx);
  free(phi); free(theta); free(w); free(shift); free(radon); free(out);
  return EXIT_SUCCESS;
}
