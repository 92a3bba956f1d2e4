/*
 * ray_grid -- drop-in for the reference program ray_grid (ray_grid/src/ray_grid.c).
 * Config order: ngrid, nslices, nshift, nphi (ray_grid.c:177-180).  Input: ngrid^3 lines
 * "%lf".  Output: "%lf, %lf, %lf, %lf\n" = z, s, phi, Pf in the order z -> s -> phi
 * (ray_grid.c:242-255), plus readme_<output>.
 */
#include "cli_common.h"

int main(int argc, char** argv) {
  cli_opts o;
  cli_parse(argc, argv, &o, 0);
  int p[4];
  printf("  Reading parameters from %s...", o.params);
  if (cli_params(o.params, p) != 0) { perror("Aborting : parameters_init.\n"); exit(EXIT_FAILURE); }
  printf("Done.\n");
  const int n = p[0], nz = p[1], nshift = p[2], nphi = p[3];
  if (n < 2 || nz < 2 || nshift < 2 || nphi < 1) { fprintf(stderr, "Aborting : invalid grid sizes\n"); exit(EXIT_FAILURE); }
  printf("  Initializing test-function data...");
  double* vol = (double*)cli_alloc(sizeof(double) * (size_t)n * n * n);
  if (cli_read_input(&o, (size_t)n * n * n, 1, vol) != 0) { perror("Aborting : tfunction_values_init.\n"); exit(EXIT_FAILURE); }
  printf("Done.\n");
  double* phi = (double*)cli_alloc(sizeof(double) * nphi);
  double* shift = (double*)cli_alloc(sizeof(double) * nshift);
  double* z = (double*)cli_alloc(sizeof(double) * nz);
  printf("  Initializing the grid in ray transforms space...");
  wrt_ray_grid_init(nshift, nphi, shift, phi);
  wrt_zslice_grid(nz, z);
  printf("Done.\n\n");
  wrt_ctx* ctx = NULL;
  int rc = wrt_init(&ctx, o.device);
  if (rc) cli_die("wrt_init", rc);
  const size_t tot = (size_t)nz * nshift * nphi;
  double* out = (double*)cli_alloc(sizeof(double) * tot);
  const double t0 = cli_now();
  wrt_stats st;
  rc = wrt_ray_grid(ctx, vol, n, z, nz, shift, nshift, phi, nphi, o.prec, out, &st);
  if (rc) cli_die("wrt_ray_grid", rc);
  cli_write_output(o.bin_out, o.output, z, nz, shift, nshift, phi, nphi, out, tot);
  printf("\n  Number of threads: %d, elapsed time: %3.1f sec (GPU kernel %.3f sec)\n", o.nthreads,
         cli_now() - t0, st.kernel_seconds);
  const char* names[] = {"number of z-slices [z]", "number of shifts [s]  ", "number of angles [phi]"};
  const int vals[] = {nz, nshift, nphi};
  cli_readme(o.output, "Ray transforms in 3D (layer-by-layer)", "[z], [s], [phi], [Pf]", names, vals, 3);
  wrt_free(ctx);
  free(vol); free(phi); free(shift); free(z); free(out);
  return EXIT_SUCCESS;
}
