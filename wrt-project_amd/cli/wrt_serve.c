/*
 * wrt_serve -- serve a seeded Poisson stream of projection (FP32), backprojection (FP32) and
 * reduction (FP64) requests on one GPU and print one JSON line: latency percentiles,
 * throughput, loads and evictions.  No reference counterpart (the reference runs one
 * program per input file); the operators are the same as the drop-in programs'.
 *
 *   wrt_serve [-q qps] [-r requests] [-s seed] [-e 64,96,128,192,256] [-m 0.5,0.3,0.2]
 *             [-D datasets] [-b budget_MiB] [-f max_inflight] [-d device] [-F]
 * -F runs every request in the plain FP64 form, one at a time (the per-request baseline).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "wrt.h"

static int parse_list(const char* s, double* v, int max) {
  int n = 0;
  while (*s && n < max) {
    char* end;
    v[n++] = strtod(s, &end);
    if (end == s) return -1;
    s = *end == ',' ? end + 1 : end;
  }
  return n;
}

int main(int argc, char** argv) {
  wrt_serve_cfg cfg;
  memset(&cfg, 0, sizeof(cfg));
  cfg.qps = 8.0;
  cfg.nreq = 64;
  cfg.seed = 1;
  cfg.mix[0] = 0.5;
  cfg.mix[1] = 0.3;
  cfg.mix[2] = 0.2;
  cfg.ndatasets = 2;
  cfg.max_inflight = 4;
  int edges[16] = {64, 96, 128, 192, 256};
  cfg.nedges = 5;
  int device = 0;
  double v[16];
  int c, k;
  while ((c = getopt(argc, argv, "hq:r:s:e:m:D:b:f:d:F")) != -1) {
    switch (c) {
      case 'q': cfg.qps = atof(optarg); break;
      case 'r': cfg.nreq = atoi(optarg); break;
      case 's': cfg.seed = strtoull(optarg, NULL, 10); break;
      case 'e':
        k = parse_list(optarg, v, 16);
        if (k < 1) { fprintf(stderr, "wrt_serve: bad edge list\n"); return EXIT_FAILURE; }
        for (int i = 0; i < k; ++i) edges[i] = (int)v[i];
        cfg.nedges = k;
        break;
      case 'm':
        if (parse_list(optarg, v, 3) != 3) { fprintf(stderr, "wrt_serve: -m wants 3 weights\n"); return EXIT_FAILURE; }
        for (int i = 0; i < 3; ++i) cfg.mix[i] = v[i];
        break;
      case 'D': cfg.ndatasets = atoi(optarg); break;
      case 'b': cfg.hbm_budget = (size_t)(atof(optarg) * 1048576.0); break;
      case 'f': cfg.max_inflight = atoi(optarg); break;
      case 'd': device = atoi(optarg); break;
      case 'F': cfg.plain = 1; break;
      case 'h':
      default:
        fprintf(c == 'h' ? stdout : stderr,
                "Usage: %s [-q qps] [-r requests] [-s seed] [-e edges] [-m radon,adjoint,reduction]\n"
                "          [-D datasets] [-b budget_MiB] [-f max_inflight] [-d device] [-F]\n",
                argv[0]);
        return c == 'h' ? EXIT_SUCCESS : EXIT_FAILURE;
    }
  }
  cfg.edges = edges;
  wrt_ctx* ctx = NULL;
  int rc = wrt_init(&ctx, device);
  if (rc != WRT_OK) {
    fprintf(stderr, "wrt_serve: %s\n", wrt_strerror(rc));
    return EXIT_FAILURE;
  }
  wrt_serve_report rep;
  rc = wrt_serve_run(ctx, &cfg, &rep, NULL, NULL);
  if (rc != WRT_OK) {
    fprintf(stderr, "wrt_serve: %s\n", wrt_strerror(rc));
    wrt_free(ctx);
    return EXIT_FAILURE;
  }
  printf("{\"requests\": %d, \"form\": \"%s\", \"qps\": %g, \"wall_s\": %.4f, \"p50_ms\": %.3f, \"p99_ms\": %.3f, "
         "\"mean_ms\": %.3f, \"max_ms\": %.3f, \"samples\": %.6g, \"samples_per_s\": %.6g, "
         "\"loads\": %d, \"evictions\": %d, \"peak_inflight\": %d, \"peak_bytes\": %zu, "
         "\"batched\": %d, \"rounds\": %d}\n",
         rep.completed, cfg.plain ? "plain f64, one at a time" : "served", cfg.qps, rep.wall_s, rep.p50_ms, rep.p99_ms, rep.mean_ms, rep.max_ms,
         rep.samples, rep.samples_per_s, rep.loads, rep.evictions, rep.peak_inflight,
         rep.peak_bytes, rep.batched, rep.rounds);
  wrt_free(ctx);
  return EXIT_SUCCESS;
}
