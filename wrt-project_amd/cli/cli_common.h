/*
 * cli_common.h -- shared parts of the drop-in command-line programs: the 4-line config
 * parser, the text readers and the option loop.
 *
 * Contract kept from the reference programs:
 *   - short options -h -p <config> -i <input> -o <output> -n <threads> (radon_reduction also
 *     -s: one-column input);
 *   - the config file holds 4 lines "<int> <comment>", each parsed with "%d %[^\n]\n"
 *     (radon_grid/src/init.c:17-82), so a line without trailing text is an error;
 *   - text outputs printed with %lf in the reference's element order, plus readme_<output>.
 * Build-only options that never change the meaning of the reference's: -P f64|f32|bf16
 * (precision, default f64 = the plain form), -d <device>.
 * Long options (--help, --parameters, --input, --output, --nthreads, --small) work here;
 * the reference declares them with a zero flag value, so they never reach its switch.
 */
#ifndef CLI_COMMON_H
#define CLI_COMMON_H

#include <errno.h>
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "wrt.h"

#define CLI_LINE 128

typedef struct {
  const char* params;
  const char* input;
  const char* output;
  int nthreads;
  int small;
  int device;
  wrt_prec prec;
} cli_opts;

static void cli_usage(FILE* f, const char* prog, int with_small) {
  fprintf(f, "Usage: %s -p (filename) -i (filename) -o (filename) -n (number)%s\n", prog,
          with_small ? " [-s]" : "");
  fprintf(f,
          "   -h --help                 Display this usage information.\n"
          "   -p --parameters filename  Read the 4-line grid configuration from file.\n"
          "   -i --input filename       Read input data from file.\n"
          "   -o --output filename      Write output data to the file.\n"
          "   -n --nthreads number      Accepted for compatibility (work runs on the GPU).\n");
  if (with_small) fprintf(f, "   -s --small                Input holds one value per line.\n");
  fprintf(f,
          "   -P --precision f64|f32|bf16  Arithmetic (default f64, the plain form).\n"
          "   -d --device number        HIP device (default 0).\n");
}

static int cli_parse(int argc, char** argv, cli_opts* o, int with_small) {
  memset(o, 0, sizeof(*o));
  o->nthreads = 1;
  o->prec = WRT_F64;
  if (argc == 1) {
    fprintf(stderr, "%s: arguments required. Try '-h' or '--help'.\n", argv[0]);
    exit(EXIT_FAILURE);
  }
  const struct option longs[] = {{"help", no_argument, 0, 'h'},
                                 {"parameters", required_argument, 0, 'p'},
                                 {"input", required_argument, 0, 'i'},
                                 {"output", required_argument, 0, 'o'},
                                 {"nthreads", required_argument, 0, 'n'},
                                 {"small", no_argument, 0, 's'},
                                 {"precision", required_argument, 0, 'P'},
                                 {"device", required_argument, 0, 'd'},
                                 {0, 0, 0, 0}};
  const char* shorts = with_small ? "hp:i:so:n:P:d:" : "hp:i:o:n:P:d:";
  int c;
  while ((c = getopt_long(argc, argv, shorts, longs, NULL)) != -1) {
    switch (c) {
      case 'h': cli_usage(stdout, argv[0], with_small); exit(EXIT_SUCCESS);
      case 'p': o->params = optarg; printf("Parameters file: %s\n", optarg); break;
      case 'i': o->input = optarg; printf("Input file: %s\n", optarg); break;
      case 'o': o->output = optarg; printf("Output file: %s\n", optarg); break;
      case 'n':
        o->nthreads = atoi(optarg);
        if (o->nthreads <= 0) {
          fprintf(stderr, "%s: the number of threads must be positive\n", argv[0]);
          exit(EXIT_FAILURE);
        }
        printf("Number of threads: %d\n", o->nthreads);
        break;
      case 's':
        if (!with_small) { cli_usage(stderr, argv[0], with_small); exit(EXIT_FAILURE); }
        o->small = 1;
        printf("Reading of the input in small format\n");
        break;
      case 'P':
        if (!strcmp(optarg, "f64")) o->prec = WRT_F64;
        else if (!strcmp(optarg, "f32")) o->prec = WRT_F32;
        else if (!strcmp(optarg, "bf16")) o->prec = WRT_BF16;
        else { fprintf(stderr, "%s: unknown precision %s\n", argv[0], optarg); exit(EXIT_FAILURE); }
        break;
      case 'd': o->device = atoi(optarg); break;
      default: cli_usage(stderr, argv[0], with_small); exit(EXIT_FAILURE);
    }
  }
  if (optind < argc) printf("Warning: non-option ARGV-arguments are omitted. See '-h' or '--help'.\n");
  if (!o->params || !o->input || !o->output) {
    fprintf(stderr, "%s: -p, -i and -o are required\n", argv[0]);
    exit(EXIT_FAILURE);
  }
  return 0;
}

/* 4 lines "%d %[^\n]\n": returns 0, errno (open failure) or -1 (parse failure) */
static int cli_params(const char* fn, int p[4]) {
  FILE* f = fopen(fn, "r");
  if (!f) {
    fprintf(stderr, "Recieved call %s while opening %s. \n", strerror(errno), fn);
    return errno;
  }
  char buf[CLI_LINE * 8];
  for (int k = 0; k < 4; ++k) {
    int cnt = fscanf(f, "%d %1023[^\n]\n", &p[k], buf);
    if (cnt == EOF) {
      fprintf(stderr, "Error: fscanf parameters %d : reached end of file before it was expected. \n", k);
      fclose(f);
      return -1;
    } else if (cnt != 2) {
      fprintf(stderr, "Error: fscanf parameters %d : matching failure, fscanf expected 1 integer value.\n", k);
      fclose(f);
      return -1;
    }
  }
  fclose(f);
  return 0;
}

/* count values "%lf\n" (cols = 1) or "%lf, %lf, %lf, %lf\n" keeping the last (cols = 4) */
static int cli_read_values(const char* fn, size_t count, int cols, double* out) {
  FILE* f = fopen(fn, "r");
  if (!f) {
    fprintf(stderr, "Recieved call %s while opening %s. \n", strerror(errno), fn);
    return errno;
  }
  for (size_t i = 0; i < count; ++i) {
    int cnt;
    if (cols == 1) {
      cnt = fscanf(f, "%lf\n", &out[i]);
    } else {
      double a, b, c;
      cnt = fscanf(f, "%lf, %lf, %lf, %lf\n", &a, &b, &c, &out[i]);
    }
    if (cnt == EOF) {
      fprintf(stderr, "Error : fscanf : reached the end of file before it was expected.\n");
      fclose(f);
      return -1;
    } else if (cnt != cols) {
      fprintf(stderr, "Error: fscanf : matching failure, expected %d floating-point value(s).\n", cols);
      fclose(f);
      return -1;
    }
  }
  fclose(f);
  return 0;
}

static double cli_now(void) {
  struct timeval t;
  gettimeofday(&t, NULL);
  return t.tv_sec + t.tv_usec * 1e-6;
}

static void cli_die(const char* what, int rc) {
  fprintf(stderr, "Aborting : %s : %s (%d)\n", what, wrt_strerror(rc), rc);
  exit(EXIT_FAILURE);
}

static void* cli_alloc(size_t bytes) {
  void* p = malloc(bytes ? bytes : 1);
  if (!p) {
    fprintf(stderr, "Aborting : out of host memory (%zu bytes)\n", bytes);
    exit(EXIT_FAILURE);
  }
  return p;
}

static void cli_readme(const char* out, const char* what, const char* format, const char* const* names,
                       const int* vals, int nv) {
  char fn[4096];
  snprintf(fn, sizeof(fn), "readme_%s", out);
  FILE* f = fopen(fn, "w");
  if (!f) return;
  fprintf(f, "  %s for data given in %s\n", what, out);
  fprintf(f, "  Format of the data: %s\n", format);
  fprintf(f, "  Parameters of the data:\n");
  for (int k = 0; k < nv; ++k) fprintf(f, "    %s: %d%s\n", names[k], vals[k], k + 1 < nv ? "," : ".");
  fclose(f);
}

#endif
