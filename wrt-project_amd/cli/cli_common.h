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
 * (precision, default f64 = the plain form), -d <device>, -b / -B (input / output in the
 * library's binary format, the value column only).  Text is read and written through the
 * library's parallel codecs (wrt_read_text, wrt_write_text_*), with the same values and bytes
 * as fscanf / fprintf.
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
  int bin_in, bin_out; /* binary input / output (wrt_read_f64 / wrt_write_f64 format) */
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
          "   -b --binary-in            Input in the library's binary format (values only).\n"
          "   -B --binary-out           Output in the library's binary format (values only).\n"
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
                                 {"binary-in", no_argument, 0, 'b'},
                                 {"binary-out", no_argument, 0, 'B'},
                                 {0, 0, 0, 0}};
  const char* shorts = with_small ? "hp:i:so:n:P:d:bB" : "hp:i:o:n:P:d:bB";
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
      case 'b': o->bin_in = 1; break;
      case 'B': o->bin_out = 1; break;
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

/* count lines "%d %[^\n]\n": returns 0, errno (open failure) or -1 (parse failure) */
static int cli_params_n(const char* fn, int* p, int count) {
  FILE* f = fopen(fn, "r");
  if (!f) {
    fprintf(stderr, "Recieved call %s while opening %s. \n", strerror(errno), fn);
    return errno;
  }
  char buf[CLI_LINE * 8];
  for (int k = 0; k < count; ++k) {
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

/* the grid programs' 4-line config */
static int cli_params(const char* fn, int p[4]) { return cli_params_n(fn, p, 4); }

/* count values "%lf\n" (cols = 1) or "%lf, %lf, %lf, %lf\n" keeping the last (cols = 4),
 * or the binary format with -b; 0 or -1 */
static int cli_read_input(const cli_opts* o, size_t count, int cols, double* out) {
  const int rc = o->bin_in ? wrt_read_f64(o->input, count, out) : wrt_read_text(o->input, count, cols, out);
  return rc == WRT_OK ? 0 : -1;
}

/* the output: 4-column text rows over the grid (a, b, c), 1-column text, or binary (-B) */
static void cli_write_output(int bin, const char* path, const double* a, int na, const double* b,
                             int nb, const double* c, int nc, const double* values, size_t count) {
  int rc;
  if (bin) rc = wrt_write_f64(path, values, count);
  else if (a) rc = wrt_write_text_grid(path, a, na, b, nb, c, nc, values);
  else rc = wrt_write_text_values(path, values, count);
  if (rc != WRT_OK) {
    fprintf(stderr, "Aborting : cannot write %s : %s\n", path, wrt_strerror(rc));
    exit(EXIT_FAILURE);
  }
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

/* ------------------------------------------------------------ analytic programs --- */
/* Options of radon_analytic / ray_analytic: the reference's -h -p -o -n (no input file: the
 * reference compiles its test function in), plus the test function, which here is chosen at
 * run time: -F <example name> (the reference's example files) or -f <term file>. */
typedef struct {
  const char* params;
  const char* output;
  const char* example;
  const char* terms_file;
  int nthreads;
  int device;
  int bin_out;
} cli_an_opts;

static void cli_an_usage(FILE* f, const char* prog) {
  fprintf(f, "Usage: %s -p (filename) -o (filename) -n (number) (-F name | -f filename)\n", prog);
  fprintf(f,
          "   -h --help                 Display this usage information.\n"
          "   -p --parameters filename  Read the grid configuration from file.\n"
          "   -o --output filename      Write output data to the file.\n"
          "   -n --nthreads number      Accepted for compatibility (work runs on the GPU).\n"
          "   -F --function name        Test function: exponential_bump or sinusoidal_bump\n"
          "                             (the reference's example test functions).\n"
          "   -f --terms filename       Test function as a sum of terms, one per line:\n"
          "                               exp_bump  cx cy cz r2max eps\n"
          "                               sine_bump cx cy cz amplitude rin rout\n"
          "                               ellipsoid cx cy cz density a1 a2 a3 r11 r12 ... r33\n"
          "                             ('#' starts a comment).\n"
          "   -B --binary-out           Output in the library's binary format (values only).\n"
          "   -d --device number        HIP device (default 0).\n");
}

static void cli_an_parse(int argc, char** argv, cli_an_opts* o) {
  memset(o, 0, sizeof(*o));
  o->nthreads = 1;
  if (argc == 1) {
    fprintf(stderr, "%s: arguments required. Try '-h' or '--help'.\n", argv[0]);
    exit(EXIT_FAILURE);
  }
  const struct option longs[] = {{"help", no_argument, 0, 'h'},
                                 {"parameters", required_argument, 0, 'p'},
                                 {"output", required_argument, 0, 'o'},
                                 {"nthreads", required_argument, 0, 'n'},
                                 {"function", required_argument, 0, 'F'},
                                 {"terms", required_argument, 0, 'f'},
                                 {"device", required_argument, 0, 'd'},
                                 {"binary-out", no_argument, 0, 'B'},
                                 {0, 0, 0, 0}};
  int c;
  while ((c = getopt_long(argc, argv, "hp:o:n:F:f:d:B", longs, NULL)) != -1) {
    switch (c) {
      case 'h': cli_an_usage(stdout, argv[0]); exit(EXIT_SUCCESS);
      case 'p': o->params = optarg; printf("Parameters file: %s\n", optarg); break;
      case 'o': o->output = optarg; printf("Output file: %s\n", optarg); break;
      case 'n':
        o->nthreads = atoi(optarg);
        if (o->nthreads <= 0) {
          fprintf(stderr, "%s: the number of threads must be positive\n", argv[0]);
          exit(EXIT_FAILURE);
        }
        printf("Number of threads: %d\n", o->nthreads);
        break;
      case 'F': o->example = optarg; break;
      case 'f': o->terms_file = optarg; break;
      case 'd': o->device = atoi(optarg); break;
      case 'B': o->bin_out = 1; break;
      default: cli_an_usage(stderr, argv[0]); exit(EXIT_FAILURE);
    }
  }
  if (optind < argc) printf("Warning: non-option ARGV-arguments are omitted. See '-h' or '--help'.\n");
  if (!o->params || !o->output || (!o->example == !o->terms_file)) {
    fprintf(stderr, "%s: -p, -o and one of -F / -f are required\n", argv[0]);
    exit(EXIT_FAILURE);
  }
}

/* a term file (format in cli_an_usage): returns the number of terms or -1 */
static int cli_read_terms(const char* fn, wrt_fn_term* t, int max) {
  FILE* f = fopen(fn, "r");
  if (!f) {
    fprintf(stderr, "Recieved call %s while opening %s. \n", strerror(errno), fn);
    return -1;
  }
  char line[1024];
  int n = 0, lineno = 0;
  while (fgets(line, sizeof(line), f)) {
    ++lineno;
    char* h = strchr(line, '#');
    if (h) *h = 0;
    char kind[32];
    int used = 0;
    if (sscanf(line, " %31s%n", kind, &used) != 1) continue; /* blank */
    int want;
    wrt_fn_term* q = t + n;
    memset(q, 0, sizeof(*q));
    if (!strcmp(kind, "exp_bump")) { q->kind = WRT_FN_EXP_BUMP; want = 5; }
    else if (!strcmp(kind, "sine_bump")) { q->kind = WRT_FN_SINE_BUMP; want = 6; }
    else if (!strcmp(kind, "ellipsoid")) { q->kind = WRT_FN_ELLIPSOID; want = 16; }
    else { fprintf(stderr, "%s:%d: unknown term '%s'\n", fn, lineno, kind); fclose(f); return -1; }
    if (n >= max) { fprintf(stderr, "%s: more than %d terms\n", fn, max); fclose(f); return -1; }
    double v[16];
    char* s = line + used;
    int got = 0;
    for (; got < want; ++got) {
      int u = 0;
      if (sscanf(s, " %lf%n", &v[got], &u) != 1) break;
      s += u;
    }
    if (got != want) {
      fprintf(stderr, "%s:%d: '%s' takes %d numbers\n", fn, lineno, kind, want);
      fclose(f);
      return -1;
    }
    for (int k = 0; k < 3; ++k) q->c[k] = v[k];
    for (int k = 3; k < want; ++k) q->p[k - 3] = v[k];
    ++n;
  }
  fclose(f);
  return n;
}

static int cli_an_terms(const cli_an_opts* o, wrt_fn_term* t, int max) {
  const int n = o->example ? wrt_fn_example(o->example, t, max) : cli_read_terms(o->terms_file, t, max);
  if (n < 0) {
    fprintf(stderr, "Aborting : unknown or malformed test function %s\n",
            o->example ? o->example : o->terms_file);
    exit(EXIT_FAILURE);
  }
  printf("Test function: %s (%d term%s)\n", o->example ? o->example : o->terms_file, n, n == 1 ? "" : "s");
  return n;
}

#endif
