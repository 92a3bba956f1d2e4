/*
 * wrt_internal.h -- private structures of the C host layer.
 */
#ifndef WRT_INTERNAL_H
#define WRT_INTERNAL_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>

#include "../../include/wrt.h"
#include "../csrc/wrt_dev.h"
#include "../csrc/wrt_fast.h"


struct wrt_ctx {
  int device;
  hipStream_t stream;
  hipEvent_t ev0, ev1, ev2, ev3;
  int cu_count;
  /* batch graph cache */
  void* batch;
  /* loads so far on this context: a job's gen is the sequence number of its last load, so
   * a job created at a freed job's address never matches a plan made for the old one */
  unsigned long long load_seq;
  /* spectral Laplacian: FFT plans of the last size (host/wrt_spectral.c) */
  void* spectral;
};
void wrt_spectral_release(wrt_ctx* c);
/* drop the context's batch plan if it lists job j (host/wrt_batch.c) */
void wrt_batch_forget(wrt_ctx* c, const struct wrt_job* j);

struct wrt_job {
  wrt_ctx* ctx;
  wrt_job_desc d;
  /* own stream (default priority) and timing events: e0 start, e1 start of the dominant
   * kernel, e3 its end, e2 end of the job's device work */
  hipStream_t st;
  hipEvent_t e0, e1, e2, e3;
  int pending; /* launched, not yet waited for */
  /* owned host copies of the grids */
  double *phi, *theta, *theta_w, *shift, *z;
  size_t in_count, out_count;
  /* device buffers */
  void* d_in;      /* input, job precision/layout */
  void* d_out;     /* output: double (F64) or float (F32/BF16) */
  double* d_tabs;  /* all small tables in one allocation */
  double* d_shift;
  double* d_z;
  double* d_dirs;  /* radon: [ndir][9]; ray: [nphi][4] */
  double* d_ttrig; /* [ntheta][2] cos, sin */
  double* d_ptrig; /* [nphi][2] cos, sin */
  double* d_wgt;   /* [ntheta] */
  double* d_fac;   /* [2*nshift] */
  double* d_spl;   /* packed spline table */
  void* d_work;    /* fast-path work lists / partials */
  size_t work_bytes;
  void* fast;      /* fast-path plan (wrt_fast.h) */
  int loaded;
  unsigned long long gen; /* ctx->load_seq at the last load (batch plans key on it) */
  size_t bytes;
};

/* a loaded fast job described for the batch planner (host/wrt_fast.c) */
typedef struct {
  wrt_batch_item item;
  const int* brick;           /* radon: non-empty bricks, heaviest first */
  const int* weight;          /* radon: their non-zero voxel counts */
  const unsigned char* bnd;   /* radon: 1 if the brick touches a cube face */
  int nbrick;
  long long* acc;             /* radon: fixed-point plane sums (units of qstep) ... */
  double qstep;
  float* out;                 /* ... and the FP32 output */
  long long nout;
  long long nvox;             /* adjoint: n^3 */
  int dgroup;                 /* radon: direction group of the per-job launch */
} wrt_fast_desc;
int wrt_fast_supported(const wrt_job_desc* d);
int wrt_fast_describe(wrt_job* j, wrt_fast_desc* fd);
/* radon: the band shape and direction group of the per-job launch */
int wrt_fast_tuning(const wrt_job* j, int* band_rows, int* dgroup);
/* radon: the FP64 accumulators rounded to FP32, as doubles (the form's output) */
int wrt_fast_fetch_acc(wrt_job* j, double* out);

/* A task of the persistent launch with its estimated device time in CU-seconds: a full
 * 32^3 brick against one direction over a shift grid of spacing 2/n, and one (voxel,
 * direction) of the FP32 backprojection, measured on MI355X (a 256^3 plane-integral pass,
 * a 256^3 backprojection of 8192 directions).  Only their ratio, the order they induce and
 * the rough length of a serve round depend on them. */
#define WRT_COST_BRICK_DIR 20e-6
#define WRT_COST_VOXEL_DIR 1.1e-10
typedef struct {
  double cost;
  wrt_batch_task t;
} wrt_costed_task;
/* the tasks of item fd (index `item` in its launch; radon: direction groups of G): with
 * ct NULL the count, else written to ct (that many entries), heaviest brick first */
long long wrt_item_tasks(const wrt_fast_desc* fd, int item, int G, wrt_costed_task* ct);
/* order by falling cost (qsort) */
int wrt_costed_desc(const void* a, const void* b);

/* the tuning cache (host/wrt_tune.c) */
int wrt_tune_enabled(void);
unsigned long long wrt_tune_hash(const long long* v, int n);
int wrt_tune_lookup(unsigned long long sig, int* v0, int* v1);
void wrt_tune_store(unsigned long long sig, int v0, int v1, double sec);

/* host-side tables (C library sin/cos, the reference's bits) */
void wrt_host_radon_dirs(int n, const double* phi, int nphi, const double* theta, int ntheta,
                         double* dirs);
void wrt_host_plane_dirs(double delta, const double* phi, int nphi, const double* theta,
                         int ntheta, double* dirs);
void wrt_host_ray_trig(const double* phi, int nphi, double* trig);

int wrt_hip(hipError_t e);
int wrt_check_grid(const double* g, int count, int min_count);

#endif
