/*
 * wrt_serve.c -- the request loop: Poisson arrivals, admission by live device memory,
 * resident datasets, continuous batching of the fast forms into rounds of the persistent
 * brick/voxel kernel, latency percentiles.
 *
 * One host thread drives everything.  Each loop pass
 *   (1) moves arrived requests to the pending list;
 *   (2) records finished work: rounds whose completion event has fired, and with them every
 *       request whose last task lay in a round no later than all rounds done so far; and
 *       the requests running on their own streams (below);
 *   (3) starts pending requests in arrival order: a request whose dataset is not resident
 *       first makes it resident (create + load) when hipMemGetInfo shows room for it (and
 *       the optional byte budget allows), evicting idle datasets least recently used
 *       first; a request whose dataset is busy waits for it.  A fast request (FP32 plane
 *       integrals, FP32 backprojection) joins the active list with its task list (bricks x
 *       direction groups, or voxel blocks), heaviest first; any other (FP64 reduction,
 *       plain mode) runs on its dataset's own stream, up to max_inflight at once;
 *   (4) while one of two round slots is idle and active requests have tasks left, builds a
 *       round of about RS_ROUND_S of device time: first every small request whole (its
 *       tasks at most a quarter of the round), then chunks of the others in arrival order,
 *       so that a request arriving while a 256^3 request runs starts within a round or
 *       two instead of queueing behind it.  A round is one launch of k_batch (the batch
 *       runner's kernel) over its tasks, after a prologue that zeroes the FP64
 *       accumulators of requests starting in it.  The two slots use two streams, so the
 *       next round fills the CUs the current one frees at its end; a round waits for the
 *       other slot's prologue, so no task of a request runs before its accumulators are
 *       zeroed.  A served plane-integral request's output is its FP64 accumulators rounded
 *       to FP32 on the host, the rounding the per-job launch does on the device.
 * Nothing here waits on the device except a dataset load and an output fetch.
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "wrt_internal.h"

static unsigned long long mix64(unsigned long long z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

/* uniform in (0, 1) from a counter */
static double unif(unsigned long long seed, unsigned long long ctr) {
  return ((double)(mix64(seed ^ mix64(ctr)) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

static int cfg_ok(const wrt_serve_cfg* c) {
  if (!c || c->nreq < 0 || !c->edges || c->nedges < 1 || c->ndatasets < 1 || c->max_inflight < 1)
    return 0;
  if (!(c->mix[0] >= 0 && c->mix[1] >= 0 && c->mix[2] >= 0) || c->mix[0] + c->mix[1] + c->mix[2] <= 0)
    return 0;
  for (int i = 0; i < c->nedges; ++i)
    if (c->edges[i] < 8 || (c->edges[i] & 1) || c->edges[i] > 1024) return 0;
  return 1;
}

int wrt_serve_describe(const wrt_serve_cfg* c, int i, wrt_serve_req* r) {
  if (!cfg_ok(c) || !r || i < 0 || i >= c->nreq) return WRT_E_ARG;
  /* arrival i = sum of i+1 exponential gaps; each gap comes from its own counter, so
   * request i does not depend on how many were drawn before it */
  double t = 0.0;
  if (c->qps > 0)
    for (int k = 0; k <= i; ++k) t += -log(unif(c->seed, 4ULL * k)) / c->qps;
  const double u = unif(c->seed, 4ULL * i + 1) * (c->mix[0] + c->mix[1] + c->mix[2]);
  r->op = u < c->mix[0] ? WRT_OP_RADON : (u < c->mix[0] + c->mix[1] ? WRT_OP_ADJOINT : WRT_OP_REDUCTION);
  r->prec = (r->op == WRT_OP_REDUCTION || c->plain) ? WRT_F64 : WRT_F32;
  int e = (int)(unif(c->seed, 4ULL * i + 2) * c->nedges);
  r->edge = c->edges[e < c->nedges ? e : c->nedges - 1];
  int ds = (int)(unif(c->seed, 4ULL * i + 3) * c->ndatasets);
  r->dataset = ds < c->ndatasets ? ds : c->ndatasets - 1;
  r->data_seed = mix64(c->seed ^ mix64(((unsigned long long)r->op << 40) ^
                                        ((unsigned long long)r->edge << 20) ^ (unsigned long long)r->dataset));
  r->arrival_s = t;
  return WRT_OK;
}

int wrt_serve_grids(wrt_ctx* ctx, int e, double* phi, double* theta, double* theta_w,
                    double* shift, double* z) {
  if (!ctx || e < 8 || (e & 1) || !phi || !theta || !theta_w || !shift || !z) return WRT_E_ARG;
  const int rc = wrt_radon_grid_init(ctx, e / 2, e / 4, e + 1, phi, theta, theta_w, shift);
  if (rc != WRT_OK) return rc;
  return wrt_zslice_grid(e, z);
}

int wrt_serve_desc(const wrt_serve_req* r, const double* phi, const double* theta,
                   const double* theta_w, const double* shift, const double* z,
                   wrt_job_desc* d) {
  if (!r || !d) return WRT_E_ARG;
  const int e = r->edge;
  memset(d, 0, sizeof(*d));
  d->op = r->op;
  d->prec = r->prec;
  d->phi = phi;
  d->nphi = e / 2;
  d->shift = shift;
  d->nshift = e + 1;
  d->theta = theta;
  d->ntheta = e / 4;
  if (r->op == WRT_OP_REDUCTION) {
    d->z = z;
    d->nz = e;
  } else {
    d->n = e;
    if (r->op == WRT_OP_ADJOINT) d->theta_w = theta_w;
  }
  return WRT_OK;
}

int wrt_serve_input(const wrt_serve_req* r, double* in) {
  if (!r || !in) return WRT_E_ARG;
  const size_t e = (size_t)r->edge;
  switch (r->op) {
    case WRT_OP_RADON: return wrt_phantom(r->edge, 1, 10, r->data_seed, 0.01, in);
    case WRT_OP_ADJOINT: return wrt_gaussian((e + 1) * (e / 2) * (e / 4), r->data_seed, in);
    case WRT_OP_REDUCTION: return wrt_gaussian(e * (e + 1) * (e / 2), r->data_seed, in);
    default: return WRT_E_ARG;
  }
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void sleep_s(double s) {
  if (s <= 0) return;
  struct timespec ts;
  ts.tv_sec = (time_t)s;
  ts.tv_nsec = (long)((s - (double)ts.tv_sec) * 1e9);
  nanosleep(&ts, NULL);
}

typedef struct {
  wrt_op op;
  int edge, dataset, grid;
  wrt_serve_req req;  /* a request of this dataset (for its description and input) */
  wrt_job_desc desc;
  double* host_in;
  wrt_job* job;
  size_t bytes;
  int busy;            /* request running on it, or -1 */
  double last_used;
} dataset_t;

typedef struct {
  int edge;
  double *phi, *theta, *w, *shift, *z;
} grid_t;

static int cmp_double(const void* a, const void* b) {
  const double x = *(const double*)a, y = *(const double*)b;
  return x < y ? -1 : (x > y);
}

/* nearest-rank percentile of sorted values */
static double pct(const double* v, int n, double p) {
  if (n <= 0) return 0;
  int k = (int)ceil(p / 100.0 * n) - 1;
  if (k < 0) k = 0;
  if (k >= n) k = n - 1;
  return v[k];
}

/* ------------------------------------------------------------------ rounds --- */
#define RS_MAIN 2              /* slots of the main rounds */
#define RS_SLOTS 3             /* ... and one for express rounds of small requests */
#define RS_MAX_ITEMS 64        /* requests in one round */
#define RS_MAX_TASKS 65536     /* tasks in one round */
#define RS_ROUND_S 8e-3        /* device time a main round aims at (the whole GPU); the
                                  environment's WRT_SERVE_ROUND_MS overrides it */
#define RS_GROUP 128           /* directions per plane-integral task at most (shorter tasks
                                  keep rounds short; WRT_SERVE_GROUP overrides it) */
#define RS_MIN_WAVES 1         /* tasks per CU a main round holds at least (WRT_SERVE_WAVES) */
#define RS_MEDIUM_S 25e-3      /* a request with at most this much device time left goes
                                  before the larger ones ... */
#define RS_AGE_S 0.15          /* ... unless they arrived this long ago */

/* a fast request on the active list */
typedef struct {
  int req, ds;
  double arrival;         /* seconds into the stream */
  wrt_fast_desc fd;
  wrt_costed_task* tasks; /* heaviest first */
  long long ntasks, next;
  double left;            /* estimated CU-seconds not yet scheduled */
  int rounds;             /* rounds holding its tasks that have not finished */
  int G;                  /* radon: directions per task */
  int zeroed;             /* accumulators zeroed by a scheduled prologue */
} active_t;

typedef struct {
  hipStream_t st;
  hipEvent_t pro, done;
  wrt_batch_item *d_items, *h_items;
  wrt_batch_task *d_tasks, *h_tasks;
  wrt_batch_seg *d_segs, *h_segs;
  int* d_counter;
  wrt_costed_task* scratch; /* the round's tasks before sorting */
  int busy;                 /* a round runs in this slot */
  int nreq, req[RS_MAX_ITEMS]; /* the requests it holds */
} round_slot;

static void slot_free(round_slot* r) {
  if (r->st) hipStreamSynchronize(r->st);
  if (r->pro) hipEventDestroy(r->pro);
  if (r->done) hipEventDestroy(r->done);
  if (r->st) hipStreamDestroy(r->st);
  hipFree(r->d_items);
  hipFree(r->d_tasks);
  hipFree(r->d_segs);
  hipFree(r->d_counter);
  hipHostFree(r->h_items);
  hipHostFree(r->h_tasks);
  hipHostFree(r->h_segs);
  free(r->scratch);
  memset(r, 0, sizeof(*r));
}

static int slot_init(round_slot* r, int cu_count) {
  memset(r, 0, sizeof(*r));
  (void)cu_count;
  if (hipStreamCreateWithFlags(&r->st, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&r->pro, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&r->done, hipEventDisableTiming) != hipSuccess)
    return WRT_E_HIP;
  if (hipMalloc((void**)&r->d_items, sizeof(wrt_batch_item) * RS_MAX_ITEMS) != hipSuccess ||
      hipMalloc((void**)&r->d_tasks, sizeof(wrt_batch_task) * RS_MAX_TASKS) != hipSuccess ||
      hipMalloc((void**)&r->d_segs, sizeof(wrt_batch_seg) * RS_MAX_ITEMS) != hipSuccess ||
      hipMalloc((void**)&r->d_counter, 256) != hipSuccess ||
      hipHostMalloc((void**)&r->h_items, sizeof(wrt_batch_item) * RS_MAX_ITEMS, 0) != hipSuccess ||
      hipHostMalloc((void**)&r->h_tasks, sizeof(wrt_batch_task) * RS_MAX_TASKS, 0) != hipSuccess ||
      hipHostMalloc((void**)&r->h_segs, sizeof(wrt_batch_seg) * RS_MAX_ITEMS, 0) != hipSuccess)
    return WRT_E_NOMEM;
  r->scratch = (wrt_costed_task*)malloc(sizeof(wrt_costed_task) * RS_MAX_TASKS);
  return r->scratch ? WRT_OK : WRT_E_NOMEM;
}

/* Build a round in slot `me` of `slots` from the active list (arrival order) and launch it;
 * nothing is launched (busy stays 0) when there is nothing to schedule.  A main round takes:
 * every small request (the rest of its work fits in a round) whole, up to twice the budget
 * in all; then the budget left to the others in arrival order: large ones (more than
 * RS_MEDIUM_S of the whole GPU left) that arrived over RS_AGE_S ago, then medium ones, then
 * the other large ones, so that a 128^3 or 192^3 request is not held up behind a 256^3 one,
 * while a large one is not put off for long; at least one task each and, round robin, until
 * the round holds min_tasks tasks.
 * An express round takes only the small requests, whole.  The round waits for every other
 * running round's prologue, so no task runs before its request's accumulators are zeroed. */
static int round_launch(wrt_ctx* ctx, round_slot* slots, int me, int express, active_t* act,
                        int nact, double round_s, int min_waves, double now) {
  round_slot* r = &slots[me];
  const double budget = round_s * ctx->cu_count;
  double left = budget;
  const long long min_tasks = express ? 0 : (long long)min_waves * ctx->cu_count;
  const double medium = RS_MEDIUM_S * ctx->cu_count;
  int sel[RS_MAX_ITEMS];
  int nsel = 0, nseg = 0;
  long long nt = 0;
  for (int i = 0; i < nact && nsel < RS_MAX_ITEMS; ++i) {
    const active_t* a = &act[i];
    if (a->next >= a->ntasks) continue;
    if (a->left <= budget && left - a->left >= -budget) {
      sel[nsel++] = i;
      left -= a->left;
    }
  }
  const int nsmall = nsel;
  if (!express)
    for (int cls = 0; cls < 3; ++cls)
      for (int i = 0; i < nact && nsel < RS_MAX_ITEMS; ++i) {
        const active_t* a = &act[i];
        int chosen = 0;
        for (int q = 0; q < nsmall; ++q) chosen |= sel[q] == i;
        if (a->next >= a->ntasks || chosen) continue;
        const int c = a->left <= medium ? 1 : (now - a->arrival > RS_AGE_S ? 0 : 2);
        if (c == cls) sel[nsel++] = i;
      }
  for (int q = 0; q < nsel && nt < RS_MAX_TASKS; ++q) {
    active_t* a = &act[sel[q]];
    long long took = 0;
    while (a->next < a->ntasks && nt < RS_MAX_TASKS && (q < nsmall || left > 0 || took == 0)) {
      wrt_costed_task t = a->tasks[a->next++];
      t.t.item = q;
      r->scratch[nt++] = t;
      if (q >= nsmall) left -= t.cost;
      a->left -= t.cost;
      ++took;
    }
  }
  for (int more = 1; more && nt < min_tasks;) { /* fill the GPU: one more task each */
    more = 0;
    for (int q = nsmall; q < nsel && nt < min_tasks && nt < RS_MAX_TASKS; ++q) {
      active_t* a = &act[sel[q]];
      if (a->next >= a->ntasks) continue;
      wrt_costed_task t = a->tasks[a->next++];
      t.t.item = q;
      r->scratch[nt++] = t;
      a->left -= t.cost;
      more = 1;
    }
  }
  if (nt == 0) return WRT_OK;
  r->nreq = nsel;
  for (int q = 0; q < nsel; ++q) {
    active_t* a = &act[sel[q]];
    wrt_batch_item it = a->fd.item;
    if (it.kind == WRT_ITEM_RADON) {
      it.ra.dgroup = a->G;
      if (!a->zeroed) {
        r->h_segs[nseg].acc = a->fd.acc;
        r->h_segs[nseg].qstep = a->fd.qstep;
        r->h_segs[nseg].out = a->fd.out;
        r->h_segs[nseg].count = a->fd.nout;
        ++nseg;
        a->zeroed = 1;
      }
    }
    r->h_items[q] = it;
    r->req[q] = a->req;
    a->rounds++;
  }
  /* heaviest first across the requests, so that the persistent workgroups end together */
  qsort(r->scratch, (size_t)nt, sizeof(wrt_costed_task), wrt_costed_desc);
  for (long long q = 0; q < nt; ++q) r->h_tasks[q] = r->scratch[q].t;
  hipStream_t st = r->st;
  if (hipMemcpyAsync(r->d_items, r->h_items, sizeof(wrt_batch_item) * (size_t)nsel,
                     hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(r->d_tasks, r->h_tasks, sizeof(wrt_batch_task) * (size_t)nt,
                     hipMemcpyHostToDevice, st) != hipSuccess ||
      (nseg > 0 && hipMemcpyAsync(r->d_segs, r->h_segs, sizeof(wrt_batch_seg) * (size_t)nseg,
                                  hipMemcpyHostToDevice, st) != hipSuccess))
    return WRT_E_HIP;
  for (int q = 0; q < RS_SLOTS; ++q)
    if (q != me && slots[q].busy && hipStreamWaitEvent(st, slots[q].pro, 0) != hipSuccess)
      return WRT_E_HIP;
  int rc = wrtk_batch_prologue(r->d_segs, nseg, r->d_counter, st);
  if (rc != WRT_OK) return rc;
  if (hipEventRecord(r->pro, st) != hipSuccess) return WRT_E_HIP;
  rc = wrtk_batch_run(r->d_items, r->d_tasks, (int)nt, r->d_counter, ctx->cu_count, st);
  if (rc != WRT_OK) return rc;
  if (hipEventRecord(r->done, st) != hipSuccess) return WRT_E_HIP;
  r->busy = 1;
  return WRT_OK;
}

/* a request that runs through the rounds: the fast forms the persistent kernel serves */
static int batched(const wrt_job_desc* d) {
  return d->prec == WRT_F32 && wrt_fast_supported(d) &&
         (d->op == WRT_OP_RADON || (d->op == WRT_OP_ADJOINT && d->n <= 1024));
}

int wrt_serve_run(wrt_ctx* ctx, const wrt_serve_cfg* c, wrt_serve_report* rep,
                  double* latency_ms, double* const* outputs) {
  if (!ctx || !cfg_ok(c) || !rep) return WRT_E_ARG;
  memset(rep, 0, sizeof(*rep));
  const int nreq = c->nreq;
  if (nreq == 0) return WRT_OK;
  if (hipSetDevice(ctx->device) != hipSuccess) return WRT_E_HIP;
  int rc = WRT_OK;
  wrt_serve_req* reqs = (wrt_serve_req*)calloc((size_t)nreq, sizeof(*reqs));
  int* req_ds = (int*)calloc((size_t)nreq, sizeof(int));
  int* pending = (int*)calloc((size_t)nreq, sizeof(int));
  int* running = (int*)calloc((size_t)nreq, sizeof(int));
  double* lat = (double*)calloc((size_t)nreq, sizeof(double));
  active_t* act = (active_t*)calloc((size_t)nreq, sizeof(active_t));
  const int maxds = 3 * c->nedges * c->ndatasets;
  dataset_t* ds = (dataset_t*)calloc((size_t)maxds, sizeof(dataset_t));
  grid_t* grids = (grid_t*)calloc((size_t)c->nedges, sizeof(grid_t));
  round_slot slots[RS_SLOTS];
  memset(slots, 0, sizeof(slots));
  int nds = 0, ngrids = 0, nact = 0;
  if (!reqs || !req_ds || !pending || !running || !lat || !act || !ds || !grids) {
    rc = WRT_E_NOMEM;
    goto done;
  }

  /* the stream, its grids and datasets; the inputs (the clients' payloads) exist before
   * the stream starts */
  for (int i = 0; i < nreq && rc == WRT_OK; ++i) {
    rc = wrt_serve_describe(c, i, &reqs[i]);
    if (rc != WRT_OK) break;
    int g = 0;
    while (g < ngrids && grids[g].edge != reqs[i].edge) ++g;
    if (g == ngrids) {
      const int e = reqs[i].edge;
      grid_t* G = &grids[ngrids];
      G->edge = e;
      G->phi = (double*)malloc(sizeof(double) * (size_t)(e / 2));
      G->theta = (double*)malloc(sizeof(double) * (size_t)(e / 4));
      G->w = (double*)malloc(sizeof(double) * (size_t)(e / 4));
      G->shift = (double*)malloc(sizeof(double) * (size_t)(e + 1));
      G->z = (double*)malloc(sizeof(double) * (size_t)e);
      ++ngrids;
      if (!G->phi || !G->theta || !G->w || !G->shift || !G->z) { rc = WRT_E_NOMEM; break; }
      rc = wrt_serve_grids(ctx, e, G->phi, G->theta, G->w, G->shift, G->z);
      if (rc != WRT_OK) break;
    }
    int k = 0;
    while (k < nds && !(ds[k].op == reqs[i].op && ds[k].edge == reqs[i].edge &&
                        ds[k].dataset == reqs[i].dataset))
      ++k;
    if (k == nds) {
      dataset_t* D = &ds[nds++];
      D->op = reqs[i].op;
      D->edge = reqs[i].edge;
      D->dataset = reqs[i].dataset;
      D->grid = g;
      D->req = reqs[i];
      D->busy = -1;
      const grid_t* G = &grids[g];
      wrt_serve_desc(&reqs[i], G->phi, G->theta, G->w, G->shift, G->z, &D->desc);
      D->bytes = wrt_job_bytes(&D->desc);
      const size_t cnt = wrt_job_input_count(&D->desc);
      /* pinned, so that a load's copy runs at full speed */
      if (hipHostMalloc((void**)&D->host_in, sizeof(double) * cnt, 0) != hipSuccess) D->host_in = NULL;
      if (!D->host_in || D->bytes == 0) { rc = D->host_in ? WRT_E_ARG : WRT_E_NOMEM; break; }
      rc = wrt_serve_input(&reqs[i], D->host_in);
    }
    req_ds[i] = k;
  }
  if (rc != WRT_OK) goto done;
  for (int q = 0; q < RS_SLOTS && rc == WRT_OK; ++q) rc = slot_init(&slots[q], ctx->cu_count);
  if (rc != WRT_OK) goto done;

  /* admission: live free device memory less a reserve, and the optional byte budget */
  const size_t reserve = (size_t)1 << 30;
  const size_t budget = c->hbm_budget ? c->hbm_budget : (size_t)-1;
  {
    size_t fr = 0, tot = 0;
    rc = wrt_mem_info(ctx, &fr, &tot);
    if (rc != WRT_OK) goto done;
    for (int k = 0; k < nds; ++k)
      if (ds[k].bytes > budget || ds[k].bytes + reserve > fr) { rc = WRT_E_ADMIT; goto done; }
  }

  double round_s = RS_ROUND_S;
  int group_cap = RS_GROUP, min_waves = RS_MIN_WAVES;
  {
    const char* e = getenv("WRT_SERVE_ROUND_MS");
    if (e && atof(e) > 0) round_s = atof(e) * 1e-3;
    e = getenv("WRT_SERVE_GROUP");
    if (e && atoi(e) >= 16) group_cap = atoi(e);
    e = getenv("WRT_SERVE_WAVES");
    if (e && atoi(e) >= 1) min_waves = atoi(e);
  }
  size_t used = 0;
  int npend = 0, nrun = 0, next = 0, completed = 0;
  long long nrounds = 0;
  double samples = 0.0;
  const double t0 = now_s();
  while (completed < nreq) {
    const double t = now_s() - t0;
    int progress = 0;
    while (next < nreq && reqs[next].arrival_s <= t) pending[npend++] = next++;

    /* finished requests on their own streams */
    for (int q = 0; q < nrun;) {
      const int r = running[q];
      dataset_t* D = &ds[req_ds[r]];
      wrt_stats st;
      const int p = wrt_job_poll(D->job, &st);
      if (p == 1) { ++q; continue; }
      if (p != WRT_OK) { rc = p; goto done; }
      const double tdone = now_s() - t0;
      lat[r] = (tdone - reqs[r].arrival_s) * 1e3;
      samples += st.logical_samples;
      if (outputs && outputs[r]) {
        rc = wrt_job_fetch(D->job, outputs[r]);
        if (rc != WRT_OK) goto done;
      }
      D->busy = -1;
      D->last_used = tdone;
      running[q] = running[--nrun];
      ++completed;
      progress = 1;
    }

    /* finished rounds, then the requests whose every round is done */
    for (int q = 0; q < RS_SLOTS; ++q) {
      if (!slots[q].busy) continue;
      const hipError_t e = hipEventQuery(slots[q].done);
      if (e == hipErrorNotReady) continue;
      if (e != hipSuccess) { rc = WRT_E_HIP; goto done; }
      for (int k = 0; k < slots[q].nreq; ++k)
        for (int i = 0; i < nact; ++i)
          if (act[i].req == slots[q].req[k]) act[i].rounds--;
      slots[q].busy = 0;
      progress = 1;
    }
    for (int i = 0; i < nact;) {
      active_t* a = &act[i];
      if (a->next < a->ntasks || a->rounds > 0) { ++i; continue; }
      const int r = a->req;
      dataset_t* D = &ds[a->ds];
      const double tdone = now_s() - t0;
      lat[r] = (tdone - reqs[r].arrival_s) * 1e3;
      const wrt_job_desc* d = &D->desc;
      samples += d->op == WRT_OP_RADON ? 16.0 * d->n * (double)d->n * (double)a->fd.nout
                                       : (double)a->fd.nvox * a->fd.item.andir;
      if (outputs && outputs[r]) {
        rc = d->op == WRT_OP_RADON ? wrt_fast_fetch_acc(D->job, outputs[r])
                                   : wrt_job_fetch(D->job, outputs[r]);
        if (rc != WRT_OK) goto done;
      }
      D->busy = -1;
      D->last_used = tdone;
      free(a->tasks);
      /* keep arrival order */
      memmove(&act[i], &act[i + 1], sizeof(active_t) * (size_t)(nact - i - 1));
      --nact;
      ++completed;
      progress = 1;
    }

    /* start pending requests in arrival order */
    const int slots_own = c->plain ? 1 : c->max_inflight;
    for (int q = 0; q < npend;) {
      const int r = pending[q];
      dataset_t* D = &ds[req_ds[r]];
      const int fast = !c->plain && batched(&D->desc);
      if (D->busy >= 0 || (!fast && nrun >= slots_own)) { ++q; continue; }
      if (!D->job) {
        for (;;) { /* room: the budget and the live free memory, evicting idle datasets */
          size_t fr = 0, tot = 0;
          rc = wrt_mem_info(ctx, &fr, &tot);
          if (rc != WRT_OK) goto done;
          if (used + D->bytes <= budget && D->bytes + reserve <= fr) break;
          int victim = -1;
          for (int k = 0; k < nds; ++k)
            if (ds[k].job && ds[k].busy < 0 && (victim < 0 || ds[k].last_used < ds[victim].last_used))
              victim = k;
          if (victim < 0) break;
          wrt_job_destroy(ds[victim].job);
          ds[victim].job = NULL;
          used -= ds[victim].bytes;
          rep->evictions++;
        }
        {
          size_t fr = 0, tot = 0;
          rc = wrt_mem_info(ctx, &fr, &tot);
          if (rc != WRT_OK) goto done;
          if (used + D->bytes > budget || D->bytes + reserve > fr) { ++q; continue; } /* later */
        }
        rc = wrt_job_create(ctx, &D->desc, &D->job);
        if (rc == WRT_OK) rc = wrt_job_load(D->job, D->host_in);
        if (rc != WRT_OK) goto done;
        used += D->bytes;
        if (used > rep->peak_bytes) rep->peak_bytes = used;
        rep->loads++;
      }
      if (fast) {
        active_t* a = &act[nact];
        memset(a, 0, sizeof(*a));
        a->req = r;
        a->ds = req_ds[r];
        a->arrival = reqs[r].arrival_s;
        rc = wrt_fast_describe(D->job, &a->fd);
        if (rc != WRT_OK) goto done;
        /* the per-job launch's direction group, at most group_cap, a multiple of 16 */
        int G = a->fd.item.kind == WRT_ITEM_RADON ? a->fd.dgroup : 1;
        if (G > group_cap) G = group_cap & ~15;
        a->G = G;
        a->ntasks = wrt_item_tasks(&a->fd, 0, G, NULL);
        a->tasks = (wrt_costed_task*)malloc(sizeof(wrt_costed_task) * (size_t)(a->ntasks + 1));
        if (!a->tasks) { rc = WRT_E_NOMEM; goto done; }
        wrt_item_tasks(&a->fd, 0, G, a->tasks);
        for (long long k = 0; k < a->ntasks; ++k) a->left += a->tasks[k].cost;
        if (a->ntasks == 0 && a->fd.item.kind == WRT_ITEM_RADON) {
          /* an all-zero volume: no task, so no prologue zeroes its planes */
          rc = wrtk_fill_f64((double*)a->fd.acc, 0.0, (size_t)a->fd.nout, ctx->stream);
          if (rc == WRT_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = WRT_E_HIP;
          if (rc != WRT_OK) goto done;
        }
        /* arrival order on the active list */
        int pos = nact;
        while (pos > 0 && act[pos - 1].req > r) --pos;
        if (pos < nact) {
          const active_t tmp = *a;
          memmove(&act[pos + 1], &act[pos], sizeof(active_t) * (size_t)(nact - pos));
          act[pos] = tmp;
        }
        ++nact;
        rep->batched++;
      } else {
        rc = wrt_job_launch(D->job);
        if (rc != WRT_OK) goto done;
        running[nrun++] = r;
      }
      D->busy = r;
      if (nact + nrun > rep->peak_inflight) rep->peak_inflight = nact + nrun;
      pending[q] = pending[--npend];
      /* keep arrival order among the rest */
      for (int a2 = q; a2 + 1 < npend && pending[a2] > pending[a2 + 1]; ++a2) {
        const int tmp = pending[a2];
        pending[a2] = pending[a2 + 1];
        pending[a2 + 1] = tmp;
      }
      progress = 1;
    }

    /* rounds: small requests at once in the express slot, then the main slots */
    for (int q = RS_SLOTS - 1; q >= 0; --q) {
      if (slots[q].busy) continue;
      rc = round_launch(ctx, slots, q, q >= RS_MAIN, act, nact, round_s, min_waves, now_s() - t0);
      if (rc != WRT_OK) goto done;
      if (!slots[q].busy) continue; /* nothing to schedule there */
      ++nrounds;
      progress = 1;
    }

    if (!progress) {
      if (nrun > 0 || nact > 0)
        sleep_s(20e-6);
      else if (next < nreq)
        sleep_s(reqs[next].arrival_s - (now_s() - t0));
    }
  }
  rep->wall_s = now_s() - t0;
  rep->completed = completed;
  rep->samples = samples;
  rep->samples_per_s = rep->wall_s > 0 ? samples / rep->wall_s : 0;
  rep->rounds = (int)nrounds;
  {
    double* sorted = (double*)malloc(sizeof(double) * (size_t)nreq);
    if (!sorted) { rc = WRT_E_NOMEM; goto done; }
    memcpy(sorted, lat, sizeof(double) * (size_t)nreq);
    qsort(sorted, (size_t)nreq, sizeof(double), cmp_double);
    double sum = 0;
    for (int i = 0; i < nreq; ++i) sum += sorted[i];
    rep->p50_ms = pct(sorted, nreq, 50);
    rep->p99_ms = pct(sorted, nreq, 99);
    rep->mean_ms = sum / nreq;
    rep->max_ms = sorted[nreq - 1];
    free(sorted);
  }
  if (latency_ms) memcpy(latency_ms, lat, sizeof(double) * (size_t)nreq);

done:
  for (int q = 0; q < RS_SLOTS; ++q) slot_free(&slots[q]);
  if (act)
    for (int i = 0; i < nact; ++i) free(act[i].tasks);
  if (ds)
    for (int k = 0; k < nds; ++k) {
      if (ds[k].job) wrt_job_destroy(ds[k].job);
      hipHostFree(ds[k].host_in);
    }
  if (grids)
    for (int g = 0; g < ngrids; ++g) {
      free(grids[g].phi);
      free(grids[g].theta);
      free(grids[g].w);
      free(grids[g].shift);
      free(grids[g].z);
    }
  free(grids);
  free(ds);
  free(act);
  free(lat);
  free(running);
  free(pending);
  free(req_ds);
  free(reqs);
  return rc;
}
