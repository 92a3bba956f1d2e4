/*
 * wrt_serve.c -- the request loop: Poisson arrivals, admission by device bytes, resident
 * datasets, overlapping launches on per-dataset streams, latency percentiles.
 *
 * One host thread drives everything.  Each loop pass (1) moves arrived requests to the
 * pending list, (2) polls the running requests and records the finished ones, (3) starts
 * pending requests in arrival order while a slot is free: a request whose dataset is not
 * resident first makes it resident (create + load, evicting idle datasets if the budget
 * asks), a request whose dataset is busy waits for it.  Nothing here waits on the device
 * except a load, which runs on the loading dataset's own stream.
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

int wrt_serve_run(wrt_ctx* ctx, const wrt_serve_cfg* c, wrt_serve_report* rep,
                  double* latency_ms, double* const* outputs) {
  if (!ctx || !cfg_ok(c) || !rep) return WRT_E_ARG;
  memset(rep, 0, sizeof(*rep));
  const int nreq = c->nreq;
  if (nreq == 0) return WRT_OK;
  int rc = WRT_OK;
  wrt_serve_req* reqs = (wrt_serve_req*)calloc((size_t)nreq, sizeof(*reqs));
  int* req_ds = (int*)calloc((size_t)nreq, sizeof(int));
  int* pending = (int*)calloc((size_t)nreq, sizeof(int));
  int* running = (int*)calloc((size_t)nreq, sizeof(int));
  double* lat = (double*)calloc((size_t)nreq, sizeof(double));
  const int maxds = 3 * c->nedges * c->ndatasets;
  dataset_t* ds = (dataset_t*)calloc((size_t)maxds, sizeof(dataset_t));
  grid_t* grids = (grid_t*)calloc((size_t)c->nedges, sizeof(grid_t));
  int nds = 0, ngrids = 0;
  if (!reqs || !req_ds || !pending || !running || !lat || !ds || !grids) {
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
      D->host_in = (double*)malloc(sizeof(double) * cnt);
      if (!D->host_in || D->bytes == 0) { rc = D->host_in ? WRT_E_ARG : WRT_E_NOMEM; break; }
      rc = wrt_serve_input(&reqs[i], D->host_in);
    }
    req_ds[i] = k;
  }
  if (rc != WRT_OK) goto done;

  size_t budget = c->hbm_budget;
  if (budget == 0) {
    size_t fr = 0, tot = 0;
    rc = wrt_mem_info(ctx, &fr, &tot);
    if (rc != WRT_OK) goto done;
    budget = (size_t)(0.9 * (double)fr);
  }
  for (int k = 0; k < nds; ++k)
    if (ds[k].bytes > budget) { rc = WRT_E_ADMIT; goto done; }

  size_t used = 0;
  int npend = 0, nrun = 0, next = 0, completed = 0;
  double samples = 0.0;
  const double t0 = now_s();
  while (completed < nreq) {
    const double t = now_s() - t0;
    int progress = 0;
    while (next < nreq && reqs[next].arrival_s <= t) pending[npend++] = next++;

    /* finished requests */
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

    /* start pending requests in arrival order */
    const int slots = c->plain ? 1 : c->max_inflight;
    for (int q = 0; q < npend && nrun < slots;) {
      const int r = pending[q];
      dataset_t* D = &ds[req_ds[r]];
      if (D->busy >= 0) { ++q; continue; }  /* its dataset serves an earlier request */
      if (!D->job) {
        while (used + D->bytes > budget) {  /* evict the least recently used idle dataset */
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
        if (used + D->bytes > budget) { ++q; continue; }  /* not now: memory frees later */
        rc = wrt_job_create(ctx, &D->desc, &D->job);
        if (rc == WRT_OK) rc = wrt_job_load(D->job, D->host_in);
        if (rc != WRT_OK) goto done;
        used += D->bytes;
        if (used > rep->peak_bytes) rep->peak_bytes = used;
        rep->loads++;
      }
      rc = wrt_job_launch(D->job);
      if (rc != WRT_OK) goto done;
      D->busy = r;
      running[nrun++] = r;
      if (nrun > rep->peak_inflight) rep->peak_inflight = nrun;
      pending[q] = pending[--npend];
      /* keep arrival order among the rest */
      for (int a = q; a + 1 < npend && pending[a] > pending[a + 1]; ++a) {
        const int tmp = pending[a];
        pending[a] = pending[a + 1];
        pending[a + 1] = tmp;
      }
      progress = 1;
    }

    if (!progress) {
      if (nrun > 0)
        sleep_s(20e-6);
      else if (next < nreq)
        sleep_s(reqs[next].arrival_s - (now_s() - t0));
    }
  }
  rep->wall_s = now_s() - t0;
  rep->completed = completed;
  rep->samples = samples;
  rep->samples_per_s = rep->wall_s > 0 ? samples / rep->wall_s : 0;
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
  if (ds)
    for (int k = 0; k < nds; ++k) {
      if (ds[k].job) wrt_job_destroy(ds[k].job);
      free(ds[k].host_in);
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
  free(lat);
  free(running);
  free(pending);
  free(req_ds);
  free(reqs);
  return rc;
}
