/*
 * wrt_batch.c -- many jobs in one submission: the fast plane-integral and backprojection
 * jobs of the list run as one persistent launch (k_batch) captured in a HIP graph, the
 * others (plain FP64 forms, reductions) on their own streams beside it.
 *
 * Plan (built on the first call for a list, replayed while the list and every job's load
 * generation stay the same):
 *   items    one per batched job: its arguments (wrt_fast_describe);
 *   tasks    plane integrals: (job, non-empty brick, direction group of G directions);
 *            backprojections: (job, up to WRT_ADJ_TASK_TILES voxel tiles);
 *            ordered by an estimate of their device time, longest first, so that the
 *            persistent workgroups, which take tasks from a counter, finish together;
 *   segments the FP64 accumulators the prologue zeroes and the epilogue rounds to FP32;
 *   graph    prologue (zero fill by kernel, no memset node) -> persistent kernel -> epilogue.
 * The direction-group size G is the tunable: the autotuner times the batch at each
 * candidate on the first call for a list shape and keeps the fastest, in memory and in the
 * tuning cache file (host/wrt_tune.c; set WRT_AUTOTUNE=0 to skip tuning and use G = 64, or
 * WRT_BATCH_DGROUP=G (>= 16) to fix G).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "wrt_internal.h"

#define HIPCHK(x)                            \
  do {                                       \
    if ((x) != hipSuccess) return WRT_E_HIP; \
  } while (0)

typedef struct {
  int njobs;
  wrt_job** jobs;
  unsigned long long* gens;
  int dgroup;
  wrt_batch_item* d_items;
  wrt_batch_task* d_tasks;
  int ntasks;
  wrt_batch_seg* d_segs;
  int nseg;
  int* d_counter;
  hipGraph_t graph;
  hipGraphExec_t exec;
  double samples;
} batch_plan;

static void plan_free(batch_plan* p) {
  if (!p) return;
  if (p->exec) hipGraphExecDestroy(p->exec);
  if (p->graph) hipGraphDestroy(p->graph);
  hipFree(p->d_items);
  hipFree(p->d_tasks);
  hipFree(p->d_segs);
  hipFree(p->d_counter);
  free(p->jobs);
  free(p->gens);
  free(p);
}

void wrt_batch_release(wrt_ctx* ctx) {
  if (!ctx) return;
  plan_free((batch_plan*)ctx->batch);
  ctx->batch = NULL;
}

void wrt_batch_forget(wrt_ctx* ctx, const wrt_job* j) {
  const batch_plan* p = ctx ? (const batch_plan*)ctx->batch : NULL;
  if (!p) return;
  for (int i = 0; i < p->njobs; ++i)
    if (p->jobs[i] == j) {
      hipStreamSynchronize(ctx->stream); /* a replay still running reads the job's buffers */
      wrt_batch_release(ctx);
      return;
    }
}

/* tasks address voxels with 32-bit offsets: larger backprojections run beside the graph */
static int batchable(const wrt_job* j) {
  return j->d.prec != WRT_F64 && wrt_fast_supported(&j->d) &&
         (j->d.op == WRT_OP_RADON || (j->d.op == WRT_OP_ADJOINT && j->d.n <= 1024));
}

int wrt_costed_desc(const void* a, const void* b) {
  const double x = ((const wrt_costed_task*)a)->cost, y = ((const wrt_costed_task*)b)->cost;
  return x < y ? 1 : (x > y ? -1 : 0);
}

long long wrt_item_tasks(const wrt_fast_desc* fd, int item, int G, wrt_costed_task* ct) {
  const wrt_batch_item* it = &fd->item;
  if (it->kind == WRT_ITEM_RADON) {
    const int ndir = it->ra.ndir, n = it->ra.n;
    const long long count = (long long)fd->nbrick * ((ndir + G - 1) / G);
    if (!ct) return count;
    const double shifts = (double)(it->ra.nshift - 1) / n; /* shift density vs 2/n */
    long long k = 0;
    for (int b = 0; b < fd->nbrick; ++b)
      for (int g = 0; g * G < ndir; ++g) {
        const int nd = (g + 1) * G <= ndir ? G : ndir - g * G;
        ct[k].cost = WRT_COST_BRICK_DIR * nd * shifts * (fd->weight[b] / 32768.0);
        ct[k].t.item = item;
        ct[k].t.a = fd->brick[b];
        ct[k].t.b = g;
        ct[k].t.c = fd->bnd[b];
        ++k;
      }
    return k;
  }
  const long long ntiles = wrt_adj_tiles(it->an, WRT_ADJ_VPT);
  const long long count = (ntiles + WRT_ADJ_TASK_TILES - 1) / WRT_ADJ_TASK_TILES;
  if (!ct) return count;
  long long k = 0;
  for (long long v = 0; v < ntiles; v += WRT_ADJ_TASK_TILES) {
    const int cnt = (int)(ntiles - v < WRT_ADJ_TASK_TILES ? ntiles - v : WRT_ADJ_TASK_TILES);
    ct[k].cost = WRT_COST_VOXEL_DIR * cnt * (64.0 * WRT_ADJ_VPT) * it->andir;
    ct[k].t.item = item;
    ct[k].t.a = (int)v;
    ct[k].t.b = cnt;
    ct[k].t.c = 0;
    ++k;
  }
  return k;
}

/* the task list for direction groups of G, longest first; returns the count or < 0 */
static int build_tasks(const wrt_fast_desc* fd, int nitems, int G, wrt_batch_task** out) {
  long long total = 0;
  for (int i = 0; i < nitems; ++i) total += wrt_item_tasks(&fd[i], i, G, NULL);
  if (total > 0x7fffffff) return WRT_E_ARG;
  wrt_costed_task* ct = (wrt_costed_task*)malloc(sizeof(wrt_costed_task) * (size_t)(total + 1));
  if (!ct) return WRT_E_NOMEM;
  long long k = 0;
  for (int i = 0; i < nitems; ++i) k += wrt_item_tasks(&fd[i], i, G, ct + k);
  qsort(ct, (size_t)k, sizeof(wrt_costed_task), wrt_costed_desc);
  wrt_batch_task* t = (wrt_batch_task*)malloc(sizeof(wrt_batch_task) * (size_t)(k + 1));
  if (!t) {
    free(ct);
    return WRT_E_NOMEM;
  }
  for (long long q = 0; q < k; ++q) t[q] = ct[q].t;
  free(ct);
  *out = t;
  return (int)k;
}

/* upload items (with direction groups of G) and tasks into the plan's buffers */
static int plan_upload(batch_plan* p, wrt_fast_desc* fd, int nitems, int G, hipStream_t st) {
  for (int i = 0; i < nitems; ++i) fd[i].item.ra.dgroup = G;
  wrt_batch_task* t = NULL;
  const int nt = build_tasks(fd, nitems, G, &t);
  if (nt < 0) return nt;
  int rc = WRT_OK;
  hipFree(p->d_tasks);
  p->d_tasks = NULL;
  wrt_batch_item* items = (wrt_batch_item*)malloc(sizeof(wrt_batch_item) * (size_t)nitems);
  if (!items) rc = WRT_E_NOMEM;
  if (rc == WRT_OK) {
    for (int i = 0; i < nitems; ++i) items[i] = fd[i].item;
    if (hipMalloc((void**)&p->d_tasks, sizeof(wrt_batch_task) * (size_t)(nt + 1)) != hipSuccess)
      rc = WRT_E_NOMEM;
  }
  if (rc == WRT_OK &&
      (hipMemcpyAsync(p->d_items, items, sizeof(wrt_batch_item) * (size_t)nitems,
                      hipMemcpyHostToDevice, st) != hipSuccess ||
       (nt > 0 && hipMemcpyAsync(p->d_tasks, t, sizeof(wrt_batch_task) * (size_t)nt,
                                 hipMemcpyHostToDevice, st) != hipSuccess) ||
       hipStreamSynchronize(st) != hipSuccess))
    rc = WRT_E_HIP;
  free(items);
  free(t);
  if (rc == WRT_OK) {
    p->ntasks = nt;
    p->dgroup = G;
  }
  return rc;
}

static int enqueue_plan(wrt_ctx* c, batch_plan* p, hipStream_t st) {
  int rc = wrtk_batch_prologue(p->d_segs, p->nseg, p->d_counter, st);
  if (rc == WRT_OK) rc = wrtk_batch_run(p->d_items, p->d_tasks, p->ntasks, p->d_counter, c->cu_count, st);
  if (rc == WRT_OK) rc = wrtk_batch_epilogue(p->d_segs, p->nseg, st);
  return rc;
}

/* ---- autotune: the plan's signature in the tuning cache (host/wrt_tune.c) ---- */
static unsigned long long signature(const wrt_fast_desc* fd, int nitems) {
  long long* w = (long long*)malloc(sizeof(long long) * (size_t)(2 + 6 * nitems));
  if (!w) return 0;
  int k = 0;
  w[k++] = 0x62617463680aLL; /* "batch": its own signature space */
  w[k++] = nitems;
  for (int i = 0; i < nitems; ++i) {
    const wrt_batch_item* it = &fd[i].item;
    w[k++] = it->kind;
    if (it->kind == WRT_ITEM_RADON) {
      w[k++] = it->ra.n;
      w[k++] = it->ra.ndir;
      w[k++] = it->ra.nshift;
      w[k++] = it->ra.bf16;
      w[k++] = fd[i].nbrick;
    } else {
      w[k++] = it->an;
      w[k++] = it->andir;
      w[k++] = it->ans;
    }
  }
  const unsigned long long h = wrt_tune_hash(w, k);
  free(w);
  return h;
}

static int time_plan(wrt_ctx* c, batch_plan* p, double* sec) {
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  const int rc = enqueue_plan(c, p, c->stream);
  if (rc != WRT_OK) return rc;
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *sec = ms * 1e-3;
  return WRT_OK;
}

static int plan_build(wrt_ctx* c, wrt_job* const* jobs, int njobs, batch_plan** out) {
  batch_plan* p = (batch_plan*)calloc(1, sizeof(batch_plan));
  if (!p) return WRT_E_NOMEM;
  int rc = WRT_OK;
  p->njobs = njobs;
  p->jobs = (wrt_job**)malloc(sizeof(wrt_job*) * (size_t)njobs);
  p->gens = (unsigned long long*)malloc(sizeof(unsigned long long) * (size_t)njobs);
  wrt_fast_desc* fd = (wrt_fast_desc*)calloc((size_t)njobs + 1, sizeof(wrt_fast_desc));
  wrt_batch_seg* segs = (wrt_batch_seg*)calloc((size_t)njobs + 1, sizeof(wrt_batch_seg));
  if (!p->jobs || !p->gens || !fd || !segs) rc = WRT_E_NOMEM;
  int nitems = 0, nseg = 0;
  for (int i = 0; rc == WRT_OK && i < njobs; ++i) {
    p->jobs[i] = jobs[i];
    p->gens[i] = jobs[i]->gen;
    if (!batchable(jobs[i])) continue;
    rc = wrt_fast_describe(jobs[i], &fd[nitems]);
    if (rc != WRT_OK) break;
    p->samples += jobs[i]->d.op == WRT_OP_RADON
                      ? 16.0 * jobs[i]->d.n * (double)jobs[i]->d.n * (double)fd[nitems].nout
                      : (double)fd[nitems].nvox * fd[nitems].item.andir;
    if (fd[nitems].item.kind == WRT_ITEM_RADON) {
      segs[nseg].acc = fd[nitems].acc;
      segs[nseg].qstep = fd[nitems].qstep;
      segs[nseg].out = fd[nitems].out;
      segs[nseg].count = fd[nitems].nout;
      ++nseg;
    }
    ++nitems;
  }
  if (rc == WRT_OK &&
      (hipMalloc((void**)&p->d_items, sizeof(wrt_batch_item) * (size_t)(nitems + 1)) != hipSuccess ||
       hipMalloc((void**)&p->d_segs, sizeof(wrt_batch_seg) * (size_t)(nseg + 1)) != hipSuccess ||
       hipMalloc((void**)&p->d_counter, 256) != hipSuccess))
    rc = WRT_E_NOMEM;
  if (rc == WRT_OK && nseg > 0 &&
      (hipMemcpyAsync(p->d_segs, segs, sizeof(wrt_batch_seg) * (size_t)nseg, hipMemcpyHostToDevice,
                      c->stream) != hipSuccess ||
       hipStreamSynchronize(c->stream) != hipSuccess))
    rc = WRT_E_HIP;
  p->nseg = nseg;

  /* direction-group size: cached, tuned, or the default */
  int G = 64;
  const char* force = getenv("WRT_BATCH_DGROUP");
  const int forced = force ? atoi(force) : 0;
  if (rc == WRT_OK && nitems > 0 && forced >= 16) {
    G = forced;
  } else if (rc == WRT_OK && nitems > 0) {
    const unsigned long long sig = signature(fd, nitems);
    const int tune = wrt_tune_enabled();
    int cached = 0, unused = 0;
    if (tune && !(wrt_tune_lookup(sig, &cached, &unused) && cached >= 16)) cached = 0;
    if (cached) {
      G = cached;
    } else if (tune) {
      static const int cand[] = {16, 32, 64, 128, 256, 512};
      double best = 1e300;
      for (int q = 0; rc == WRT_OK && q < (int)(sizeof(cand) / sizeof(cand[0])); ++q) {
        double sec = 0;
        rc = plan_upload(p, fd, nitems, cand[q], c->stream);
        if (rc == WRT_OK) rc = time_plan(c, p, &sec);
        if (rc == WRT_OK && sec < best) {
          best = sec;
          G = cand[q];
        }
      }
      if (rc == WRT_OK) wrt_tune_store(sig, G, 0, best);
    }
  }
  if (rc == WRT_OK) rc = plan_upload(p, fd, nitems, G, c->stream);

  /* capture prologue -> persistent kernel -> epilogue */
  if (rc == WRT_OK) {
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      rc = WRT_E_HIP;
    } else {
      const int erc = enqueue_plan(c, p, c->stream);
      hipGraph_t g = NULL;
      const hipError_t e = hipStreamEndCapture(c->stream, &g);
      p->graph = g;
      if (erc != WRT_OK) rc = erc;
      else if (e != hipSuccess || !g) rc = WRT_E_HIP;
      else if (hipGraphInstantiate(&p->exec, g, NULL, NULL, 0) != hipSuccess) rc = WRT_E_HIP;
    }
  }
  free(fd);
  free(segs);
  if (rc != WRT_OK) {
    plan_free(p);
    return rc;
  }
  *out = p;
  return WRT_OK;
}

static int plan_matches(const batch_plan* p, wrt_job* const* jobs, int njobs) {
  if (!p || p->njobs != njobs) return 0;
  for (int i = 0; i < njobs; ++i)
    if (p->jobs[i] != jobs[i] || p->gens[i] != jobs[i]->gen) return 0;
  return 1;
}

int wrt_batch_run(wrt_ctx* ctx, wrt_job* const* jobs, int njobs, wrt_stats* st) {
  if (!ctx || njobs < 0 || (njobs > 0 && !jobs)) return WRT_E_ARG;
  for (int i = 0; i < njobs; ++i) {
    if (!jobs[i] || jobs[i]->ctx != ctx) return WRT_E_ARG;
    if (!jobs[i]->loaded || jobs[i]->pending) return WRT_E_STATE;
  }
  HIPCHK(hipSetDevice(ctx->device));
  if (st) memset(st, 0, sizeof(*st));
  int nbatch = 0;
  for (int i = 0; i < njobs; ++i) nbatch += batchable(jobs[i]);

  /* the others run on their own streams beside the graph */
  int rc = WRT_OK;
  for (int i = 0; rc == WRT_OK && i < njobs; ++i)
    if (!batchable(jobs[i])) rc = wrt_job_launch(jobs[i]);
  if (rc != WRT_OK) return rc;

  double gsec = 0, samples = 0;
  if (nbatch > 0) {
    batch_plan* p = (batch_plan*)ctx->batch;
    if (!plan_matches(p, jobs, njobs)) {
      wrt_batch_release(ctx);
      rc = plan_build(ctx, jobs, njobs, &p);
      if (rc == WRT_OK) ctx->batch = p;
    }
    if (rc == WRT_OK) {
      if (hipEventRecord(ctx->ev0, ctx->stream) != hipSuccess ||
          hipGraphLaunch(p->exec, ctx->stream) != hipSuccess ||
          hipEventRecord(ctx->ev1, ctx->stream) != hipSuccess ||
          hipEventSynchronize(ctx->ev1) != hipSuccess)
        rc = WRT_E_HIP;
      float ms = 0;
      if (rc == WRT_OK && hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) != hipSuccess) rc = WRT_E_HIP;
      gsec = ms * 1e-3;
      samples = p->samples;
    }
  }
  double solo = 0;
  for (int i = 0; i < njobs; ++i)
    if (!batchable(jobs[i])) {
      wrt_stats s;
      const int r = wrt_job_wait(jobs[i], &s);
      if (rc == WRT_OK) rc = r;
      if (r == WRT_OK) {
        samples += s.logical_samples;
        if (s.seconds > solo) solo = s.seconds;
      }
    }
  if (rc != WRT_OK) return rc;
  if (st) {
    st->kernel_seconds = gsec;
    st->seconds = gsec > solo ? gsec : solo;
    st->logical_samples = samples;
    st->outputs = 0;
    for (int i = 0; i < njobs; ++i) st->outputs += (long long)jobs[i]->out_count;
  }
  return WRT_OK;
}

/* the direction-group size of the current plan (0 if none) */
int wrt_batch_dgroup(wrt_ctx* ctx) {
  const batch_plan* p = ctx ? (const batch_plan*)ctx->batch : NULL;
  return p ? p->dgroup : 0;
}
