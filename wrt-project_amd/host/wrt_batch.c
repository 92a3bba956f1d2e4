/*
 * wrt_batch.c -- many jobs in one submission.
 */
#include <stdlib.h>

#include "wrt_internal.h"

void wrt_batch_release(wrt_ctx* ctx) { (void)ctx; }

int wrt_batch_run(wrt_ctx* ctx, wrt_job* const* jobs, int njobs, wrt_stats* st) {
  if (!ctx || (njobs > 0 && !jobs)) return WRT_E_ARG;
  double secs = 0;
  for (int i = 0; i < njobs; ++i) {
    wrt_stats s = {0};
    int rc = wrt_job_run(jobs[i], &s);
    if (rc != WRT_OK) return rc;
    secs += s.seconds;
  }
  if (st) st->seconds = secs;
  return WRT_OK;
}
