/*
 * wrt_fast.c -- host side of the fast (FP32 / BF16-brick) forms.
 */
#include <stdlib.h>

#include "wrt_internal.h"

int wrt_fast_supported(const wrt_job_desc* d) { (void)d; return 0; }
size_t wrt_fast_bytes(const wrt_job_desc* d) { (void)d; return 0; }
int wrt_fast_create(wrt_job* j) { (void)j; return WRT_E_ARG; }
int wrt_fast_load(wrt_job* j, const double* in) { (void)j; (void)in; return WRT_E_ARG; }
int wrt_fast_run(wrt_job* j, wrt_stats* st) { (void)j; (void)st; return WRT_E_ARG; }
int wrt_fast_fetch(wrt_job* j, double* out) { (void)j; (void)out; return WRT_E_ARG; }
void wrt_fast_destroy(wrt_job* j) { (void)j; }
