// k_noalloc.hip -- the fixed device scratch of the no-allocation plane-integral entry point
// (wrt_radon_grid_noalloc, host/wrt_noalloc.c) and its small kernels.
//
// The scratch is one file-scope device array of WRT_NOALLOC_BYTES.  Each GPU the library
// runs on holds its own copy, at its own address, until the process ends.  In host code the
// array's name is not a device address: wrtk_noalloc_setup (called by wrt_init with the
// device current) takes that address once per device with hipGetSymbolAddress, and every
// later use goes through the pointer it keeps.  Calls that overlap on one GPU share the
// array and race; the entry point clears what must start at zero on every call.
#include "wrt_common.hpp"
#include "wrt_fast.h"

static __device__ __align__(256) unsigned char scratch[WRT_NOALLOC_BYTES];

#define NA_MAX_DEVICES 64
static void* na_ptr[NA_MAX_DEVICES];

extern "C" int wrtk_noalloc_setup(int device) {
  if (device < 0 || device >= NA_MAX_DEVICES) return WRT_E_ARG;
  if (na_ptr[device]) return WRT_OK;
  void* p = NULL;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return WRT_E_HIP;
  na_ptr[device] = p;
  return WRT_OK;
}

extern "C" void* wrtk_noalloc_scratch(int device) {
  return (device >= 0 && device < NA_MAX_DEVICES) ? na_ptr[device] : NULL;
}

__global__ __launch_bounds__(256) void k_na_zero(unsigned long long* __restrict__ p, size_t count) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

extern "C" int wrtk_noalloc_zero(void* p, size_t words, hipStream_t st) {
  if (!words) return WRT_OK;
  size_t blocks = (words + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  k_na_zero<<<(unsigned)blocks, 256, 0, st>>>((unsigned long long*)p, words);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

// The task list of the brick kernel from the bricks' non-zero counts: (brick, group) for
// every non-empty brick and direction group, interior bricks first, then bricks on a cube
// face; counts[0], counts[1] = their numbers.  One workgroup: the bricks are listed in
// index order in LDS, then the tasks written in parallel.
__global__ __launch_bounds__(256) void k_na_tasks(const int* __restrict__ flags, int n, int nb,
                                                  int ngroups, int* __restrict__ tasks,
                                                  int* __restrict__ counts) {
  __shared__ int list[WRT_NOALLOC_MAX_BRICKS];
  __shared__ int nint, nall;
  const int nbr = nb * nb * nb;
  if (threadIdx.x == 0) {
    int k = 0;
    for (int pass = 0; pass < 2; ++pass) {
      for (int b = 0; b < nbr; ++b) {
        if (!flags[b]) continue;
        const int bx = b / (nb * nb), by = (b / nb) % nb, bz = b % nb;
        const int bnd = bx == 0 || by == 0 || bz == 0 || (bx + 1) * WRT_BRICK >= n ||
                        (by + 1) * WRT_BRICK >= n || (bz + 1) * WRT_BRICK >= n;
        if (bnd == pass) list[k++] = b;
      }
      if (pass == 0) nint = k;
    }
    nall = k;
  }
  __syncthreads();
  const int total = nall * ngroups;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    tasks[2 * t] = list[t / ngroups];
    tasks[2 * t + 1] = t % ngroups;
  }
  if (threadIdx.x == 0) {
    counts[0] = nint * ngroups;
    counts[1] = (nall - nint) * ngroups;
  }
}

extern "C" int wrtk_noalloc_tasks(const int* flags, int n, int nb, int ngroups, int* tasks,
                                  int* counts, hipStream_t st) {
  if (nb < 1 || nb * nb * nb > WRT_NOALLOC_MAX_BRICKS || ngroups < 1) return WRT_E_ARG;
  k_na_tasks<<<1, 256, 0, st>>>(flags, n, nb, ngroups, tasks, counts);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

// the fixed-point plane sums (units of qstep) -> the FP32 form's output, kept as doubles in
// the same 8-byte slots for the host copy
__global__ __launch_bounds__(256) void k_na_round(long long* __restrict__ acc, double qstep,
                                                  size_t count) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    acc[i] = __double_as_longlong((double)(float)((double)acc[i] * qstep));
}

extern "C" int wrtk_noalloc_round(long long* acc, double qstep, size_t count, hipStream_t st) {
  if (!count) return WRT_OK;
  size_t blocks = (count + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  k_na_round<<<(unsigned)blocks, 256, 0, st>>>(acc, qstep, count);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}
