// k_batch.hip -- one persistent launch over the tasks of many jobs (plane integrals and
// backprojections of volumes of any size), the form the batch runner captures in a HIP
// graph.
//
// The grid is one workgroup of 1024 threads per CU (the brick task's LDS allows no more).
// A workgroup takes the next task from a device counter, runs it, and takes another until
// the list is exhausted, so uneven tasks balance across the CUs; the host orders the list
// longest first.  A plane-integral task is the brick task of the per-job launch
// (radon_brick.hpp) with the job's arguments; a backprojection task is up to
// WRT_ADJ_TASK_TILES voxel tiles of one job, one per wavefront, WRT_ADJ_VPT voxels per
// thread, with the directions' cubic rows staged in the brick's LDS (adjoint_staged_tiles,
// in chunks of WRT_ADJ_PF float4 per thread, half the per-job launch's: the one register
// allocation of this kernel has no room for more).  Results are the
// per-job launches', bit for bit: every voxel's sum runs in the same order, and the plane
// sums are exact 64-bit fixed-point sums, whatever the order of their sections.
#include "adjoint_voxel.hpp"
#include "radon_brick.hpp"

// zero the fixed-point plane sums (segment blockIdx.y < nseg) and the task counter
__global__ __launch_bounds__(256) void k_batch_prologue(const wrt_batch_seg* __restrict__ segs,
                                                        int nseg, int* __restrict__ counter) {
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *counter = 0;
  if ((int)blockIdx.y >= nseg) return;
  const wrt_batch_seg sg = segs[blockIdx.y];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < sg.count;
       i += (long long)gridDim.x * blockDim.x)
    sg.acc[i] = 0ll;
}

__global__ __launch_bounds__(256) void k_batch_epilogue(const wrt_batch_seg* __restrict__ segs) {
  const wrt_batch_seg sg = segs[blockIdx.y];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < sg.count;
       i += (long long)gridDim.x * blockDim.x)
    sg.out[i] = (float)((double)sg.acc[i] * sg.qstep);
}

__global__ __launch_bounds__(FR_WAVES * 64) void k_batch(const wrt_batch_item* __restrict__ items,
                                                         const wrt_batch_task* __restrict__ tasks,
                                                         int ntasks, int* __restrict__ counter) {
  // 16-byte aligned: a backprojection task stages float4 cubic rows in L.sb
  __shared__ __attribute__((aligned(16))) radon_lds L;
  static_assert(sizeof(L.sb) >= ADJ_LDS_F4 * sizeof(float4), "the staged rows fit L.sb");
  __shared__ int s_task;
  for (;;) {
    __syncthreads();  // the previous task is done with L and s_task
    if (threadIdx.x == 0) s_task = atomicAdd(counter, 1);
    __syncthreads();
    // wave-uniform, so that the task and its item are read into scalar registers
    const int ti = __builtin_amdgcn_readfirstlane(s_task);
    if (ti >= ntasks) break;  // every workgroup reaches this: the counter only grows
    const wrt_batch_task t = tasks[ti];
    const wrt_batch_item& it = items[t.item];
    if (it.kind == WRT_ITEM_RADON) {
      // one code path for every plane-integral task: face brick, BF16 and power-of-two edge
      // are run-time flags (each inlined variant would add its own live ranges to the one
      // register allocation of this kernel)
      radon_brick_task<false>(it.ra, t.a, t.b, it.h, it.dc, it.inv_dc, L, t.c != 0,
                              it.ra.bf16 != 0, it.pow2 != 0);
    } else if (adjoint_chunk(it.ans, WRT_ADJ_PF) >= 4) {
      // tile aorder[a + w] for wavefront w < b of the task's tiles, the cubic rows staged in
      // LDS
      adjoint_staged_tiles<WRT_ADJ_VPT, false, WRT_ADJ_PF>(
          (const WRT_CONST float4*)it.adirs, (const WRT_CONST double4*)it.atrig, it.andir,
          (const WRT_GLOBAL float4*)it.acub, it.ans, (float)it.as0, (float)(1.0 / it.ads),
          (float)it.ads, it.an, (const WRT_CONST int*)it.aorder, (long long)t.a, t.b,
          adjoint_chunk(it.ans, WRT_ADJ_PF),
          (WRT_LDS float4*)L.sb, (WRT_GLOBAL float*)it.aout);
    } else {
      // rows too long for LDS: tile a + w for wavefront w < b, gathers from memory
      const int w = threadIdx.x >> 6;
      if (w < t.b) {
        long long k[WRT_ADJ_VPT];
        bool valid[WRT_ADJ_VPT];
        adjoint_tile_voxels<WRT_ADJ_VPT>((long long)t.a + w, threadIdx.x & 63, it.an, k, valid);
        float r[WRT_ADJ_VPT];
        adjoint_voxels_f32<WRT_ADJ_VPT, false>(
            (const WRT_CONST float4*)it.adirs, (const WRT_CONST double4*)it.atrig, it.andir,
            (const WRT_GLOBAL float4*)it.acub, it.ans, (float)it.as0, (float)(1.0 / it.ads),
            (float)it.ads, it.an, k, r);
        WRT_GLOBAL float* const out = (WRT_GLOBAL float*)it.aout;
#pragma unroll
        for (int v = 0; v < WRT_ADJ_VPT; ++v)
          if (valid[v]) out[k[v]] = r[v];
      }
    }
  }
}

extern "C" int wrtk_batch_prologue(const wrt_batch_seg* segs, int nseg, int* counter,
                                   hipStream_t st) {
  if (nseg < 0 || nseg > 65535) return WRT_E_ARG;
  // with no segment, one block still resets the counter
  k_batch_prologue<<<dim3(64, (unsigned)(nseg > 0 ? nseg : 1)), 256, 0, st>>>(segs, nseg, counter);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

extern "C" int wrtk_batch_run(const wrt_batch_item* items, const wrt_batch_task* tasks,
                              int ntasks, int* counter, int grid, hipStream_t st) {
  if (ntasks <= 0) return WRT_OK;
  if (grid < 1) return WRT_E_ARG;
  k_batch<<<(unsigned)grid, FR_WAVES * 64, 0, st>>>(items, tasks, ntasks, counter);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

extern "C" int wrtk_batch_epilogue(const wrt_batch_seg* segs, int nseg, hipStream_t st) {
  if (nseg <= 0) return WRT_OK;
  if (nseg > 65535) return WRT_E_ARG;
  k_batch_epilogue<<<dim3(64, (unsigned)nseg), 256, 0, st>>>(segs);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}
