/*
 * wrt_fast.h -- launch interface of the fast (FP32 / BF16-brick) kernels.
 *
 * Brick layout in HBM: the volume is cut into WRT_BRICK^3 bricks, brick-major
 * (brick id = (bx*nb + by)*nb + bz, nb = ceil(n / WRT_BRICK)), each brick row-major
 * [x][y][z] inside, zero-padded past the volume edge.  Values are FP32 or BF16 (the upper
 * 16 bits of the FP32 pattern, rounded to nearest even).
 */
#ifndef WRT_FAST_H
#define WRT_FAST_H

#include <hip/hip_runtime_api.h>

#define WRT_BRICK 32

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  const void* bricks; /* brick-major volume, FP32 or BF16                     */
  int n;              /* volume edge                                          */
  int nb;             /* bricks per axis                                      */
  int bf16;           /* 1: bricks hold BF16                                  */
  const int* tasks;   /* [ntasks][2] = brick id, direction group              */
  int ntasks;
  const double* dirs; /* [ndir][9] n, a = delta*v1, b = delta*v2              */
  int ndir;
  int dgroup;         /* directions per group                                 */
  const double* shift;
  int nshift;
  double* acc;        /* [nshift][ndir] FP64 accumulators, zeroed by caller   */
  double rho2;        /* squared support radius: every sample farther from the
                         origin reads a zero cell (cell-centre radius of the
                         farthest non-zero cell plus half a cell diagonal)     */
} wrt_radon_fast_args;

/* FP64 vol[x][y][z] -> bricks; flags[brick] += its number of non-zero values (flags zeroed by the caller);
 * r2max = largest squared cell-centre radius of a non-zero cell (as the bit pattern of a
 * non-negative double, so unsigned max orders it), zeroed by the caller */
int wrtk_brickify(const double* vol, int n, int nb, int bf16, void* bricks, int* flags,
                  unsigned long long* r2max, hipStream_t st);
/* boundary = 1: the tasks' bricks touch the cube's faces (cube test and edge clamp) */
int wrtk_radon_fast(const wrt_radon_fast_args* a, int boundary, hipStream_t st);
/* out[i] = (float)acc[i] */
int wrtk_acc_to_f32(const double* acc, float* out, size_t count, hipStream_t st);

#ifdef __cplusplus
}
#endif
#endif
