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
  const int* tasks;   /* [ntasks][2] = brick id, direction group (transpose:
                         2 * brick id + half brick)                           */
  int ntasks;
  const double* dirs; /* [ndir][9] n, a = delta*v1, b = delta*v2              */
  int ndir;
  int dgroup;         /* directions per group                                 */
  const double* shift; /* ascending (the kernel bisects it for a brick's shifts) */
  const int* sidx;    /* [nshift] output row of the j-th smallest shift: the
                         caller's grid may be in any order                    */
  int nshift;
  long long* acc;     /* [nshift][ndir] plane sums in 64-bit fixed point, units
                         of qstep (wrt_plane_qstep), zeroed by the caller      */
  double qmul;        /* delta^2 / qstep: a section adds rint(its FP32 sample
                         sum * qmul)                                          */
  int band_rows;      /* rows per band of the walk: 4, 8 or 16 (0: 8); the forward
                         per-job launch only, the others walk 8                */
  double rho2;        /* squared support radius: every sample farther from the
                         origin reads a zero cell (cell-centre radius of the
                         farthest non-zero cell plus half a cell diagonal)     */
  /* transpose only */
  const float* g;     /* [nshift][ndir] plane values (rows as acc's)          */
  double gscale;      /* fixed-point scale: a section adds rint(g * gscale)   */
  long long* part;    /* [ngroups][nb^3][WRT_BRICK^3] per-group brick partials,
                         fixed point                                          */
} wrt_radon_fast_args;

/* FP64 vol[x][y][z] -> the bricks of the brick rows bx in [bx0, bx0 + nbx) (vol holds the
 * planes x >= 32 bx0: the whole volume for bx0 = 0, nbx = nb); flags[brick] += its number of
 * non-zero values (flags zeroed by the caller); r2max[0] = largest squared cell-centre
 * radius of a non-zero cell (as the bit pattern of a non-negative double, so unsigned max
 * orders it), r2max[1] = largest |value| as stored in the bricks (the bits of a
 * non-negative float; infinity if a value is not finite in FP32), both zeroed by the caller */
int wrtk_brickify(const double* vol, int n, int nb, int bx0, int nbx, int bf16, void* bricks,
                  int* flags, unsigned long long* r2max, hipStream_t st);
/* boundary = 1: the tasks' bricks touch the cube's faces (cube test and edge clamp) */
int wrtk_radon_fast(const wrt_radon_fast_args* a, int boundary, hipStream_t st);
/* out[i] = (float)acc[i] */
int wrtk_acc_to_f32(const double* acc, float* out, size_t count, hipStream_t st);
/* out[i] = (float)(acc[i] * qstep): the fixed-point plane sums rounded to FP32 */
int wrtk_fix_to_f32(const long long* acc, double qstep, float* out, size_t count, hipStream_t st);

/* The fixed-point unit of a volume's plane sums: the power of two qstep >= 16 vmax 2^-50
 * (vmax the largest |brick value|).  A plane's sum is at most 16 vmax in magnitude (16 n^2
 * lattice samples of weight delta^2 = 1/n^2), so every section's and every plane's sum is
 * below 2^50 units: a section's value converts exactly through binary64 (|x| < 2^51) and the
 * 64-bit sums cannot overflow.  The unit is at most 2^-46 of the largest possible sum. */
static inline double wrt_plane_qstep(float vmax) {
  double q = 1.0;
  const double want = 16.0 * (double)vmax * (1.0 + 1.0 / 1048576.0) / 1125899906842624.0;
  if (!(want > 0.0)) return 1.0;  /* an all-zero volume */
  while (q < want) q *= 2.0;
  while (q * 0.5 >= want) q *= 0.5;
  return q;
}

/* Exact transpose of the plane integrals from FP32 plane values (a.g, a.gscale, a.part set;
 * a.rho2 >= 3 so that no row is clipped): every task (brick, half, group) stores its half
 * brick of 64-bit fixed-point sums to part[group][brick] (cells with z in [16 half, 16 half
 * + 16)).  The caller chooses gscale so that no cell's total over all groups exceeds 2^62. */
int wrtk_radon_T_fast(const wrt_radon_fast_args* a, int boundary, hipStream_t st);
/* out[x][y][z] = (float)(scale * sum over groups of part[group][brick][cell]): the integer
 * sum is exact, one rounding to FP64 and one to FP32 */
int wrtk_radon_T_reduce(const long long* part, int ngroups, int n, int nb, double scale,
                        float* out, hipStream_t st);

/* ---- the no-allocation plane integrals (k_noalloc.hip, host/wrt_noalloc.c) ---- */
#define WRT_NOALLOC_BYTES ((size_t)64 << 20) /* the fixed device scratch per GPU      */
#define WRT_NOALLOC_MAX_BRICKS 512           /* n <= 256 (the scratch binds first)   */
/* once per device, with the device current: the scratch's device address */
int wrtk_noalloc_setup(int device);
/* that address (NULL before setup) */
void* wrtk_noalloc_scratch(int device);
/* zero `words` 8-byte words */
int wrtk_noalloc_zero(void* p, size_t words, hipStream_t st);
/* the brick kernel's task list from the bricks' non-zero counts: interior tasks, then face
 * tasks; counts[0], counts[1] = their numbers */
int wrtk_noalloc_tasks(const int* flags, int n, int nb, int ngroups, int* tasks, int* counts,
                       hipStream_t st);
/* acc[i] (fixed point, units of qstep) -> the double (float)(acc[i] * qstep), in place */
int wrtk_noalloc_round(long long* acc, double qstep, size_t count, hipStream_t st);

/* ---- the persistent batch launch: many jobs' tasks in one grid ---- */
enum { WRT_ITEM_RADON = 0, WRT_ITEM_ADJOINT = 1 };
typedef struct {
  int kind;                 /* WRT_ITEM_RADON: brick tasks; WRT_ITEM_ADJOINT: voxel tasks  */
  int pow2;                 /* radon: n is a power of two                                 */
  double h, dc, inv_dc;     /* radon: n/2, RN(2/n), 1/RN(2/n)                            */
  wrt_radon_fast_args ra;   /* radon (tasks/ntasks unused)                               */
  const void* adirs;        /* adjoint: [ndir] float4 (n.x, n.y, n.z, weight)            */
  const void* atrig;        /* adjoint: [ndir] double4 (sin t, cos p, sin p, cos t)      */
  const void* acub;         /* adjoint: [ndir][ns-1] float4 cubics                       */
  float* aout;              /* adjoint: n^3 [z][x][y]                                    */
  const int* aorder;        /* adjoint: the tiles in launch order (wrt_adj_tile_order)     */
  int andir, ans, an;
  double as0, ads;          /* adjoint: uniform shift grid s_i = s0 + i*ds               */
} wrt_batch_item;

/* Backprojection voxels go to wavefronts in tiles of 4 (x) x 4 (y) x 4*vpt (z) voxels,
 * vpt voxels per thread (adjoint_voxel.hpp); the tile count of an n^3 volume: */
static inline long long wrt_adj_tiles(int n, int vpt) {
  const long long txy = (long long)((n + 3) / 4);
  return txy * txy * (long long)((n + 4 * vpt - 1) / (4 * vpt));
}
/* The tiles of an n^3 backprojection with vpt voxels per thread in launch order: first the
 * tiles with a voxel at radius 1 - 1e-4 or beyond (the general mode, which takes about 1.5
 * times the others' work), then the rest, each in index order.  The 16 tiles of a staged
 * workgroup wait for each other at every chunk of directions, so a workgroup of one kind
 * does not idle behind the other kind's work; the longer kind goes first. */
static inline void wrt_adj_tile_order(int n, int vpt, int* order) {
  const int ntx = (n + 3) / 4, ntz = (n + 4 * vpt - 1) / (4 * vpt);
  const double r2 = (1.0 - 1e-4) * (1.0 - 1e-4);
  long long k = 0;
  for (int pass = 0; pass < 2; ++pass)
    for (int tz = 0; tz < ntz; ++tz)
      for (int tx = 0; tx < ntx; ++tx)
        for (int ty = 0; ty < ntx; ++ty) {
          /* the largest |coordinate| of the tile's voxel centres along each axis (indices
           * clamped into the volume, as the kernel clamps them) */
          double m2 = 0.0;
          const int lo[3] = {4 * tx, 4 * ty, 4 * vpt * tz}, len[3] = {4, 4, 4 * vpt};
          for (int c = 0; c < 3; ++c) {
            const int a = lo[c] < n - 1 ? lo[c] : n - 1;
            const int b = lo[c] + len[c] - 1 < n - 1 ? lo[c] + len[c] - 1 : n - 1;
            const double xa = (-1.0) + (1.0 / n) + a * (2.0 / n);
            const double xb = (-1.0) + (1.0 / n) + b * (2.0 / n);
            m2 += xa * xa > xb * xb ? xa * xa : xb * xb;
          }
          if ((m2 <= r2) == (pass == 1)) order[k++] = (tz * ntx + tx) * ntx + ty;
        }
}

/* voxels per thread of a backprojection task in the persistent kernel; a task is up to one
 * tile per wavefront of its 1024-thread workgroup, which stages the directions' cubic rows
 * in LDS in chunks of WRT_ADJ_PF float4 per thread (4096 float4; 2 voxels and 4 float4 per
 * thread keep the kernel within 128 registers, with a few values spilled and reloaded only
 * outside the sample loops) */
#define WRT_ADJ_VPT 2
#define WRT_ADJ_PF 4
#define WRT_ADJ_TASK_TILES 16

/* radon: (item, brick, direction group, 1 if the brick touches a cube face);
 * adjoint: (item, first tile, tile count <= WRT_ADJ_TASK_TILES, 0) */
typedef struct {
  int item, a, b, c;
} wrt_batch_task;

/* the fixed-point plane sums of a radon item and its FP32 output */
typedef struct {
  long long* acc;
  float* out;
  long long count;
  double qstep;
} wrt_batch_seg;

/* zero every segment's accumulator and the task counter */
int wrtk_batch_prologue(const wrt_batch_seg* segs, int nseg, int* counter, hipStream_t st);
/* grid workgroups of 1024 threads take tasks from *counter until none is left */
int wrtk_batch_run(const wrt_batch_item* items, const wrt_batch_task* tasks, int ntasks,
                   int* counter, int grid, hipStream_t st);
/* round every segment's accumulator to its FP32 output */
int wrtk_batch_epilogue(const wrt_batch_seg* segs, int nseg, hipStream_t st);

#ifdef __cplusplus
}
#endif
#endif
