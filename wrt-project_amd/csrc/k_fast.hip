// k_fast.hip -- per-job launches of the fast plane integrals (one workgroup per task of the
// job's list), the brick conversion and the accumulator rounding.  The brick task itself is
// in radon_brick.hpp.
#include "radon_brick.hpp"

template <bool BND, bool BF16, bool POW2, int BR>
__global__ __launch_bounds__(FR_WAVES * 64) void k_radon_fast(wrt_radon_fast_args a, double h,
                                                              double dc, double inv_dc) {
  __shared__ radon_lds L;
  radon_brick_task<false, BR>(a, a.tasks[2 * blockIdx.x], a.tasks[2 * blockIdx.x + 1], h, dc,
                              inv_dc, L, BND, BF16, POW2);
}

template <bool BND, bool BF16, int BR>
static int launch_pow2(const wrt_radon_fast_args* a, bool pow2, hipStream_t st) {
  const double n = a->n;
  const double dc = 2.0 / a->n;
  const double inv = 1.0 / dc;
  const double h = n / 2.0;
  if (pow2)
    k_radon_fast<BND, BF16, true, BR><<<a->ntasks, FR_WAVES * 64, 0, st>>>(*a, h, dc, inv);
  else
    k_radon_fast<BND, BF16, false, BR><<<a->ntasks, FR_WAVES * 64, 0, st>>>(*a, h, dc, inv);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

template <int BR>
static int launch_rows(const wrt_radon_fast_args* a, int boundary, hipStream_t st) {
  const bool pow2 = (a->n & (a->n - 1)) == 0;
  if (boundary)
    return a->bf16 ? launch_pow2<true, true, BR>(a, pow2, st) : launch_pow2<true, false, BR>(a, pow2, st);
  return a->bf16 ? launch_pow2<false, true, BR>(a, pow2, st) : launch_pow2<false, false, BR>(a, pow2, st);
}

extern "C" int wrtk_radon_fast(const wrt_radon_fast_args* a, int boundary, hipStream_t st) {
  if (a->ntasks <= 0) return WRT_OK;
  switch (a->band_rows) {
    case 0:
    case 8: return launch_rows<8>(a, boundary, st);
    case 4: return launch_rows<4>(a, boundary, st);
    case 16: return launch_rows<16>(a, boundary, st);
    default: return WRT_E_ARG;
  }
}

// ------------------------------------------------- transpose of the plane integrals ---
// One workgroup per (brick, half brick, direction group): the brick task in transpose mode
// adds every section's fixed-point value into the half brick's 64-bit LDS accumulators,
// which are then stored whole to the group's slot in the brick's own [x][y][z] order.  No
// global atomics, and every sum is an exact integer sum: the result is the same bits for
// any schedule.
template <bool BND, bool POW2>
__global__ __launch_bounds__(FR_WAVES * 64) void k_radon_T_fast(wrt_radon_fast_args a, double h,
                                                                double dc, double inv_dc) {
  __shared__ radon_T_lds L;
  const int bh = a.tasks[2 * blockIdx.x], grp = a.tasks[2 * blockIdx.x + 1];
  const int brick = bh >> 1, zh = bh & 1;
  radon_brick_task<true>(a, brick, grp, h, dc, inv_dc, L, BND, false, POW2, zh);
  __syncthreads();
  const size_t nbr = (size_t)a.nb * a.nb * a.nb;
  long long* __restrict__ dst = a.part + ((size_t)grp * nbr + brick) * (FB * FB * FB) + FT_Z * zh;
  for (int i = threadIdx.x; i < FB * FB * FT_Z; i += blockDim.x) {
    const int lx = i >> 9, ly = (i >> 4) & 31, lz = i & 15;
    dst[(lx * FB + ly) * FB + lz] = (long long)L.acc[lx * FT_PX + ly * FT_PY + lz];
  }
}

template <bool BND>
static int launch_T(const wrt_radon_fast_args* a, bool pow2, hipStream_t st) {
  const double dc = 2.0 / a->n, inv = 1.0 / dc, h = a->n / 2.0;
  if (pow2)
    k_radon_T_fast<BND, true><<<a->ntasks, FR_WAVES * 64, 0, st>>>(*a, h, dc, inv);
  else
    k_radon_T_fast<BND, false><<<a->ntasks, FR_WAVES * 64, 0, st>>>(*a, h, dc, inv);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

extern "C" int wrtk_radon_T_fast(const wrt_radon_fast_args* a, int boundary, hipStream_t st) {
  if (a->ntasks <= 0) return WRT_OK;
  if (!a->g || !a->part || a->rho2 < 3.0 || !(a->gscale > 0.0)) return WRT_E_ARG;
  const bool pow2 = (a->n & (a->n - 1)) == 0;
  return boundary ? launch_T<true>(a, pow2, st) : launch_T<false>(a, pow2, st);
}

__global__ __launch_bounds__(256) void k_radon_T_reduce(const long long* __restrict__ part,
                                                        int ngroups, int n, int nb,
                                                        double scale, float* __restrict__ out) {
  const size_t total = (size_t)n * n * n;
  const size_t slot = (size_t)nb * nb * nb * (FB * FB * FB);
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (size_t)gridDim.x * blockDim.x) {
    const int z = (int)(k % n), y = (int)((k / n) % n), x = (int)(k / ((size_t)n * n));
    const size_t brick = ((size_t)(x / FB) * nb + y / FB) * nb + z / FB;
    const size_t cell = brick * (FB * FB * FB) + ((x % FB) * FB + y % FB) * FB + z % FB;
    unsigned long long acc = 0ull;  // two's complement: wrapping adds give the exact sum
    for (int g = 0; g < ngroups; ++g) acc += (unsigned long long)part[(size_t)g * slot + cell];
    out[k] = (float)((double)(long long)acc * scale);
  }
}

extern "C" int wrtk_radon_T_reduce(const long long* part, int ngroups, int n, int nb,
                                   double scale, float* out, hipStream_t st) {
  const size_t total = (size_t)n * n * n;
  size_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (!blocks || ngroups < 1) return WRT_E_ARG;
  k_radon_T_reduce<<<(unsigned)blocks, 256, 0, st>>>(part, ngroups, n, nb, scale, out);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

// ------------------------------------------------------------------- brick loader ---
__device__ __forceinline__ unsigned short f32_to_bf16_rne(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// bricks of the rows bx in [bx0, bx0 + nbx); vol holds the planes x >= 32 bx0
__global__ __launch_bounds__(256) void k_brickify(const double* __restrict__ vol, int n, int nb,
                                                  int bx0, int nbx, int bf16,
                                                  void* __restrict__ bricks,
                                                  int* __restrict__ flags,
                                                  unsigned long long* __restrict__ r2max) {
  const size_t first = (size_t)bx0 * nb * nb * (FB * FB * FB);
  const size_t total = first + (size_t)nbx * nb * nb * (FB * FB * FB);
  const int x0 = bx0 * FB;
  double r2 = 0.0;
  float vmax = 0.0f;
  for (size_t i = first + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t brick = i / (FB * FB * FB);
    const int li = (int)(i % (FB * FB * FB));
    const int bx = (int)(brick / ((size_t)nb * nb)), by = (int)((brick / nb) % nb), bz = (int)(brick % nb);
    const int x = bx * FB + (li >> 10), y = by * FB + ((li >> 5) & 31), z = bz * FB + (li & 31);
    double v = 0.0;
    if (x < n && y < n && z < n) v = vol[((size_t)(x - x0) * n + y) * n + z];
    const float f = (float)v;
    float fa;
    if (bf16) {
      const unsigned short h = f32_to_bf16_rne(f);
      ((unsigned short*)bricks)[i] = h;
      fa = fabsf(__uint_as_float((unsigned)h << 16));
    } else {
      ((float*)bricks)[i] = f;
      fa = fabsf(f);
    }
    // a NaN counts as infinite, so that the caller sees every non-finite value
    vmax = fmaxf(vmax, fa <= 3.4028234663852886e38f ? fa : __builtin_inff());
    // one count per wavefront: the 64 lanes of a wave always share a brick
    const unsigned long long nz = __ballot(v != 0.0);
    if ((threadIdx.x & 63) == 0 && nz) atomicAdd(&flags[brick], (int)__popcll(nz));
    if (v != 0.0) {
      const double cx = -1.0 + (2.0 * x + 1.0) / n, cy = -1.0 + (2.0 * y + 1.0) / n,
                   cz = -1.0 + (2.0 * z + 1.0) / n;
      r2 = fmax(r2, cx * cx + cy * cy + cz * cz);
    }
  }
  // non-negative doubles order like their bit patterns
  unsigned long long b = (unsigned long long)__double_as_longlong(r2);
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_down(b, off, 64);
    b = o > b ? o : b;
  }
  if ((threadIdx.x & 63) == 0 && b) atomicMax(r2max, b);
  // and the largest |value| (non-negative floats order like their bit patterns too)
  unsigned vb = __float_as_uint(vmax);
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned o = __shfl_down(vb, off, 64);
    vb = o > vb ? o : vb;
  }
  if ((threadIdx.x & 63) == 0 && vb) atomicMax(r2max + 1, (unsigned long long)vb);
}

extern "C" int wrtk_brickify(const double* vol, int n, int nb, int bx0, int nbx, int bf16,
                             void* bricks, int* flags, unsigned long long* r2max, hipStream_t st) {
  if (bx0 < 0 || nbx < 1 || bx0 + nbx > nb) return WRT_E_ARG;
  const size_t total = (size_t)nbx * nb * nb * (FB * FB * FB);
  size_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  k_brickify<<<(unsigned)blocks, 256, 0, st>>>(vol, n, nb, bx0, nbx, bf16, bricks, flags, r2max);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

__global__ void k_acc_to_f32(const double* __restrict__ acc, float* __restrict__ out, size_t count) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = (float)acc[i];
}

__global__ void k_fix_to_f32(const long long* __restrict__ acc, double qstep,
                             float* __restrict__ out, size_t count) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = (float)((double)acc[i] * qstep);  // exact: |acc| < 2^53, qstep a power of two
}

extern "C" int wrtk_fix_to_f32(const long long* acc, double qstep, float* out, size_t count,
                               hipStream_t st) {
  if (count == 0) return WRT_OK;
  size_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  k_fix_to_f32<<<(unsigned)blocks, 256, 0, st>>>(acc, qstep, out, count);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

extern "C" int wrtk_acc_to_f32(const double* acc, float* out, size_t count, hipStream_t st) {
  size_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (!blocks) return WRT_OK;
  k_acc_to_f32<<<(unsigned)blocks, 256, 0, st>>>(acc, out, count);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}
