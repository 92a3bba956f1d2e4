// k_spectral.hip -- the frequency-domain factor of the spectral Laplacian.
//
// laplace2dfft.m / laplace3dfft.m (radon_inversion/backprojection/) multiply the DFT of the
// image by U = -(2 pi / period)^2 * (sum over the axes of k^2), where the signed frequency of
// index i is i for i <= n/2 and i - n otherwise.  Here the DFT is the real-to-complex one: the
// last axis keeps indices 0..n/2 (all <= n/2, so their frequency is the index).  The 1/n^d
// normalisation of the inverse transform is folded into the same multiply.
#include "wrt_common.hpp"
#include "wrt_dev.h"

__device__ __forceinline__ double sfreq(long long i, int n) {
  return (double)(2 * i <= n ? i : i - n);
}

// spec: [n]..[n][n/2+1] complex (dims = 2 or 3), multiplied in place by U * scale
__global__ __launch_bounds__(256) void k_laplace_factor(double2* __restrict__ spec, int n, int dims,
                                                        double coef, double scale) {
  const long long half = n / 2 + 1;
  const long long total = (dims == 3 ? (long long)n * n : (long long)n) * half;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const long long l = t % half;
    const long long r = t / half;
    double k2 = (double)l * (double)l;
    if (dims == 3) {
      const double fj = sfreq(r % n, n), fi = sfreq(r / n, n);
      k2 += fi * fi + fj * fj;
    } else {
      const double fi = sfreq(r, n);
      k2 += fi * fi;
    }
    const double m = coef * k2 * scale;
    double2 v = spec[t];
    v.x *= m;
    v.y *= m;
    spec[t] = v;
  }
}

extern "C" int wrtk_laplace_factor(void* spec, int n, int dims, double coef, double scale,
                                   hipStream_t st) {
  if (n < 1 || (dims != 2 && dims != 3)) return WRT_E_ARG;
  const long long total = (dims == 3 ? (long long)n * n : (long long)n) * (n / 2 + 1);
  long long blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  k_laplace_factor<<<(unsigned)blocks, 256, 0, st>>>((double2*)spec, n, dims, coef, scale);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}
