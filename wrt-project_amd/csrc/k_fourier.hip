// k_fourier.hip -- Fourier-slice (Chang-type) inversion of Radon data in 3D and 2D, FP64.
//
// rtft3d.m (radon_inversion/rtft3d.m:3-137) Fourier transforms every direction's Radon data
// along s (zero-padded to ntotal = n points), multiplies by exp(2 pi i rsupp k dfreq) and
// fftshift-s; with the jacobian weight 0.5 dshift dfreq dphi w_theta rho_m^2 this gives the
// coefficient a[d][m] of the node rho_m u_d, rho_m = (m - n/2) dfreq.  nfft_reconstruct_3d.m
// (:19-53) then evaluates, by an adjoint NFFT,
//   f(p) = Re sum_d sum_m a[d][m] exp(2 pi i rho_m (p . u_d))
// at p = k * 2/ngrid, k in [-ngrid/2, ngrid/2)^3, and chang3d.m (:3-30) divides by w0.
// In 2D (rtft2d.m:3-62, nfft_reconstruct_2d.m:19-45, chang2d.m:3-26) the directions are
// (cos phi, sin phi), the weight 0.5 dshift dfreq dphi |rho_m| and the grid ngrid^2; the
// kernels take the radial factor (rho^2 or |rho|) as a table and the grid as nx x ny x nz.
//
// Only the real part is kept, and Re(a e^{-i psi}) = Re(conj(a) e^{i psi}), so the nodes at
// -rho and +rho fold into one coefficient: with z = exp(2 pi i dfreq t), t = p . u_d,
//   Re sum_m a_m exp(2 pi i rho_m t) = Re sum_{j=0}^{n/2} c_j z^j,
//   c_0 = a_{n/2},  c_j = a_{n/2+j} + conj(a_{n/2-j}) (0 < j < n/2),  c_{n/2} = conj(a_0),
// an identity (no symmetry of the data is assumed).  The polynomial is evaluated exactly by
// Horner's scheme, 4 FP64 multiply-adds per term, instead of the NFFT's approximation.
//
// Kernels: k_rtft -- one workgroup per direction: a direct DFT of its (at most n) shift samples
// with the exact twiddle table of n points, the shift correction, the weight and the fold;
// k_fourier_adjoint -- each thread evaluates 4 voxels along z (4 independent Horner chains
// sharing every coefficient load) over one slice of the directions; the slices' partial sums
// are added in a fixed order by k_fourier_sum, so the result does not depend on scheduling.
#include "wrt_common.hpp"
#include "wrt_dev.h"

// weighted, shift-corrected coefficient of the node m (fftshift: m holds frequency index k)
__device__ __forceinline__ double2 rtft_value(const double* __restrict__ sig, int ns, int m, int n,
                                              const double2* __restrict__ tw,
                                              const double2* __restrict__ corr,
                                              const double* __restrict__ rw, double wd) {
  const int k = (m + n / 2) % n;
  double re = 0.0, im = 0.0;
  int q = 0;  // (j * k) mod n, stepped exactly
  for (int j = 0; j < ns; ++j) {
    const double2 w = tw[q];
    re += sig[j] * w.x;
    im += sig[j] * w.y;
    q += k;
    if (q >= n) q -= n;
  }
  const double2 c = corr[k];
  const double g = wd * rw[m];
  return make_double2(g * (re * c.x - im * c.y), g * (re * c.y + im * c.x));
}

// radon [nshift][ndir]; tw [n] = exp(-2 pi i q / n); corr [n] = exp(2 pi i rsupp k dfreq);
// jw [ndir] = 0.5 dshift dfreq dphi (w_theta); rw [n] = rho_m^2 (3D) or |rho_m| (2D);
// out coef [ndir][n/2 + 1] (folded)
__global__ __launch_bounds__(256) void k_rtft(const double* __restrict__ radon, int nshift, int ndir,
                                              const double2* __restrict__ tw,
                                              const double2* __restrict__ corr,
                                              const double* __restrict__ jw,
                                              const double* __restrict__ rw, int n,
                                              double2* __restrict__ coef) {
  extern __shared__ double sig[];
  const int d = blockIdx.x;
  const int ns = min(nshift, n);  // fft(x, n) keeps the first n samples
  for (int j = threadIdx.x; j < ns; j += blockDim.x) sig[j] = radon[(size_t)j * ndir + d];
  __syncthreads();
  const double wd = jw[d];
  const int h = n / 2;
  for (int j = threadIdx.x; j <= h; j += blockDim.x) {
    double2 c;
    if (j == 0) {
      c = rtft_value(sig, ns, h, n, tw, corr, rw, wd);
    } else {
      const double2 lo = rtft_value(sig, ns, h - j, n, tw, corr, rw, wd);
      c = make_double2(lo.x, -lo.y);
      if (j < h) {
        const double2 hi = rtft_value(sig, ns, h + j, n, tw, corr, rw, wd);
        c.x += hi.x;
        c.y += hi.y;
      }
    }
    coef[(size_t)d * (h + 1) + j] = c;
  }
}

extern "C" int wrtk_rtft(const double* radon, int nshift, int ndir, const void* tw, const void* corr,
                         const double* jw, const double* rw, int n, void* coef, hipStream_t st) {
  if (nshift < 1 || ndir < 1 || n < 2 || (n & 1)) return WRT_E_ARG;
  const size_t lds = (size_t)(nshift < n ? nshift : n) * sizeof(double);
  if (lds > 65536) return WRT_E_ARG;
  k_rtft<<<ndir, 256, lds, st>>>(radon, nshift, ndir, (const double2*)tw, (const double2*)corr, jw,
                                 rw, n, (double2*)coef);
  WRT_LAUNCH_CHECK();
  return WRT_OK;
}

#define FA_VPT 4  // voxels per thread, consecutive in z
#define FA_UNROLL 8

// coef [ndir][nc] folded; dirs [ndir][4] = u (3) and pad; directions [d0, d1) of the slice;
// out [nx][ny][nz] at p = ((ix, iy, iz) - (nx, ny, nz)/2) * 2/ngrid (2D: nx = 1, u.x = 0):
// scale * the slice's sum (partial sums when scale = 1)
__global__ __launch_bounds__(256) void k_fourier_adjoint(const double2* __restrict__ coef, int nc,
                                                         double dfreq,
                                                         const double4* __restrict__ dirs,
                                                         int ndir, int nslice, int nx, int ny,
                                                         int nz, int ngrid, double scale,
                                                         double* __restrict__ out) {
  const int nzq = (nz + FA_VPT - 1) / FA_VPT;
  const long long tq = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nq = (long long)nx * ny * nzq;
  if (tq >= nq) return;
  const int s = blockIdx.y;
  const int d0 = (int)((long long)ndir * s / nslice), d1 = (int)((long long)ndir * (s + 1) / nslice);
  const int izq = (int)(tq % nzq), iy = (int)((tq / nzq) % ny), ix = (int)(tq / ((long long)nzq * ny));
  const double dx = 2.0 / ngrid;
  const double px = (ix - nx / 2) * dx, py = (iy - ny / 2) * dx;
  double pz[FA_VPT], acc[FA_VPT];
#pragma unroll
  for (int k = 0; k < FA_VPT; ++k) {
    pz[k] = (izq * FA_VPT + k - nz / 2) * dx;
    acc[k] = 0.0;
  }
  const double w1 = 2.0 * M_PI * dfreq;  // z = exp(i w1 t)
  for (int d = d0; d < d1; ++d) {
    const double4 u = dirs[d];
    const double txy = px * u.x + py * u.y;
    double zc[FA_VPT], zs[FA_VPT], hr[FA_VPT], hi[FA_VPT];
    const double2* __restrict__ a = coef + (size_t)d * nc;
    const double2 top = a[nc - 1];
#pragma unroll
    for (int k = 0; k < FA_VPT; ++k) {
      sincos(w1 * (txy + pz[k] * u.z), &zs[k], &zc[k]);
      hr[k] = top.x;
      hi[k] = top.y;
    }
    int m = nc - 2;
    for (; m >= FA_UNROLL - 1; m -= FA_UNROLL) {
      double2 am[FA_UNROLL];
#pragma unroll
      for (int r = 0; r < FA_UNROLL; ++r) am[r] = a[m - r];
#pragma unroll
      for (int r = 0; r < FA_UNROLL; ++r)
#pragma unroll
        for (int k = 0; k < FA_VPT; ++k) {
          const double nr = fma(hr[k], zc[k], fma(-hi[k], zs[k], am[r].x));
          hi[k] = fma(hr[k], zs[k], fma(hi[k], zc[k], am[r].y));
          hr[k] = nr;
        }
    }
    for (; m >= 0; --m) {
      const double2 am = a[m];
#pragma unroll
      for (int k = 0; k < FA_VPT; ++k) {
        const double nr = fma(hr[k], zc[k], fma(-hi[k], zs[k], am.x));
        hi[k] = fma(hr[k], zs[k], fma(hi[k], zc[k], am.y));
        hr[k] = nr;
      }
    }
#pragma unroll
    for (int k = 0; k < FA_VPT; ++k) acc[k] += hr[k];
  }
  double* o = out + (size_t)s * nx * ny * nz + ((size_t)ix * ny + iy) * nz;
#pragma unroll
  for (int k = 0; k < FA_VPT; ++k)
    if (izq * FA_VPT + k < nz) o[izq * FA_VPT + k] = acc[k] * scale;
}

// out[v] = scale * sum_s part[s][v], slices in order
__global__ __launch_bounds__(256) void k_fourier_sum(const double* __restrict__ part, int nslice,
                                                     long long nvox, double scale,
                                                     double* __restrict__ out) {
  const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvox) return;
  double acc = 0.0;
  for (int s = 0; s < nslice; ++s) acc += part[(size_t)s * nvox + v];
  out[v] = acc * scale;
}

// part: nslice * nx * ny * nz doubles when nslice > 1 (unused otherwise)
extern "C" int wrtk_fourier_adjoint(const void* coef, int n, double dfreq, const void* dirs, int ndir,
                                    int nx, int ny, int nz, int ngrid, int nslice, double inv_w0,
                                    double* part, double* out, hipStream_t st) {
  const long long nvox = (long long)nx * ny * nz;
  if (nx < 1 || ny < 1 || nz < 1 || ngrid < 1 || n < 2 || (n & 1) || ndir < 1 || nslice < 1 ||
      nslice > ndir || nslice > 65535 || (nslice > 1 && !part))
    return WRT_E_ARG;
  const long long nq = (long long)nx * ny * ((nz + FA_VPT - 1) / FA_VPT);
  const long long blocks = (nq + 255) / 256;
  if (blocks >= (1LL << 31)) return WRT_E_ARG;
  const dim3 grid((unsigned)blocks, (unsigned)nslice);
  k_fourier_adjoint<<<grid, 256, 0, st>>>((const double2*)coef, n / 2 + 1, dfreq,
                                          (const double4*)dirs, ndir, nslice, nx, ny, nz, ngrid,
                                          nslice > 1 ? 1.0 : inv_w0, nslice > 1 ? part : out);
  WRT_LAUNCH_CHECK();
  if (nslice > 1) {
    k_fourier_sum<<<(unsigned)((nvox + 255) / 256), 256, 0, st>>>(part, nslice, nvox, inv_w0, out);
    WRT_LAUNCH_CHECK();
  }
  return WRT_OK;
}
