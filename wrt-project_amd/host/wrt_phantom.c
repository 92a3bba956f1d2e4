/*
 * wrt_phantom.c -- seeded synthetic inputs (no reference counterpart): ellipsoid phantoms
 * and Gaussian data.  Random numbers come from splitmix64, so a (seed, index) pair gives
 * the same value on every host.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "../../include/wrt.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static uint64_t sm64_next(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static uint64_t sm64_hash(uint64_t a) {
  uint64_t s = a;
  return sm64_next(&s);
}
static double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

int wrt_phantom_params(int nell, unsigned long long seed, double* P) {
  if (nell < 0 || (nell > 0 && !P)) return WRT_E_ARG;
  uint64_t s = (uint64_t)seed ^ 0x5DEECE66DULL;
  for (int e = 0; e < nell; ++e) {
    double* p = P + 16 * e;
    double c[3];
    do {
      for (int k = 0; k < 3; ++k) c[k] = u01(sm64_next(&s)) - 0.5;
    } while (c[0] * c[0] + c[1] * c[1] + c[2] * c[2] >= 0.25);
    for (int k = 0; k < 3; ++k) p[k] = c[k];
    for (int k = 0; k < 3; ++k) p[3 + k] = 0.05 + 0.35 * u01(sm64_next(&s));
    /* uniform random rotation from a uniform unit quaternion (Shoemake) */
    const double u1 = u01(sm64_next(&s)), u2 = u01(sm64_next(&s)), u3 = u01(sm64_next(&s));
    const double a = sqrt(1.0 - u1), b = sqrt(u1);
    const double qx = a * sin(2 * M_PI * u2), qy = a * cos(2 * M_PI * u2);
    const double qz = b * sin(2 * M_PI * u3), qw = b * cos(2 * M_PI * u3);
    double* R = p + 6;
    R[0] = 1 - 2 * (qy * qy + qz * qz);
    R[1] = 2 * (qx * qy - qz * qw);
    R[2] = 2 * (qx * qz + qy * qw);
    R[3] = 2 * (qx * qy + qz * qw);
    R[4] = 1 - 2 * (qx * qx + qz * qz);
    R[5] = 2 * (qy * qz - qx * qw);
    R[6] = 2 * (qx * qz - qy * qw);
    R[7] = 2 * (qy * qz + qx * qw);
    R[8] = 1 - 2 * (qx * qx + qy * qy);
    p[15] = -0.5 + 1.5 * u01(sm64_next(&s));
  }
  return WRT_OK;
}

static double gauss_at(uint64_t seed, uint64_t i) {
  const uint64_t base = sm64_hash(seed ^ 0xA0761D6478BD642FULL);
  const double u1 = 1.0 - u01(sm64_hash(base + 2 * i));
  const double u2 = u01(sm64_hash(base + 2 * i + 1));
  return sqrt(-2.0 * log(u1)) * cos(2 * M_PI * u2);
}

int wrt_gaussian(size_t count, unsigned long long seed, double* out) {
  if (!out && count) return WRT_E_ARG;
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)count; ++i) out[i] = gauss_at(seed, (uint64_t)i);
  return WRT_OK;
}

int wrt_phantom(int n, int cell_centred, int nell, unsigned long long seed, double sigma,
                double* vol) {
  if (n < 2 || nell < 0 || nell > 1000 || !vol) return WRT_E_ARG;
  double* P = (double*)malloc(sizeof(double) * 16 * (size_t)(nell ? nell : 1));
  if (!P) return WRT_E_NOMEM;
  wrt_phantom_params(nell, seed, P);
  double* coord = (double*)malloc(sizeof(double) * (size_t)n);
  if (!coord) { free(P); return WRT_E_NOMEM; }
  for (int i = 0; i < n; ++i)
    coord[i] = cell_centred ? (-1.0 + (2.0 * i + 1.0) / n) : (-1.0 + 2.0 * i / (n - 1));
  const uint64_t tseed = (uint64_t)seed * 0x9E3779B97F4A7C15ULL + 1;
#pragma omp parallel for schedule(static)
  for (int ix = 0; ix < n; ++ix) {
    const double x = coord[ix];
    for (int iy = 0; iy < n; ++iy) {
      const double y = coord[iy];
      double* row = vol + ((size_t)ix * n + iy) * n;
      for (int iz = 0; iz < n; ++iz) {
        const double z = coord[iz];
        double v = 0.0;
        if (x * x + y * y + z * z <= 1.0) {
          for (int e = 0; e < nell; ++e) {
            const double* p = P + 16 * e;
            const double dx = x - p[0], dy = y - p[1], dz = z - p[2];
            const double* R = p + 6;
            const double u = (R[0] * dx + R[3] * dy + R[6] * dz) / p[3];
            const double w = (R[1] * dx + R[4] * dy + R[7] * dz) / p[4];
            const double t = (R[2] * dx + R[5] * dy + R[8] * dz) / p[5];
            if (u * u + w * w + t * t <= 1.0) v += p[15];
          }
          if (sigma > 0) v += sigma * gauss_at(tseed, ((uint64_t)ix * n + iy) * n + iz);
        }
        row[iz] = v;
      }
    }
  }
  free(coord);
  free(P);
  return WRT_OK;
}
