// Grid interpolation shared by the particle kernels and their adjoints, and
// the XCD-contiguous block mapping of the gather kernels.
#pragma once
#include "mpf_internal.h"

namespace mpf {

// Logical block index with the hardware's round-robin dispatch of blocks over
// the 8 XCDs undone: XCD x runs the contiguous range of logical blocks
// [x q + min(x, r), ...) (q, r = grid size / 8, remainder), so the neighbour
// cells and particles that adjacent blocks share are read into one XCD's L2.
// A bijection of [0, gridDim.x).
__device__ __forceinline__ int64_t xcd_block() {
  const int64_t nb = gridDim.x, b = blockIdx.x;
  const int64_t q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// trilinear, spray_k.cu:111-122
__device__ __forceinline__ float trilinear(const float* __restrict__ u, float x, float y, float z, int w, int hh,
                                           int d) {
  x = fmaxf(0.0f, fminf(x, (float)w));
  y = fmaxf(0.0f, fminf(y, (float)hh));
  z = fmaxf(0.0f, fminf(z, (float)d));
  int i = (int)fminf(x, (float)(w - 2));
  int j = (int)fminf(y, (float)(hh - 2));
  int k = (int)fminf(z, (float)(d - 2));
  const int64_t sj = d, si = (int64_t)hh * d;
  const float* b = u + i * si + j * sj + k;
  float ix1 = (float)(i + 1) - x, ix0 = x - (float)i;
  float jy1 = (float)(j + 1) - y, jy0 = y - (float)j;
  float kz1 = (float)(k + 1) - z, kz0 = z - (float)k;
  return kz1 * (jy1 * (ix1 * b[0] + ix0 * b[si]) + jy0 * (ix1 * b[sj] + ix0 * b[si + sj])) +
         kz0 * (jy1 * (ix1 * b[1] + ix0 * b[si + 1]) + jy0 * (ix1 * b[sj + 1] + ix0 * b[si + sj + 1]));
}
// getParticleVelFromGrid, spray_k.cu:227-240
__device__ __forceinline__ f3 vel_from_grid(const Grid& g, f3 p, const float* ux, const float* uy, const float* uz) {
  float x = (p.x - g.gx0) / g.h, y = (p.y - g.gy0) / g.h, z = (p.z - g.gz0) / g.h;
  f3 v;
  v.x = trilinear(ux, x, y - 0.5f, z - 0.5f, g.nx + 1, g.ny, g.nz);
  v.y = trilinear(uy, x - 0.5f, y, z - 0.5f, g.nx, g.ny + 1, g.nz);
  v.z = trilinear(uz, x - 0.5f, y - 0.5f, z, g.nx, g.ny, g.nz + 1);
  return v;
}
// getScaleFromFrid, spray_k.cu:139-151
__device__ __forceinline__ float scalar_from_grid(const Grid& g, f3 p, const float* f) {
  float x = (p.x - g.gx0) / g.h, y = (p.y - g.gy0) / g.h, z = (p.z - g.gz0) / g.h;
  return trilinear(f, x - 0.5f, y - 0.5f, z - 0.5f, g.nx, g.ny, g.nz);
}


// Trilinear value and its gradient with respect to (x, y, z) in grid units,
// following the clamps of trilinear(): a clamped coordinate has zero
// derivative; the base index is held fixed (piecewise-linear derivative).
__device__ __forceinline__ double trilinear_grad(const float* __restrict__ u, double x, double y, double z, int w,
                                                 int hh, int d, double& gx, double& gy, double& gz) {
  const bool cx = x < 0.0 || x > (double)w, cy = y < 0.0 || y > (double)hh, cz = z < 0.0 || z > (double)d;
  x = fmax(0.0, fmin(x, (double)w));
  y = fmax(0.0, fmin(y, (double)hh));
  z = fmax(0.0, fmin(z, (double)d));
  int i = (int)fmin(x, (double)(w - 2));
  int j = (int)fmin(y, (double)(hh - 2));
  int k = (int)fmin(z, (double)(d - 2));
  const int64_t sj = d, si = (int64_t)hh * d;
  const float* b = u + i * si + j * sj + k;
  const double u000 = b[0], u100 = b[si], u010 = b[sj], u110 = b[si + sj];
  const double u001 = b[1], u101 = b[si + 1], u011 = b[sj + 1], u111 = b[si + sj + 1];
  const double ix1 = (double)(i + 1) - x, ix0 = x - (double)i;
  const double jy1 = (double)(j + 1) - y, jy0 = y - (double)j;
  const double kz1 = (double)(k + 1) - z, kz0 = z - (double)k;
  const double a00 = ix1 * u000 + ix0 * u100, a10 = ix1 * u010 + ix0 * u110;
  const double a01 = ix1 * u001 + ix0 * u101, a11 = ix1 * u011 + ix0 * u111;
  const double b0 = jy1 * a00 + jy0 * a10, b1 = jy1 * a01 + jy0 * a11;
  gx = cx ? 0.0 : kz1 * (jy1 * (u100 - u000) + jy0 * (u110 - u010)) + kz0 * (jy1 * (u101 - u001) + jy0 * (u111 - u011));
  gy = cy ? 0.0 : kz1 * (a10 - a00) + kz0 * (a11 - a01);
  gz = cz ? 0.0 : b1 - b0;
  return kz1 * b0 + kz0 * b1;
}

// The 8 corner weights of trilinear() at (x, y, z) (grid units) and the base
// index, for scattering a cotangent back to the grid.
__device__ __forceinline__ void trilinear_weights(double x, double y, double z, int w, int hh, int d, int& i, int& j,
                                                  int& k, double (&wt)[8]) {
  x = fmax(0.0, fmin(x, (double)w));
  y = fmax(0.0, fmin(y, (double)hh));
  z = fmax(0.0, fmin(z, (double)d));
  i = (int)fmin(x, (double)(w - 2));
  j = (int)fmin(y, (double)(hh - 2));
  k = (int)fmin(z, (double)(d - 2));
  const double ix1 = (double)(i + 1) - x, ix0 = x - (double)i;
  const double jy1 = (double)(j + 1) - y, jy0 = y - (double)j;
  const double kz1 = (double)(k + 1) - z, kz0 = z - (double)k;
  // corner order: (di,dj,dk) = 000,100,010,110,001,101,011,111
  wt[0] = kz1 * jy1 * ix1; wt[1] = kz1 * jy1 * ix0; wt[2] = kz1 * jy0 * ix1; wt[3] = kz1 * jy0 * ix0;
  wt[4] = kz0 * jy1 * ix1; wt[5] = kz0 * jy1 * ix0; wt[6] = kz0 * jy0 * ix1; wt[7] = kz0 * jy0 * ix0;
}

}  // namespace mpf
