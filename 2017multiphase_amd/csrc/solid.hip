// Rigid motion of the solid particles: cspray::solidmotion, spray.cu:2691-2774
// (solidCollisionWithBound, accumulate_GPU_k, compute_cI_k, computeVelSolid_k,
// computePosSolid_k, spray_k.cu:5167-5252, 5478-5508).
//
// The body's translation T (mean velocity), centre rg0 (mean position) and
// rotation R (the reference's estimate: sum of r x v over the particles,
// divided by the square of the mean |r| and by the count) come from
// reductions over the SOLID particles.  The reference adds float block sums
// on the host; here each of a fixed grid of blocks sums its grid-stride share
// in double, the block sums are added in block order by one block, and the
// results are rounded to float where the reference's are float.
#include "mpf_internal.h"

namespace mpf {

namespace {

constexpr int SM_BLOCKS = 1024, SM_THREADS = 256, SM_NV = 8;

__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float len3(f3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Block sum of SM_NV doubles per thread, in a fixed tree order.
__device__ __forceinline__ void block_sum(double (&v)[SM_NV], double* __restrict__ out) {
  __shared__ double sh[SM_NV][SM_THREADS];
#pragma unroll
  for (int q = 0; q < SM_NV; ++q) sh[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int s = SM_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int q = 0; q < SM_NV; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < SM_NV; ++q) out[q] = sh[q][0];
}

// pass A: count and sum of positions of the SOLID particles
__global__ void __launch_bounds__(SM_THREADS) k_solid_sums_a(int64_t P, const f3* __restrict__ pos,
                                                               const int8_t* __restrict__ flag,
                                                               double* __restrict__ part) {
  double v[SM_NV] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t p = (int64_t)blockIdx.x * SM_THREADS + threadIdx.x; p < P; p += (int64_t)SM_BLOCKS * SM_THREADS) {
    if (flag[p] != SOLID) continue;
    const f3 x = pos[p];
    v[0] += 1.0;
    v[1] += x.x;
    v[2] += x.y;
    v[3] += x.z;
  }
  block_sum(v, part + (int64_t)blockIdx.x * SM_NV);
}

// solidCollisionWithBound, spray_k.cu:5478-5508: solid particles past the
// wall margin are pushed back through their velocity (scaled by the count)
__global__ void k_solid_bound(Grid g, int64_t P, const f3* __restrict__ pos, f3* __restrict__ vel,
                              const int8_t* __restrict__ flag, float bounce, int nsol, float ss) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P || flag[p] != SOLID) return;
  const float m = g.h + 0.3f * ss;
  const f3 tmin = mk3(g.gx0 + m, g.gy0 + m, g.gz0 + m);
  const f3 tmax = mk3(g.gx1 - m, g.gy1 - m, g.gz1 - m);
  const f3 x = pos[p];
  f3 v = vel[p];
  const float fn = (float)nsol;
  if (x.x < tmin.x) v.x += (tmin.x - x.x) * bounce * fn;
  if (x.x > tmax.x) v.x -= (x.x - tmax.x) * bounce * fn;
  if (x.y < tmin.y) v.y += (tmin.y - x.y) * bounce * fn;
  if (x.y > tmax.y) v.y -= (x.y - tmax.y) * bounce * fn;
  if (x.z < tmin.z) v.z += (tmin.z - x.z) * bounce * fn;
  if (x.z > tmax.z) v.z -= (x.z - tmax.z) * bounce * fn;
  vel[p] = v;
}

// pass B: sums of velocity, of r x v and of |r| (compute_cI_k, spray_k.cu:
// 5188-5210: particles within 1e-6 of the centre add nothing to the last two)
__global__ void __launch_bounds__(SM_THREADS) k_solid_sums_b(int64_t P, const f3* __restrict__ pos,
                                                               const f3* __restrict__ vel,
                                                               const int8_t* __restrict__ flag, f3 rg0,
                                                               double* __restrict__ part) {
  double v[SM_NV] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t p = (int64_t)blockIdx.x * SM_THREADS + threadIdx.x; p < P; p += (int64_t)SM_BLOCKS * SM_THREADS) {
    if (flag[p] != SOLID) continue;
    const f3 x = pos[p], u = vel[p];
    v[0] += u.x;
    v[1] += u.y;
    v[2] += u.z;
    const f3 r = sub3(x, rg0);
    const float dis = len3(r);
    if (dis > 1e-6f) {
      const f3 c = cross3(r, u);
      v[3] += c.x;
      v[4] += c.y;
      v[5] += c.z;
      v[6] += dis;
    }
  }
  block_sum(v, part + (int64_t)blockIdx.x * SM_NV);
}

// block sums added in block order (one block; the SM_NV totals land after
// the block sums)
__global__ void __launch_bounds__(SM_THREADS) k_solid_total(double* __restrict__ part) {
  double v[SM_NV] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < SM_BLOCKS; b += SM_THREADS)
#pragma unroll
    for (int q = 0; q < SM_NV; ++q) v[q] += part[(int64_t)b * SM_NV + q];
  block_sum(v, part + (int64_t)SM_BLOCKS * SM_NV);
}

struct Rot {
  float m[9];  // x00 x01 x02 x10 x11 x12 x20 x21 x22
};

// computeVelSolid_k + computePosSolid_k, spray_k.cu:5220-5252
__global__ void k_solid_move(int64_t P, f3* __restrict__ pos, f3* __restrict__ vel, const int8_t* __restrict__ flag,
                             f3 rg0, f3 rg, f3 R, f3 T, Rot rm) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P || flag[p] != SOLID) return;
  const f3 x = pos[p], u = vel[p];
  const f3 vh = add3(cross3(R, sub3(x, rg)), T);
  vel[p] = mk3(0.5f * (u.x + vh.x), 0.5f * (u.y + vh.y), 0.5f * (u.z + vh.z));
  const f3 t = sub3(x, rg0);
  const float* m = rm.m;
  const f3 r = mk3(t.x * m[0] + t.y * m[3] + t.z * m[6], t.x * m[1] + t.y * m[4] + t.z * m[7],
                   t.x * m[2] + t.y * m[5] + t.z * m[8]);
  pos[p] = add3(r, rg);
}

__global__ void k_solid_zero_vel(int64_t P, f3* __restrict__ vel, const int8_t* __restrict__ flag) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < P && flag[p] == SOLID) vel[p] = mk3(0.0f, 0.0f, 0.0f);
}

}  // namespace

// The host arithmetic of cspray::solidmotion (spray.cu:2714-2766) on the
// reduced totals, in float as the reference does it.
void solid_motion_params(const double* A, const double* B, float dt, int lock_y, float* rg0, float* rg, float* R,
                         float* T, float* rm, int* nsol) {
  const int n = (int)A[0];
  *nsol = n;
  if (n <= 0) return;
  const float fn = (float)n;
  float t[3] = {(float)B[0] / fn, (float)B[1] / fn, (float)B[2] / fn};
  if (lock_y) t[1] = 0.0f;
  for (int a = 0; a < 3; ++a) {
    T[a] = t[a];
    rg0[a] = (float)A[1 + a] / fn;
    rg[a] = rg0[a] + dt * t[a];
  }
  float ti = (float)B[6];
  if (ti <= 0) ti = 1.0f;
  ti /= fn;
  for (int a = 0; a < 3; ++a) R[a] = (float)B[3 + a] / ti / ti / fn;
  // rotation of the positions about R by |R| dt (Rodrigues; the reference
  // builds it for -|R| dt and applies its transpose)
  float ax[3] = {R[0], R[1], R[2]};
  float theta = -sqrtf(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]) * dt;
  if (std::fabs(theta) > 1e-6) {
    const float inv = 1.0f / sqrtf(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
    for (int a = 0; a < 3; ++a) ax[a] = ax[a] * inv;
  } else {
    ax[0] = 1.0f, ax[1] = 0.0f, ax[2] = 0.0f;
    theta = 0.0f;
  }
  const float x = ax[0], y = ax[1], z = ax[2];
  const float c = cosf(theta), s = sinf(theta);
  rm[0] = c + (1 - c) * x * x, rm[1] = (1 - c) * x * y - s * z, rm[2] = (1 - c) * x * z + s * y;
  rm[3] = (1 - c) * y * x + s * z, rm[4] = c + (1 - c) * y * y, rm[5] = (1 - c) * y * z - s * x;
  rm[6] = (1 - c) * z * x - s * y, rm[7] = (1 - c) * z * y + s * x, rm[8] = c + (1 - c) * z * z;
}

int run_solid_motion(Ctx& c) {
  const int cur = c.cur;
  const int64_t P = c.P;
  if (c.cfg.fix_solids) {  // solidmotion_fixed, spray.cu:2685-2689
    launch(c, "solid_zero_vel", dim3(nblocks(P, 256)), dim3(256), 0, k_solid_zero_vel, P, c.vel[cur],
           (const int8_t*)c.flag[cur]);
    return MPF_OK;
  }
  if (P == 0) return MPF_OK;
  double* part = c.partials;  // the solver's partial-sum space, free outside a solve
  static_assert((SM_BLOCKS + 1) * SM_NV <= 1024 * 48, "solid-motion sums exceed the partial-sum space");
  launch(c, "solid_sums_a", dim3(SM_BLOCKS), dim3(SM_THREADS), 0, k_solid_sums_a, P, (const f3*)c.pos[cur],
         (const int8_t*)c.flag[cur], part);
  launch(c, "solid_total", dim3(1), dim3(SM_THREADS), 0, k_solid_total, part);
  double A[SM_NV], B[SM_NV];
  if (!hip_ok(c, hipMemcpyAsync(c.red_host, part + (int64_t)SM_BLOCKS * SM_NV, SM_NV * 8, hipMemcpyDeviceToHost,
                                c.stream), "solid_motion") ||
      !hip_ok(c, hipStreamSynchronize(c.stream), "solid_motion"))
    return MPF_HIP_ERROR;
  for (int q = 0; q < SM_NV; ++q) A[q] = c.red_host[q];
  const int nsol = (int)A[0];
  if (nsol <= 0) return MPF_OK;  // the reference returns when there is no solid
  launch(c, "solid_bound", dim3(nblocks(P, 256)), dim3(256), 0, k_solid_bound, c.g, P, (const f3*)c.pos[cur],
         c.vel[cur], (const int8_t*)c.flag[cur], c.cfg.solid_bounce, nsol, c.cfg.samplespace);
  const float fn = (float)nsol;
  const f3 rg0 = f3{(float)A[1] / fn, (float)A[2] / fn, (float)A[3] / fn};
  launch(c, "solid_sums_b", dim3(SM_BLOCKS), dim3(SM_THREADS), 0, k_solid_sums_b, P, (const f3*)c.pos[cur],
         (const f3*)c.vel[cur], (const int8_t*)c.flag[cur], rg0, part);
  launch(c, "solid_total", dim3(1), dim3(SM_THREADS), 0, k_solid_total, part);
  if (!hip_ok(c, hipMemcpyAsync(c.red_host, part + (int64_t)SM_BLOCKS * SM_NV, SM_NV * 8, hipMemcpyDeviceToHost,
                                c.stream), "solid_motion") ||
      !hip_ok(c, hipStreamSynchronize(c.stream), "solid_motion"))
    return MPF_HIP_ERROR;
  for (int q = 0; q < SM_NV; ++q) B[q] = c.red_host[q];
  float r0[3], r1[3], R[3], T[3], rm[9];
  int n = 0;
  solid_motion_params(A, B, c.cfg.dt, c.cfg.rigid_lock_y, r0, r1, R, T, rm, &n);
  Rot rot;
  for (int q = 0; q < 9; ++q) rot.m[q] = rm[q];
  launch(c, "solid_move", dim3(nblocks(P, 256)), dim3(256), 0, k_solid_move, P, c.pos[cur], c.vel[cur],
         (const int8_t*)c.flag[cur], f3{r0[0], r0[1], r0[2]}, f3{r1[0], r1[1], r1[2]}, f3{R[0], R[1], R[2]},
         f3{T[0], T[1], T[2]}, rot);
  c.sorted = c.own_sort = c.mc_sorted = false;
  return MPF_OK;
}

}  // namespace mpf
