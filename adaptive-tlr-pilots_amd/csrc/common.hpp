// Shared device helpers for the TLR likelihood kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define TLR_WAVE 64

namespace tlr {

// Matern nu=3/2 covariance sigma2 (1 + x) exp(-x), x = sqrt(3) h / r, h = |s - t| with
// h = sqrt(max(dx^2 + dy^2, 0)) exactly as the reference's dist() (codes/11_tlr_rank_rules.py:56-58).
struct MaternParams {
  double sigma2;
  double inv_range_sqrt3;  // sqrt(3) / range
  double nugget;           // sigma2 * eta, added on the diagonal of diagonal tiles only
};

__device__ __forceinline__ double matern32(const MaternParams& mp, double x0, double y0, double x1, double y1) {
  double dx = x0 - x1, dy = y0 - y1;
  double d2 = dx * dx + dy * dy;
  double h = sqrt(fmax(d2, 0.0));
  double x = h * mp.inv_range_sqrt3;
  return mp.sigma2 * ((1.0 + x) * exp(-x));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, TLR_WAVE);
  return v;
}

// Sum over groups of G consecutive lanes (G a power of two <= 64).
template <int G>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = G / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, TLR_WAVE);
  return v;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (TLR_WAVE - 1); }
__device__ __forceinline__ int wave_id() { return threadIdx.x / TLR_WAVE; }

// Index of the off-diagonal tile (i, j), i > j, in the row-major lower block triangle.
__host__ __device__ __forceinline__ int64_t tri_index(int i, int j) { return (int64_t)i * (i - 1) / 2 + j; }

// Inverse of tri_index: tile t -> (i, j).  i is found by a float guess and integer correction.
__host__ __device__ __forceinline__ void tri_decode(int64_t t, int* pi, int* pj) {
  int i = (int)((1.0 + sqrt(1.0 + 8.0 * (double)t)) * 0.5);
  while ((int64_t)i * (i - 1) / 2 > t) --i;
  while ((int64_t)(i + 1) * i / 2 <= t) ++i;
  *pi = i;
  *pj = (int)(t - (int64_t)i * (i - 1) / 2);
}

}  // namespace tlr
