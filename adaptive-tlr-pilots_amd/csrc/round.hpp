// Storage precision of a low-rank factor pair (the "mixed" rule of codes/17_tlr_cholesky_pilot.py:37-82).
//
// A pair (U, V) with T = U V^T is stored in the lowest of FP8-E4M3, FP16, FP32 whose rounding error
// ||U V^T - fl(U) fl(V)^T||_F fits the budget; otherwise it stays FP64.  fl(U) scales U so that max|U| maps to
// half the format's largest value (224 for FP8, 32752 for FP16; FP32 unscaled), rounds, and scales back
// (17:62-69).  Arithmetic stays FP64: the rounded values replace U and V in place.
//
// The conversions are exact emulations in FP64 (every step is a power-of-two scaling or one correctly
// rounded operation): FP32 and FP16 are IEEE round-to-nearest-even with their subnormal ranges; FP8 is 17's
// definition (exponent floor(log2|x|) clipped to [-9, 8], three fraction bits, half-even, saturation at 448).
//
// The error is computed without forming T: with dU = U - fl(U), dV = V - fl(V) (exact differences),
// U V^T - fl(U) fl(V)^T = dU V^T + fl(U) dV^T, so for row i of U
//   ||D_i||^2 = dU_i G1 dU_i^T + 2 dU_i G2 fl(U)_i^T + fl(U)_i G3 fl(U)_i^T,
//   G1 = V^T V, G2 = V^T dV, G3 = dV^T dV (r x r each).
#pragma once
#include "common.hpp"

namespace tlr {

__device__ __forceinline__ double round_ieee(double x, int frac_bits, int emin) {
  const double a = fabs(x);
  if (a == 0.0 || !(a <= 1.7976931348623157e308)) return x;
  int e = ilogb(a);
  if (e < emin) e = emin;                       // subnormal range: fixed quantum 2^(emin - frac_bits)
  const double q = ldexp(1.0, e - frac_bits);
  return rint(x / q) * q;
}

__device__ __forceinline__ double round_e4m3_17(double x) {
  const double a = fabs(x);
  if (a == 0.0) return 0.0;
  int e = ilogb(a);
  e = e < -9 ? -9 : (e > 8 ? 8 : e);
  double v = rint(ldexp(a, -e) * 8.0) * 0.125;
  v = ldexp(v, e);
  if (v > 448.0) v = 448.0;
  return x < 0.0 ? -v : v;
}

// Value of x in format p (8, 16, 32) after multiplying by scale and before dividing by it.
__device__ __forceinline__ double round_fmt(double x, int p) {
  if (p == 32) return round_ieee(x, 23, -126);
  if (p == 16) return round_ieee(x, 10, -14);
  return round_e4m3_17(x);
}

// fl(x) of a factor whose largest magnitude is mx (scale precomputed as in 17's stored()).
__device__ __forceinline__ double stored_value(double x, int p, double scale, bool raw) {
  if (raw) return x;
  return p < 32 ? round_fmt(x * scale, p) / scale : round_fmt(x, p);
}

__device__ __forceinline__ double fmt_scale(int p, double mx) {
  const double cap = p == 8 ? 224.0 : 32752.0;
  return p < 32 ? cap / mx : 1.0;
}

// Sum over the work group in a fixed order (wave sums, then waves in order); every thread gets the result.
// red: LDS, at least (blockDim.x / 64 + 1) doubles.  Must be called by every thread.
__device__ __forceinline__ double wg_sum(double v, double* red) {
  v = wave_sum(v);
  const int nw = blockDim.x / 64;
  __syncthreads();
  if (lane_id() == 0) red[wave_id()] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += red[w];
    red[nw] = s;
  }
  __syncthreads();
  return red[nw];
}

__device__ __forceinline__ double wg_max(double v, double* red) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int nw = blockDim.x / 64;
  __syncthreads();
  if (lane_id() == 0) red[wave_id()] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s = fmax(s, red[w]);
    red[nw] = s;
  }
  __syncthreads();
  return red[nw];
}

// Choose and apply the storage precision of (U, V): U is mu x r (leading dimension ldu), V is mv x r.
// Returns 8, 16, 32 or 64.  gram: at least 3 r^2 doubles (global or LDS) private to the work group; red:
// LDS reduction buffer (wg_sum).  Every thread of the work group must call it.
__device__ int round_pair_wg(double* U, int64_t ldu, int mu, double* V, int64_t ldv, int mv, int r, double budget,
                             double* gram, double* red) {
  if (r <= 0) return 64;
  const int tid = threadIdx.x, nt = blockDim.x;
  double mu_l = 0.0, mv_l = 0.0;
  for (int e = tid; e < mu * r; e += nt) mu_l = fmax(mu_l, fabs(U[(e % mu) + (int64_t)(e / mu) * ldu]));
  for (int e = tid; e < mv * r; e += nt) mv_l = fmax(mv_l, fabs(V[(e % mv) + (int64_t)(e / mv) * ldv]));
  const double mxU = wg_max(mu_l, red), mxV = wg_max(mv_l, red);
  double* G1 = gram;
  double* G2 = gram + r * r;
  double* G3 = gram + 2 * r * r;
  // G1 = V^T V does not depend on the format
  for (int e = tid; e < r * r; e += nt) {
    const int a = e % r, c = e / r;
    const double* va = V + (int64_t)a * ldv;
    const double* vc = V + (int64_t)c * ldv;
    double s = 0.0;
    for (int p = 0; p < mv; ++p) s = fma(va[p], vc[p], s);
    G1[e] = s;
  }
  for (int p = 8; p <= 32; p *= 2) {
    const bool rawU = !(mxU > 0.0), rawV = !(mxV > 0.0);
    const double sU = rawU ? 1.0 : fmt_scale(p, mxU), sV = rawV ? 1.0 : fmt_scale(p, mxV);
    __syncthreads();
    for (int e = tid; e < r * r; e += nt) {
      const int a = e % r, c = e / r;
      const double* va = V + (int64_t)a * ldv;
      const double* vc = V + (int64_t)c * ldv;
      double s2 = 0.0, s3 = 0.0;
      for (int q = 0; q < mv; ++q) {
        const double dva = va[q] - stored_value(va[q], p, sV, rawV);
        const double dvc = vc[q] - stored_value(vc[q], p, sV, rawV);
        s2 = fma(va[q], dvc, s2);
        s3 = fma(dva, dvc, s3);
      }
      G2[e] = s2;
      G3[e] = s3;
    }
    __syncthreads();
    double part = 0.0;
    for (int i = tid; i < mu; i += nt) {
      double s = 0.0;
      for (int a = 0; a < r; ++a) {
        const double ua = U[i + (int64_t)a * ldu];
        const double ra = stored_value(ua, p, sU, rawU), da = ua - ra;
        for (int c = 0; c < r; ++c) {
          const double uc = U[i + (int64_t)c * ldu];
          const double rc = stored_value(uc, p, sU, rawU), dc = uc - rc;
          s += da * G1[a + c * r] * dc + 2.0 * da * G2[a + c * r] * rc + ra * G3[a + c * r] * rc;
        }
      }
      part += s;
    }
    const double err2 = wg_sum(part, red);
    if (sqrt(fmax(err2, 0.0)) <= budget) {
      for (int e = tid; e < mu * r; e += nt) {
        double& x = U[(e % mu) + (int64_t)(e / mu) * ldu];
        x = stored_value(x, p, sU, rawU);
      }
      for (int e = tid; e < mv * r; e += nt) {
        double& x = V[(e % mv) + (int64_t)(e / mv) * ldv];
        x = stored_value(x, p, sV, rawV);
      }
      __syncthreads();
      return p;
    }
  }
  return 64;
}

}  // namespace tlr
