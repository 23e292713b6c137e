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

// The two elementary functions of an entry, written for the arguments it has (the general library routines
// spend most of their instructions on range scaling and special classes that never occur here).  Together
// they take about 36 VALU instructions per entry against about 62, and on gfx950 the FP64 VALU work of the
// entries and the FP64 matrix-core work of the products that consume them do not overlap.

// sqrt(d2) for finite d2 >= 0: the hardware reciprocal square root (relative error about 2^-23), one
// Goldschmidt step (g = d2 y, h = y / 2 refined together; error squared) and one Newton correction of the
// root (Markstein's formulation): within 0.5 ulp on the host with y perturbed by up to 2^-22.  d2 is raised
// to at least 1e-300 first (0 and NaN included, as fmax does): a distance below 1e-100 gives an entry equal
// to its value at distance 0 to the last bit, since e^-x and 1 + x both round to 1 there.
__device__ __forceinline__ double sqrt_dist(double d2) {
  d2 = fmax(d2, 1e-300);
  const double y = __builtin_amdgcn_rsq(d2);
  double g = d2 * y, h = 0.5 * y;
  const double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  const double e = fma(-g, g, d2);
  return fma(e, h, g);
}

// a * b + c with the constant c read from a scalar register pair by the three-operand form: left to the
// compiler, a Horner step with a constant addend becomes a copy of the constant into a vector register and
// the two-address v_fmac (two instructions instead of one).  c must be the same in every lane (a scalar
// register holds one value per wave): call it with literal constants only.
__device__ __forceinline__ double fma_sc(double a, double b, double c) {
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
  return d;
}

// e^(-x) for finite x >= 0: Cody-Waite reduction -x = n ln 2 + r, n = rint(-x / ln 2), |r| <= ln 2 / 2, with
// ln 2 split into a head of 31 significant bits (n times it exact for |n| < 2^22) and its tail; e^r by the
// Taylor polynomial of degree 13 in Horner form (truncation below 6e-18 relative); scaled by 2^n, which
// underflows to 0 past x of about 745.  Within 0.89 ulp of e^(-x) over 2^24 arguments in [0, 800].
__device__ __forceinline__ double exp_neg(double x) {
  const double n = rint(x * -1.4426950408889634);
  double r = fma(n, -0x1.62e42ffp-1, -x);
  r = fma(n, 4.2009150726810846e-11, r);   // minus the tail of ln 2 (ln 2 - head = -4.2e-11)
  double p = 1.6059043836821613e-10;        // 1 / 13!
  p = fma_sc(p, r, 2.08767569878681e-09);
  p = fma_sc(p, r, 2.505210838544172e-08);
  p = fma_sc(p, r, 2.755731922398589e-07);
  p = fma_sc(p, r, 2.7557319223985893e-06);
  p = fma_sc(p, r, 2.48015873015873e-05);
  p = fma_sc(p, r, 0.0001984126984126984);
  p = fma_sc(p, r, 0.001388888888888889);
  p = fma_sc(p, r, 0.008333333333333333);
  p = fma_sc(p, r, 0.041666666666666664);
  p = fma_sc(p, r, 0.16666666666666666);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)n);
}

__device__ __forceinline__ double matern32(const MaternParams& mp, double x0, double y0, double x1, double y1) {
  double dx = x0 - x1, dy = y0 - y1;
  double d2 = dx * dx + dy * dy;
  double h = sqrt_dist(d2);
  double x = h * mp.inv_range_sqrt3;
  return mp.sigma2 * ((1.0 + x) * exp_neg(x));
}

// A double moved between lanes by a DPP pattern within rows of 16 lanes (two 32-bit moves).
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const long long x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(x & 0xffffffffll), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(x >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over aligned groups of TP (2, 4, 8 or 16) lanes; every lane of a group gets the group's sum.  The
// exchanges are DPP moves (quad permutations, then mirrors within 8 and 16 lanes), with no LDS round trip.
template <int TP>
__device__ __forceinline__ double group_sum_dpp(double v) {
  static_assert(TP == 2 || TP == 4 || TP == 8 || TP == 16, "group of 2, 4, 8 or 16 lanes");
  v += dpp_move<0xB1>(v);                    // quad_perm(1, 0, 3, 2)
  if (TP >= 4) v += dpp_move<0x4E>(v);       // quad_perm(2, 3, 0, 1)
  if (TP >= 8) v += dpp_move<0x141>(v);      // row_half_mirror: lane l <-> 7 - l within 8
  if (TP >= 16) v += dpp_move<0x140>(v);     // row_mirror: lane l <-> 15 - l within 16
  return v;
}

// The same for a float (one 32-bit move per exchange).
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

template <int TP>
__device__ __forceinline__ float group_sum_dpp(float v) {
  static_assert(TP == 2 || TP == 4 || TP == 8 || TP == 16, "group of 2, 4, 8 or 16 lanes");
  v += dpp_move<0xB1>(v);
  if (TP >= 4) v += dpp_move<0x4E>(v);
  if (TP >= 8) v += dpp_move<0x141>(v);
  if (TP >= 16) v += dpp_move<0x140>(v);
  return v;
}

__device__ __forceinline__ float readlane_f(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
  const long long x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(x & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(x >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes of a wave (all lanes active); every lane gets the sum.  Row sums of 16 lanes by
// DPP exchanges, then the four row sums read from lanes 0, 16, 32, 48 and added in a fixed order.
__device__ __forceinline__ double wave_sum(double v) {
  v = group_sum_dpp<16>(v);
  return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}

__device__ __forceinline__ float wave_sum(float v) {
  v = group_sum_dpp<16>(v);
  return (readlane_f(v, 0) + readlane_f(v, 16)) + (readlane_f(v, 32) + readlane_f(v, 48));
}

// Arithmetic precision of the factorization's recompressions: unit roundoff of the working type (the
// Jacobi tolerance and noise floor scale with it), and the rounding of product operands to bfloat16
// (round to nearest even on the upper 16 bits of the float; NaN and infinity kept) for the BF16 mode, whose
// products then accumulate in float.
template <class T> struct Prec;
template <> struct Prec<double> {
  static constexpr double eps = 2.220446049250313e-16;    // 2^-52
  static constexpr double tiny2 = 4.930380657631324e-32;   // eps^2
};
template <> struct Prec<float> {
  static constexpr float eps = 1.1920928955078125e-07f;    // 2^-23
  static constexpr float tiny2 = 1.4210854715202004e-14f;  // eps^2
};

__device__ __forceinline__ float round_bf16(float x) {
  unsigned u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return x;
  u += 0x7fffu + ((u >> 16) & 1u);
  return __uint_as_float(u & 0xffff0000u);
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
