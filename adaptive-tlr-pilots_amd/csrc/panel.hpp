// Work-group-wide small dense linear algebra kept in registers / LDS:
//  * RegPanel: an m x n panel (m <= 64*RPL, n <= CPW*NW) distributed over NW waves; wave w owns the
//    columns j = w + NW*jj (jj < CPW), lane l owns the rows p = l + 64*t (t < RPL).  Householder QR with
//    LAPACK dlarfg/dgeqr2 conventions, explicit Q in place (dorg2r), all without spilling the panel.
//  * panel_coef / panel_axpy: projections of the panel against a global matrix (block Gram-Schmidt).
//  * jacobi_svd: one-sided (Hestenes) Jacobi on the columns of a small square matrix, round-robin
//    ordering, rotations accumulated into J.
//  Every routine is templated on the working type T (double, or float for the reduced-precision
//  recompressions); global operands stay double and are converted on load and store.  BF = true rounds the
//  operands of a product to bfloat16 (products accumulate in T = float).
#pragma once
#include "common.hpp"

namespace tlr {

__device__ __forceinline__ double tsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float tsqrt(float x) { return __fsqrt_rn(x); }

// x converted to the working type, rounded to bfloat16 when BF (only for T = float)
template <class T, bool BF = false>
__device__ __forceinline__ T cvt(double x) {
  static_assert(!BF || sizeof(T) == 4, "bfloat16 operands need float accumulation");
  if constexpr (BF) return round_bf16((float)x);
  else return (T)x;
}
template <class T, bool BF = false>
__device__ __forceinline__ T cvt(float x) {
  static_assert(!BF || sizeof(T) == 4, "bfloat16 operands need float accumulation");
  if constexpr (BF) return round_bf16(x);
  else return (T)x;
}

// Sums over the 64 lanes of a wave of K = 4 or 8 values per lane; every lane gets all K totals.  A
// reduce-scatter: each halving exchange (lanes l and l^1, l and l^2, and for K = 8 lanes l and l-4 within the
// row of 16) leaves a lane half of the values it held, summed with its partner's copies, so that after two (K =
// 4) or three (K = 8) of them each lane holds one value, the one selected by its lane bits 0-2; exchanges with
// lanes 4 (K = 4) and 8 apart complete the row, two cross-row exchanges the wave, and one read per value
// broadcasts the totals.  Every exchange within a row is a DPP move.  Compared with K separate wave_sum calls
// this moves K + 4 values per lane instead of 4 K and reads 2 K lanes instead of 8 K; the additions run in
// another order (the same sums, rounded differently).  Every lane of the wave must be active.
template <int K>
__device__ __forceinline__ void wave_sum_multi(double (&d)[K]) {
  static_assert(K == 4 || K == 8, "4 or 8 values");
  constexpr int H = K / 2, Q = K / 4;
  const int l = lane_id();
  const bool b0 = l & 1, b1 = (l >> 1) & 1, b2 = (l >> 2) & 1;
  double e[H];
#pragma unroll
  for (int i = 0; i < H; ++i) {   // lanes with bit 0 set keep values H.., the others 0..H-1
    const double send = b0 ? d[i] : d[i + H];
    const double keep = b0 ? d[i + H] : d[i];
    e[i] = keep + dpp_move<0xB1>(send);   // quad_perm(1, 0, 3, 2): lane l^1
  }
  double f[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {   // bit 1 selects the upper half of what is left
    const double send = b1 ? e[i] : e[i + Q];
    const double keep = b1 ? e[i + Q] : e[i];
    f[i] = keep + dpp_move<0x4E>(send);   // quad_perm(2, 3, 0, 1): lane l^2
  }
  double g;
  if constexpr (K == 8) {   // bit 2 selects one of the two; lane l-4 (mod 16) has the other bit 2
    const double send = b2 ? f[0] : f[1];
    const double keep = b2 ? f[1] : f[0];
    g = keep + dpp_move<0x124>(send);     // row_ror:4
  } else {
    g = f[0];
    g += dpp_move<0x124>(g);              // lanes 4 apart hold the same value index
  }
  g += dpp_move<0x128>(g);                // row_ror:8: the row's 16 lanes
  g += __shfl_xor(g, 16);
  g += __shfl_xor(g, 32);
  // value c ended in the lanes whose bits (0, 1[, 2]) spell c from its most significant bit down
#pragma unroll
  for (int c = 0; c < K; ++c) {
    const int lane = K == 8 ? (((c >> 2) & 1) | (((c >> 1) & 1) << 1) | ((c & 1) << 2))
                            : (((c >> 1) & 1) | ((c & 1) << 1));
    d[c] = readlane_d(g, lane);
  }
}

template <int RPL, int NW, int CPW_ = 8, class T = double>
struct RegPanel {
  static constexpr int CPW = CPW_;
  static constexpr int NT = 64 * NW;
  T a[CPW][RPL];

  __device__ __forceinline__ static int col_of(int w, int jj) { return w + NW * jj; }

  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj)
#pragma unroll
      for (int t = 0; t < RPL; ++t) a[jj][t] = (T)0;
  }

  // A: column-major, leading dimension lda; entries outside m x n read as zero.
  template <bool BF = false, class S>
  __device__ __forceinline__ void load(const S* A, int64_t lda, int m, int n) {
    const int w = wave_id(), l = lane_id();
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) {
      const int j = col_of(w, jj);
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int p = l + 64 * t;
        a[jj][t] = (j < n && p < m) ? cvt<T, BF>(A[p + (int64_t)j * lda]) : (T)0;
      }
    }
  }

  template <class S>
  __device__ __forceinline__ void store(S* A, int64_t lda, int m, int n) const {
    const int w = wave_id(), l = lane_id();
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) {
      const int j = col_of(w, jj);
      if (j < n) {
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
          const int p = l + 64 * t;
          if (p < m) A[p + (int64_t)j * lda] = (S)a[jj][t];
        }
      }
    }
  }

  // Copy column slot jc (wave-uniform) into x.
  // Apply the reflector in vb (rows >= c nonzero) with factor tc to every own column j > c.
  __device__ __forceinline__ void apply_reflector(int c, int n, const T* vb, T tc) {
    const int w = wave_id(), l = lane_id();
    T v[RPL];
#pragma unroll
    for (int t = 0; t < RPL; ++t) v[t] = vb[l + 64 * t];
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) {
      const int j = col_of(w, jj);
      if (j > c && j < n) {
        T d = (T)0;
#pragma unroll
        for (int t = 0; t < RPL; ++t) d += v[t] * a[jj][t];
        d = wave_sum(d) * tc;
#pragma unroll
        for (int t = 0; t < RPL; ++t) a[jj][t] -= d * v[t];
      }
    }
  }

  // Householder QR (dgeqr2): reflector c stored below the diagonal of column c (v_c = 1 implicit),
  // R on and above the diagonal, tau[c] in LDS.  vbuf: LDS scratch of 2*64*RPL doubles.
  // Requires m >= n.  Every thread of the work group must call it.  Column c = wo + NW*jc is owned
  // by wave wo in register slot jc; the slot loop is unrolled so every register index is static.
  __device__ __forceinline__ void geqr2(int m, int n, T* vbuf, T* tau) {
    const int w = wave_id(), l = lane_id();
#pragma unroll
    for (int jc = 0; jc < CPW; ++jc) {
      for (int wo = 0; wo < NW; ++wo) {
        const int c = wo + NW * jc;
        if (c >= n) break;
        T* vb = vbuf + (c & 1) * (64 * RPL);
        if (w == wo) {
          T part = (T)0, alpha_l = (T)0;
#pragma unroll
          for (int t = 0; t < RPL; ++t) {
            const int p = l + 64 * t;
            const T x = a[jc][t];
            if (p > c && p < m) part += x * x;
            if (p == c) alpha_l = x;
          }
          const T sigma = wave_sum(part);
          const T alpha = wave_sum(alpha_l);  // exactly one lane holds it
          T beta, tc, scal;
          if (sigma == (T)0) {
            tc = (T)0; beta = alpha; scal = (T)0;
          } else {
            const T nrm = tsqrt(alpha * alpha + sigma);
            beta = alpha >= (T)0 ? -nrm : nrm;
            tc = (beta - alpha) / beta;
            scal = (T)1 / (alpha - beta);
          }
#pragma unroll
          for (int t = 0; t < RPL; ++t) {
            const int p = l + 64 * t;
            T v = (T)0;
            if (p == c) v = (T)1;
            else if (p > c && p < m) v = a[jc][t] * scal;
            vb[p] = v;
            if (p == c) a[jc][t] = beta;
            else if (p > c) a[jc][t] = (p < m) ? v : (T)0;
          }
          if (l == 0) tau[c] = tc;
        }
        __syncthreads();
        apply_reflector(c, n, vb, tau[c]);
      }
    }
    __syncthreads();
  }

  // apply_reflector with the wave's CPW dot products summed together (wave_sum_multi; T = double, CPW = 4 or
  // 8).  Every wave must call it (the sums need all 64 lanes); a wave with no column left skips the work.
  __device__ __forceinline__ void apply_reflector_mv(int c, int n, const T* vb, T tc) {
    static_assert(sizeof(T) == 8 && (CPW == 4 || CPW == 8), "double panel, 4 or 8 columns per wave");
    const int w = wave_id(), l = lane_id();
    bool act = false;
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) act |= (col_of(w, jj) > c && col_of(w, jj) < n);
    if (!act) return;
    T v[RPL];
#pragma unroll
    for (int t = 0; t < RPL; ++t) v[t] = vb[l + 64 * t];
    T d[CPW];
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) {
      const int j = col_of(w, jj);
      T s = (T)0;
      if (j > c && j < n) {
#pragma unroll
        for (int t = 0; t < RPL; ++t) s += v[t] * a[jj][t];
      }
      d[jj] = s;
    }
    wave_sum_multi<CPW>(d);
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) {
      const int j = col_of(w, jj);
      if (j > c && j < n) {
        const T dd = d[jj] * tc;
#pragma unroll
        for (int t = 0; t < RPL; ++t) a[jj][t] -= dd * v[t];
      }
    }
  }

  // geqr2 with the reflector applications of apply_reflector_mv, and the diagonal entry read from its lane
  // (c mod 64) instead of summed over the wave; same reflectors and R as geqr2 up to rounding.
  __device__ __forceinline__ void geqr2_mv(int m, int n, T* vbuf, T* tau) {
    const int w = wave_id(), l = lane_id();
#pragma unroll
    for (int jc = 0; jc < CPW; ++jc) {
      for (int wo = 0; wo < NW; ++wo) {
        const int c = wo + NW * jc;
        if (c >= n) break;
        T* vb = vbuf + (c & 1) * (64 * RPL);
        if (w == wo) {
          T part = (T)0, xc = a[jc][0];
#pragma unroll
          for (int t = 0; t < RPL; ++t) {
            const int p = l + 64 * t;
            const T x = a[jc][t];
            if (p > c && p < m) part += x * x;
            if (t == (c >> 6)) xc = x;
          }
          const T sigma = wave_sum(part);
          const T alpha = readlane_d(xc, c & 63);
          T beta, tc, scal;
          if (sigma == (T)0) {
            tc = (T)0; beta = alpha; scal = (T)0;
          } else {
            const T nrm = tsqrt(alpha * alpha + sigma);
            beta = alpha >= (T)0 ? -nrm : nrm;
            tc = (beta - alpha) / beta;
            scal = (T)1 / (alpha - beta);
          }
#pragma unroll
          for (int t = 0; t < RPL; ++t) {
            const int p = l + 64 * t;
            T v = (T)0;
            if (p == c) v = (T)1;
            else if (p > c && p < m) v = a[jc][t] * scal;
            vb[p] = v;
            if (p == c) a[jc][t] = beta;
            else if (p > c) a[jc][t] = (p < m) ? v : (T)0;
          }
          if (l == 0) tau[c] = tc;
        }
        __syncthreads();
        apply_reflector_mv(c, n, vb, tau[c]);
      }
    }
    __syncthreads();
  }

  // org2r with the reflector applications of apply_reflector_mv.
  __device__ __forceinline__ void org2r_mv(int m, int n, T* vbuf, const T* tau) {
    const int w = wave_id(), l = lane_id();
#pragma unroll
    for (int jc = CPW - 1; jc >= 0; --jc) {
      for (int wo = NW - 1; wo >= 0; --wo) {
        const int c = wo + NW * jc;
        if (c >= n) continue;
        T* vb = vbuf + (c & 1) * (64 * RPL);
        const T tc = tau[c];
        if (w == wo) {
#pragma unroll
          for (int t = 0; t < RPL; ++t) {
            const int p = l + 64 * t;
            T v = (T)0;
            if (p == c) v = (T)1;
            else if (p > c && p < m) v = a[jc][t];
            vb[p] = v;
            a[jc][t] = (p == c) ? (T)1 - tc : ((p > c && p < m) ? -tc * v : (T)0);
          }
        }
        __syncthreads();
        apply_reflector_mv(c, n, vb, tc);
      }
    }
    __syncthreads();
  }

  // Write the n x n upper triangle R (zeros below) to R (column-major, ldr).  Call after geqr2.
  __device__ __forceinline__ void store_r(T* R, int ldr, int n) const {
    const int w = wave_id(), l = lane_id();
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) {
      const int j = col_of(w, jj);
      if (j < n) {
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
          const int p = l + 64 * t;
          if (p < n) R[p + j * ldr] = (p <= j) ? a[jj][t] : (T)0;
        }
      }
    }
  }

  // Same, transposed: Rt[j + p*ldr] = R[p][j].
  __device__ __forceinline__ void store_rt(T* Rt, int ldr, int n) const {
    const int w = wave_id(), l = lane_id();
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) {
      const int j = col_of(w, jj);
      if (j < n) {
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
          const int p = l + 64 * t;
          if (p < n) Rt[j + p * ldr] = (p <= j) ? a[jj][t] : (T)0;
        }
      }
    }
  }

  // Explicit Q (m x n) in place from the reflectors left by geqr2 (dorg2r), columns in reverse order.
  __device__ __forceinline__ void org2r(int m, int n, T* vbuf, const T* tau) {
    const int w = wave_id(), l = lane_id();
#pragma unroll
    for (int jc = CPW - 1; jc >= 0; --jc) {
      for (int wo = NW - 1; wo >= 0; --wo) {
        const int c = wo + NW * jc;
        if (c >= n) continue;
        T* vb = vbuf + (c & 1) * (64 * RPL);
        const T tc = tau[c];
        if (w == wo) {
#pragma unroll
          for (int t = 0; t < RPL; ++t) {
            const int p = l + 64 * t;
            T v = (T)0;
            if (p == c) v = (T)1;
            else if (p > c && p < m) v = a[jc][t];
            vb[p] = v;
            // column c of Q = H_c e_c = e_c - tau v
            a[jc][t] = (p == c) ? (T)1 - tc : ((p > c && p < m) ? -tc * v : (T)0);
          }
        }
        __syncthreads();
        apply_reflector(c, n, vb, tc);
      }
    }
    __syncthreads();
  }
};

// C[a + j*ldc] = sc[a] * (A[:, a]^T P[:, j]) for a < na, j < n; A global (m x na, column-major, lda),
// sc optional column scale.  Each wave produces the entries of its own panel columns.  The columns of A
// are fetched AU at a time so that several loads are in flight per wave.
// BF: the entries of A are rounded to bfloat16 (T = float).
template <bool BF = false, int RPL, int NW, int CPW, class T, class S>
__device__ __forceinline__ void panel_coef(const RegPanel<RPL, NW, CPW, T>& P, const S* A, int64_t lda, int m,
                                           int na, int n, const T* sc, T* C, int ldc) {
  constexpr int AU = CPW <= 4 ? 2 : 1;
  const int w = wave_id(), l = lane_id();
  for (int a0 = 0; a0 < na; a0 += AU) {
    T av[AU][RPL];
#pragma unroll
    for (int u = 0; u < AU; ++u)
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int p = l + 64 * t;
        av[u][t] = (a0 + u < na && p < m) ? cvt<T, BF>(A[p + (int64_t)(a0 + u) * lda]) : (T)0;
      }
#pragma unroll
    for (int u = 0; u < AU; ++u) {
      const int a = a0 + u;
      if (a < na) {
        const T s = sc ? sc[a] : (T)1;
#pragma unroll
        for (int jj = 0; jj < CPW; ++jj) {
          const int j = RegPanel<RPL, NW, CPW, T>::col_of(w, jj);
          if (j < n) {
            T d = (T)0;
#pragma unroll
            for (int t = 0; t < RPL; ++t) d += av[u][t] * P.a[jj][t];
            d = wave_sum(d);
            if (l == 0) C[a + j * ldc] = d * s;
          }
        }
      }
    }
  }
}

// P[:, j] += alpha * sum_a A[:, a] sc[a] C[a + j*ldc] for j < n (C must be visible to every wave).
// Columns of A fetched AU at a time; the accumulation order over a is sequential.
// BF: the entries of A and of C are rounded to bfloat16 (T = float; alpha and sc must be exact in it).
template <bool BF = false, int RPL, int NW, int CPW, class T, class S>
__device__ __forceinline__ void panel_axpy(RegPanel<RPL, NW, CPW, T>& P, const S* A, int64_t lda, int m, int na,
                                           int n, const T* sc, const T* C, int ldc, T alpha) {
  constexpr int AU = CPW <= 4 ? 2 : 1;
  const int w = wave_id(), l = lane_id();
  for (int a0 = 0; a0 < na; a0 += AU) {
    T av[AU][RPL];
#pragma unroll
    for (int u = 0; u < AU; ++u)
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int p = l + 64 * t;
        av[u][t] = (a0 + u < na && p < m) ? cvt<T, BF>(A[p + (int64_t)(a0 + u) * lda]) : (T)0;
      }
#pragma unroll
    for (int u = 0; u < AU; ++u) {
      const int a = a0 + u;
      if (a < na) {
        const T s = alpha * (sc ? sc[a] : (T)1);
#pragma unroll
        for (int jj = 0; jj < CPW; ++jj) {
          const int j = RegPanel<RPL, NW, CPW, T>::col_of(w, jj);
          if (j < n) {
            const T f = s * cvt<T, BF>(C[a + j * ldc]);
#pragma unroll
            for (int t = 0; t < RPL; ++t) P.a[jj][t] = fma(f, av[u][t], P.a[jj][t]);
          }
        }
      }
    }
  }
}

// Orders the LDS accesses of one wave: the stores before it are seen by every lane's loads after it (a
// wave's LDS operations execute in issue order; the fences keep the compiler from moving them across).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Rotation of the column pair (p, q) in one-sided Jacobi: given al = ||g_p||^2, be = ||g_q||^2 and
// ga = g_p . g_q, returns false when the pair is already orthogonal to the tolerance (|ga| <= tol
// sqrt(al be)) or either column is below the noise floor, else the cosine and sine of the rotation that
// makes the columns orthogonal: t = tan(angle) = sign(d) h / (|d| + sqrt(d^2 + h^2)) with d = be - al and
// h = 2 ga (t = 1 when d = 0), cs = 1 / sqrt(1 + t^2), sn = cs t.
template <class T>
__device__ __forceinline__ bool jacobi_rotation(T al, T be, T ga, T floor2, T tol, T* cs, T* sn) {
  if (!(ga != (T)0 && al > floor2 && be > floor2 && ga * ga > tol * tol * al * be)) return false;
  const T d = be - al, h = (T)2 * ga;
  const T t = d == (T)0 ? (T)1 : (d > (T)0 ? h : -h) / (fabs(d) + tsqrt(fma(d, d, h * h)));
  const T c = (T)1 / tsqrt(fma(t, t, (T)1));
  *cs = c;
  *sn = c * t;
  return true;
}

// Largest finite value of T (column norms beyond it are reported as non-finite).
__device__ __forceinline__ constexpr double huge_of(double) { return 1.7976931348623157e308; }
__device__ __forceinline__ constexpr float huge_of(float) { return 3.4028234663852886e38f; }

// jacobi_svd for n <= 128 / TP (TP = 2, 4, 8 or 16 lanes per column pair), run by one wave alone: its 64
// lanes hold every column pair of a round (rows tl, tl + TP, ... of both columns and of J in registers), so
// rounds are separated by a wave-level fence instead of a work-group barrier, each column is read once per
// round, and the group sums are DPP exchanges.  Same ordering, rotation, noise floor and stopping rule as
// jacobi_svd.  G, J, sig and perm must be in LDS (the wave-level fence orders LDS accesses only).  Only one
// wave may call it; the caller synchronises the work group after.
template <int TP, class T>
__device__ int jacobi_svd_w0(T* G, int ldg, T* J, int ldj, int n, T* sig, int* perm, int max_sweeps) {
  constexpr int UR = 128 / (TP * TP) > 0 ? 128 / (TP * TP) : 1;   // rows per lane
  constexpr bool HOIST = UR <= 8;   // J's rows loaded with G's (registers permitting)
  const int l = lane_id();
  for (int e = l; e < n * n; e += 64) {
    const int r = e % n, c = e / n;
    J[r + c * ldj] = (r == c) ? (T)1 : (T)0;
  }
  T cn = (T)0;
  if (l < n)
    for (int i = 0; i < n; ++i) { const T g = G[i + l * ldg]; cn = fma(g, g, cn); }
  const T floor2 = Prec<T>::tiny2 * wave_sum(cn);   // eps^2 ||G||_F^2
  const int np = (n + 1) & ~1, npairs = np / 2, m = np - 1;
  const int pi = l / TP, tl = l % TP;
  const T tol = Prec<T>::eps * tsqrt((T)n);
  int sweeps = -1;
  wave_lds_sync();
  for (int s = 0; s < max_sweeps && n > 1; ++s) {
    bool rotated = false;
    // round r pairs (r + pi) mod m with (r - pi) mod m (pair 0: r with m); both advance by one per round
    int pp = pi % m, qq = pi <= m ? (m - pi) % m : 0;
    for (int r = 0; r < m; ++r) {
      int p = n, q = n;
      if (pi < npairs) {
        if (pi == 0) { p = r; q = m; } else { p = pp; q = qq; }
        if (p > q) { const int x = p; p = q; q = x; }
      }
      pp = pp + 1 == m ? 0 : pp + 1;
      qq = qq + 1 == m ? 0 : qq + 1;
      if (q < n) {
        T gp[UR], gq[UR], jp[HOIST ? UR : 1], jq[HOIST ? UR : 1];
#pragma unroll
        for (int u = 0; u < UR; ++u) {
          const int i = tl + TP * u;
          gp[u] = i < n ? G[i + p * ldg] : (T)0;
          gq[u] = i < n ? G[i + q * ldg] : (T)0;
          if (HOIST) {
            jp[HOIST ? u : 0] = i < n ? J[i + p * ldj] : (T)0;
            jq[HOIST ? u : 0] = i < n ? J[i + q * ldj] : (T)0;
          }
        }
        T al = (T)0, be = (T)0, ga = (T)0;
#pragma unroll
        for (int u = 0; u < UR; ++u) {
          al = fma(gp[u], gp[u], al); be = fma(gq[u], gq[u], be); ga = fma(gp[u], gq[u], ga);
        }
        al = group_sum_dpp<TP>(al); be = group_sum_dpp<TP>(be); ga = group_sum_dpp<TP>(ga);
        T cs, sn;
        if (jacobi_rotation(al, be, ga, floor2, tol, &cs, &sn)) {
#pragma unroll
          for (int u = 0; u < UR; ++u) {
            const int i = tl + TP * u;
            if (i < n) {
              G[i + p * ldg] = cs * gp[u] - sn * gq[u];
              G[i + q * ldg] = sn * gp[u] + cs * gq[u];
              const T a = HOIST ? jp[HOIST ? u : 0] : J[i + p * ldj];
              const T c = HOIST ? jq[HOIST ? u : 0] : J[i + q * ldj];
              J[i + p * ldj] = cs * a - sn * c;
              J[i + q * ldj] = sn * a + cs * c;
            }
          }
          rotated = true;
        }
      }
      wave_lds_sync();
    }
    if (!__any(rotated)) { sweeps = s + 1; break; }
  }
  if (n <= 1) sweeps = 0;
  if (l < n) {
    T acc = (T)0;
    for (int i = 0; i < n; ++i) { const T g = G[i + l * ldg]; acc += g * g; }
    const T sv = tsqrt(acc);
    sig[l] = (sv == sv && sv <= huge_of(sv)) ? sv : (T)-1;
  }
  wave_lds_sync();
  bool bad = false;
  if (l < n) {
    const T sj = sig[l];
    int rank = 0;
    for (int i = 0; i < n; ++i) {
      const T si = sig[i];
      rank += (si > sj) || (si == sj && i < l);
    }
    perm[rank] = l;
    bad = sj < (T)0;
  }
  wave_lds_sync();
  return __any(bad) ? -2 : sweeps;
}

// One-sided Jacobi SVD of the n x n matrix G (column-major, ldg; n <= 2*(NT/TP)), rotations
// accumulated into J (initialised to I here).  On return the columns of G are mutually orthogonal,
// G_out = G_in J, sig[j] = ||G_out[:, j]|| and perm[0..n) lists columns by descending sig (ties by
// index).  TP threads cooperate on one column pair.  Returns the number of sweeps (-1: no convergence
// within max_sweeps, -2: a non-finite column norm).  G, J, sig, perm may live in LDS or global memory; flag is an LDS int.
// sig must hold n entries (it also carries the column norms used for the noise floor).
template <int TP, class T>
__device__ int jacobi_svd(T* G, int ldg, T* J, int ldj, int n, T* sig, int* perm, int* flag, int max_sweeps) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int e = tid; e < n * n; e += nt) {
    const int r = e % n, c = e / n;
    J[r + c * ldj] = (r == c) ? (T)1 : (T)0;
  }
  const int np = (n + 1) & ~1;      // players, padded to even
  const int npairs = np / 2;
  const int pi = tid / TP, tl = tid % TP;
  const T tol = Prec<T>::eps * tsqrt((T)n);
  // Columns whose squared norm is below (eps ||G||_F)^2 are rounding noise relative to the matrix: they
  // are not rotated (their singular values stay below eps ||G||_F either way), which keeps noise-level
  // columns from being stirred sweep after sweep by the rotations of the large ones.
  __shared__ T jac_floor;
  for (int c = tid; c < n; c += nt) {
    T acc = (T)0;
    for (int i = 0; i < n; ++i) { const T g = G[i + c * ldg]; acc = fma(g, g, acc); }
    sig[c] = acc;
  }
  __syncthreads();
  if (tid == 0) {
    T fro2 = (T)0;
    for (int c = 0; c < n; ++c) fro2 += sig[c];
    jac_floor = Prec<T>::tiny2 * fro2;   // eps^2 ||G||_F^2
  }
  __syncthreads();
  const T floor2 = jac_floor;
  int sweeps = -1;
  for (int s = 0; s < max_sweeps && n > 1; ++s) {
    if (tid == 0) *flag = 0;
    __syncthreads();
    for (int r = 0; r < np - 1; ++r) {
      if (pi < npairs) {
        int p, q;
        if (pi == 0) { p = np - 1; q = r; }
        else { p = (r + pi) % (np - 1); q = (r - pi + (np - 1)) % (np - 1); }
        if (p > q) { const int x = p; p = q; q = x; }
        if (q < n) {
          T al = (T)0, be = (T)0, ga = (T)0;
          for (int i = tl; i < n; i += TP) {
            const T gp = G[i + p * ldg], gq = G[i + q * ldg];
            al += gp * gp; be += gq * gq; ga += gp * gq;
          }
          al = group_sum_dpp<TP>(al); be = group_sum_dpp<TP>(be); ga = group_sum_dpp<TP>(ga);
          T cs, sn;
          if (jacobi_rotation(al, be, ga, floor2, tol, &cs, &sn)) {
            for (int i = tl; i < n; i += TP) {
              const T gp = G[i + p * ldg], gq = G[i + q * ldg];
              G[i + p * ldg] = cs * gp - sn * gq;
              G[i + q * ldg] = sn * gp + cs * gq;
              const T jp = J[i + p * ldj], jq = J[i + q * ldj];
              J[i + p * ldj] = cs * jp - sn * jq;
              J[i + q * ldj] = sn * jp + cs * jq;
            }
            if (tl == 0) *flag = 1;
          }
        }
      }
      __syncthreads();
    }
    const int rotated = *flag;
    __syncthreads();
    if (!rotated) { sweeps = s + 1; break; }
  }
  if (n <= 1) sweeps = 0;
  for (int j = tid; j < n; j += nt) {
    T acc = (T)0;
    for (int i = 0; i < n; ++i) { const T g = G[i + j * ldg]; acc += g * g; }
    const T sv = tsqrt(acc);
    // a non-finite column norm is ranked last (as -1) so perm stays a permutation; the caller sees it
    // through the returned status
    sig[j] = (sv == sv && sv <= huge_of(sv)) ? sv : (T)-1;
  }
  __syncthreads();
  for (int j = tid; j < n; j += nt) {
    const T sj = sig[j];
    int rank = 0;
    for (int i = 0; i < n; ++i) {
      const T si = sig[i];
      rank += (si > sj) || (si == sj && i < j);
    }
    perm[rank] = j;
  }
  __syncthreads();
  for (int j = 0; j < n; ++j)
    if (sig[j] < (T)0) return -2;
  return sweeps;
}

// jacobi_svd for n <= TP * UR with (n + 1) / 2 * TP <= blockDim.x (TP = 2, 4, 8 or 16 lanes per column
// pair): each pair's rows tl, tl + TP, ... of both columns (and of J when UR <= 8) are read once per round
// into registers, the group sums are DPP exchanges, and the work group synchronises once per round.  Same
// ordering, rotation, noise floor and stopping rule as jacobi_svd.  Every thread of the work group calls it.
template <int TP, int UR, class T>
__device__ int jacobi_svd_mw(T* G, int ldg, T* J, int ldj, int n, T* sig, int* perm, int max_sweeps) {
  constexpr bool HOIST = UR <= 8;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int e = tid; e < n * n; e += nt) {
    const int r = e % n, c = e / n;
    J[r + c * ldj] = (r == c) ? (T)1 : (T)0;
  }
  __shared__ T jac_floor_mw;
  for (int c = tid; c < n; c += nt) {
    T acc = (T)0;
    for (int i = 0; i < n; ++i) { const T g = G[i + c * ldg]; acc = fma(g, g, acc); }
    sig[c] = acc;
  }
  __syncthreads();
  if (tid == 0) {
    T fro2 = (T)0;
    for (int c = 0; c < n; ++c) fro2 += sig[c];
    jac_floor_mw = Prec<T>::tiny2 * fro2;   // eps^2 ||G||_F^2
  }
  __syncthreads();
  const T floor2 = jac_floor_mw;
  const int np = (n + 1) & ~1, npairs = np / 2, m = np - 1;
  const int pi = tid / TP, tl = tid % TP;
  const T tol = Prec<T>::eps * tsqrt((T)n);
  int sweeps = -1;
  for (int s = 0; s < max_sweeps && n > 1; ++s) {
    int rotated = 0;
    int pp = pi % m, qq = pi <= m ? (m - pi) % m : 0;
    for (int r = 0; r < m; ++r) {
      int p = n, q = n;
      if (pi < npairs) {
        if (pi == 0) { p = r; q = m; } else { p = pp; q = qq; }
        if (p > q) { const int x = p; p = q; q = x; }
      }
      pp = pp + 1 == m ? 0 : pp + 1;
      qq = qq + 1 == m ? 0 : qq + 1;
      if (q < n) {
        T gp[UR], gq[UR], jp[HOIST ? UR : 1], jq[HOIST ? UR : 1];
#pragma unroll
        for (int u = 0; u < UR; ++u) {
          const int i = tl + TP * u;
          gp[u] = i < n ? G[i + p * ldg] : (T)0;
          gq[u] = i < n ? G[i + q * ldg] : (T)0;
          if (HOIST) {
            jp[HOIST ? u : 0] = i < n ? J[i + p * ldj] : (T)0;
            jq[HOIST ? u : 0] = i < n ? J[i + q * ldj] : (T)0;
          }
        }
        T al = (T)0, be = (T)0, ga = (T)0;
#pragma unroll
        for (int u = 0; u < UR; ++u) {
          al = fma(gp[u], gp[u], al); be = fma(gq[u], gq[u], be); ga = fma(gp[u], gq[u], ga);
        }
        al = group_sum_dpp<TP>(al); be = group_sum_dpp<TP>(be); ga = group_sum_dpp<TP>(ga);
        T cs, sn;
        if (jacobi_rotation(al, be, ga, floor2, tol, &cs, &sn)) {
#pragma unroll
          for (int u = 0; u < UR; ++u) {
            const int i = tl + TP * u;
            if (i < n) {
              G[i + p * ldg] = cs * gp[u] - sn * gq[u];
              G[i + q * ldg] = sn * gp[u] + cs * gq[u];
              const T a = HOIST ? jp[HOIST ? u : 0] : J[i + p * ldj];
              const T c = HOIST ? jq[HOIST ? u : 0] : J[i + q * ldj];
              J[i + p * ldj] = cs * a - sn * c;
              J[i + q * ldj] = sn * a + cs * c;
            }
          }
          rotated = 1;
        }
      }
      __syncthreads();
    }
    // the round's barrier above ordered every G and J access; this one only combines the flags
    if (!__syncthreads_or(rotated)) { sweeps = s + 1; break; }
  }
  if (n <= 1) sweeps = 0;
  for (int j = tid; j < n; j += nt) {
    T acc = (T)0;
    for (int i = 0; i < n; ++i) { const T g = G[i + j * ldg]; acc += g * g; }
    const T sv = tsqrt(acc);
    sig[j] = (sv == sv && sv <= huge_of(sv)) ? sv : (T)-1;
  }
  __syncthreads();
  for (int j = tid; j < n; j += nt) {
    const T sj = sig[j];
    int rank = 0;
    for (int i = 0; i < n; ++i) {
      const T si = sig[i];
      rank += (si > sj) || (si == sj && i < j);
    }
    perm[rank] = j;
  }
  __syncthreads();   // a full barrier with its memory fence: callers read sig and perm next
  for (int j = 0; j < n; ++j)
    if (sig[j] < (T)0) return -2;
  return sweeps;
}

// jacobi_svd through its faster forms: one wave (jacobi_svd_w0; 16, 8 or 4 lanes per column pair) when
// n <= 32, in which case G, J, sig and perm must be in LDS; with 512 threads, 16 lanes per pair up to
// n = 64 and 8 lanes per pair up to n = 128 (jacobi_svd_mw).  Every thread of the work group calls it; flag
// is an LDS int (it also carries the one-wave result to the other waves).
template <int TP, class T>
__device__ int jacobi_svd_auto(T* G, int ldg, T* J, int ldj, int n, T* sig, int* perm, int* flag, int max_sweeps) {
  if (n <= 32) {
    if (wave_id() == 0) {
      int r;
      if (n <= 8) r = jacobi_svd_w0<16>(G, ldg, J, ldj, n, sig, perm, max_sweeps);
      else if (n <= 16) r = jacobi_svd_w0<8>(G, ldg, J, ldj, n, sig, perm, max_sweeps);
      else r = jacobi_svd_w0<4>(G, ldg, J, ldj, n, sig, perm, max_sweeps);
      if (lane_id() == 0) *flag = r;
    }
    __syncthreads();
    return *flag;
  }
  if (blockDim.x >= 512) {
    if (n <= 64) return jacobi_svd_mw<16, 4>(G, ldg, J, ldj, n, sig, perm, max_sweeps);
    if (n <= 128) return jacobi_svd_mw<8, 16>(G, ldg, J, ldj, n, sig, perm, max_sweeps);
  }
  return jacobi_svd<TP>(G, ldg, J, ldj, n, sig, perm, flag, max_sweeps);
}

}  // namespace tlr
