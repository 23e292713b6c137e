// Work-group-wide small dense linear algebra kept in registers / LDS:
//  * RegPanel: an m x n panel (m <= 64*RPL, n <= CPW*NW) distributed over NW waves; wave w owns the
//    columns j = w + NW*jj (jj < CPW), lane l owns the rows p = l + 64*t (t < RPL).  Householder QR with
//    LAPACK dlarfg/dgeqr2 conventions, explicit Q in place (dorg2r), all without spilling the panel.
//  * panel_coef / panel_axpy: projections of the panel against a global matrix (block Gram-Schmidt).
//  * jacobi_svd: one-sided (Hestenes) Jacobi on the columns of a small square matrix, round-robin
//    ordering, rotations accumulated into J.
#pragma once
#include "common.hpp"

namespace tlr {

template <int RPL, int NW, int CPW_ = 8>
struct RegPanel {
  static constexpr int CPW = CPW_;
  static constexpr int NT = 64 * NW;
  double a[CPW][RPL];

  __device__ __forceinline__ static int col_of(int w, int jj) { return w + NW * jj; }

  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj)
#pragma unroll
      for (int t = 0; t < RPL; ++t) a[jj][t] = 0.0;
  }

  // A: column-major, leading dimension lda; entries outside m x n read as zero.
  __device__ __forceinline__ void load(const double* A, int64_t lda, int m, int n) {
    const int w = wave_id(), l = lane_id();
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) {
      const int j = col_of(w, jj);
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int p = l + 64 * t;
        a[jj][t] = (j < n && p < m) ? A[p + (int64_t)j * lda] : 0.0;
      }
    }
  }

  __device__ __forceinline__ void store(double* A, int64_t lda, int m, int n) const {
    const int w = wave_id(), l = lane_id();
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) {
      const int j = col_of(w, jj);
      if (j < n) {
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
          const int p = l + 64 * t;
          if (p < m) A[p + (int64_t)j * lda] = a[jj][t];
        }
      }
    }
  }

  // Copy column slot jc (wave-uniform) into x.
  // Apply the reflector in vb (rows >= c nonzero) with factor tc to every own column j > c.
  __device__ __forceinline__ void apply_reflector(int c, int n, const double* vb, double tc) {
    const int w = wave_id(), l = lane_id();
    double v[RPL];
#pragma unroll
    for (int t = 0; t < RPL; ++t) v[t] = vb[l + 64 * t];
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) {
      const int j = col_of(w, jj);
      if (j > c && j < n) {
        double d = 0.0;
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
  __device__ __forceinline__ void geqr2(int m, int n, double* vbuf, double* tau) {
    const int w = wave_id(), l = lane_id();
#pragma unroll
    for (int jc = 0; jc < CPW; ++jc) {
      for (int wo = 0; wo < NW; ++wo) {
        const int c = wo + NW * jc;
        if (c >= n) break;
        double* vb = vbuf + (c & 1) * (64 * RPL);
        if (w == wo) {
          double part = 0.0, alpha_l = 0.0;
#pragma unroll
          for (int t = 0; t < RPL; ++t) {
            const int p = l + 64 * t;
            const double x = a[jc][t];
            if (p > c && p < m) part += x * x;
            if (p == c) alpha_l = x;
          }
          const double sigma = wave_sum(part);
          const double alpha = wave_sum(alpha_l);  // exactly one lane holds it
          double beta, tc, scal;
          if (sigma == 0.0) {
            tc = 0.0; beta = alpha; scal = 0.0;
          } else {
            const double nrm = sqrt(alpha * alpha + sigma);
            beta = alpha >= 0.0 ? -nrm : nrm;
            tc = (beta - alpha) / beta;
            scal = 1.0 / (alpha - beta);
          }
#pragma unroll
          for (int t = 0; t < RPL; ++t) {
            const int p = l + 64 * t;
            double v = 0.0;
            if (p == c) v = 1.0;
            else if (p > c && p < m) v = a[jc][t] * scal;
            vb[p] = v;
            if (p == c) a[jc][t] = beta;
            else if (p > c) a[jc][t] = (p < m) ? v : 0.0;
          }
          if (l == 0) tau[c] = tc;
        }
        __syncthreads();
        apply_reflector(c, n, vb, tau[c]);
      }
    }
    __syncthreads();
  }

  // Write the n x n upper triangle R (zeros below) to R (column-major, ldr).  Call after geqr2.
  __device__ __forceinline__ void store_r(double* R, int ldr, int n) const {
    const int w = wave_id(), l = lane_id();
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) {
      const int j = col_of(w, jj);
      if (j < n) {
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
          const int p = l + 64 * t;
          if (p < n) R[p + j * ldr] = (p <= j) ? a[jj][t] : 0.0;
        }
      }
    }
  }

  // Same, transposed: Rt[j + p*ldr] = R[p][j].
  __device__ __forceinline__ void store_rt(double* Rt, int ldr, int n) const {
    const int w = wave_id(), l = lane_id();
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) {
      const int j = col_of(w, jj);
      if (j < n) {
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
          const int p = l + 64 * t;
          if (p < n) Rt[j + p * ldr] = (p <= j) ? a[jj][t] : 0.0;
        }
      }
    }
  }

  // Explicit Q (m x n) in place from the reflectors left by geqr2 (dorg2r), columns in reverse order.
  __device__ __forceinline__ void org2r(int m, int n, double* vbuf, const double* tau) {
    const int w = wave_id(), l = lane_id();
#pragma unroll
    for (int jc = CPW - 1; jc >= 0; --jc) {
      for (int wo = NW - 1; wo >= 0; --wo) {
        const int c = wo + NW * jc;
        if (c >= n) continue;
        double* vb = vbuf + (c & 1) * (64 * RPL);
        const double tc = tau[c];
        if (w == wo) {
#pragma unroll
          for (int t = 0; t < RPL; ++t) {
            const int p = l + 64 * t;
            double v = 0.0;
            if (p == c) v = 1.0;
            else if (p > c && p < m) v = a[jc][t];
            vb[p] = v;
            // column c of Q = H_c e_c = e_c - tau v
            a[jc][t] = (p == c) ? 1.0 - tc : ((p > c && p < m) ? -tc * v : 0.0);
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
template <int RPL, int NW, int CPW>
__device__ __forceinline__ void panel_coef(const RegPanel<RPL, NW, CPW>& P, const double* A, int64_t lda, int m,
                                           int na, int n, const double* sc, double* C, int ldc) {
  constexpr int AU = CPW <= 4 ? 2 : 1;
  const int w = wave_id(), l = lane_id();
  for (int a0 = 0; a0 < na; a0 += AU) {
    double av[AU][RPL];
#pragma unroll
    for (int u = 0; u < AU; ++u)
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int p = l + 64 * t;
        av[u][t] = (a0 + u < na && p < m) ? A[p + (int64_t)(a0 + u) * lda] : 0.0;
      }
#pragma unroll
    for (int u = 0; u < AU; ++u) {
      const int a = a0 + u;
      if (a < na) {
        const double s = sc ? sc[a] : 1.0;
#pragma unroll
        for (int jj = 0; jj < CPW; ++jj) {
          const int j = RegPanel<RPL, NW, CPW>::col_of(w, jj);
          if (j < n) {
            double d = 0.0;
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
template <int RPL, int NW, int CPW>
__device__ __forceinline__ void panel_axpy(RegPanel<RPL, NW, CPW>& P, const double* A, int64_t lda, int m, int na,
                                           int n, const double* sc, const double* C, int ldc, double alpha) {
  constexpr int AU = CPW <= 4 ? 2 : 1;
  const int w = wave_id(), l = lane_id();
  for (int a0 = 0; a0 < na; a0 += AU) {
    double av[AU][RPL];
#pragma unroll
    for (int u = 0; u < AU; ++u)
#pragma unroll
      for (int t = 0; t < RPL; ++t) {
        const int p = l + 64 * t;
        av[u][t] = (a0 + u < na && p < m) ? A[p + (int64_t)(a0 + u) * lda] : 0.0;
      }
#pragma unroll
    for (int u = 0; u < AU; ++u) {
      const int a = a0 + u;
      if (a < na) {
        const double s = alpha * (sc ? sc[a] : 1.0);
#pragma unroll
        for (int jj = 0; jj < CPW; ++jj) {
          const int j = RegPanel<RPL, NW, CPW>::col_of(w, jj);
          if (j < n) {
            const double f = s * C[a + j * ldc];
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
__device__ __forceinline__ bool jacobi_rotation(double al, double be, double ga, double floor2, double tol,
                                                double* cs, double* sn) {
  if (!(ga != 0.0 && al > floor2 && be > floor2 && ga * ga > tol * tol * al * be)) return false;
  const double d = be - al, h = 2.0 * ga;
  const double t = d == 0.0 ? 1.0 : (d > 0.0 ? h : -h) / (fabs(d) + sqrt(fma(d, d, h * h)));
  const double c = 1.0 / sqrt(fma(t, t, 1.0));
  *cs = c;
  *sn = c * t;
  return true;
}

// jacobi_svd for n <= 128 / TP (TP = 2, 4, 8 or 16 lanes per column pair), run by one wave alone: its 64
// lanes hold every column pair of a round (rows tl, tl + TP, ... of both columns and of J in registers), so
// rounds are separated by a wave-level fence instead of a work-group barrier, each column is read once per
// round, and the group sums are DPP exchanges.  Same ordering, rotation, noise floor and stopping rule as
// jacobi_svd.  G, J, sig and perm must be in LDS (the wave-level fence orders LDS accesses only).  Only one
// wave may call it; the caller synchronises the work group after.
template <int TP>
__device__ int jacobi_svd_w0(double* G, int ldg, double* J, int ldj, int n, double* sig, int* perm,
                             int max_sweeps) {
  constexpr int UR = 128 / (TP * TP) > 0 ? 128 / (TP * TP) : 1;   // rows per lane
  constexpr bool HOIST = UR <= 8;   // J's rows loaded with G's (registers permitting)
  const int l = lane_id();
  for (int e = l; e < n * n; e += 64) {
    const int r = e % n, c = e / n;
    J[r + c * ldj] = (r == c) ? 1.0 : 0.0;
  }
  double cn = 0.0;
  if (l < n)
    for (int i = 0; i < n; ++i) { const double g = G[i + l * ldg]; cn = fma(g, g, cn); }
  const double floor2 = 4.930380657631324e-32 * wave_sum(cn);   // eps^2 ||G||_F^2
  const int np = (n + 1) & ~1, npairs = np / 2, m = np - 1;
  const int pi = l / TP, tl = l % TP;
  const double tol = 2.220446049250313e-16 * sqrt((double)n);
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
        double gp[UR], gq[UR], jp[HOIST ? UR : 1], jq[HOIST ? UR : 1];
#pragma unroll
        for (int u = 0; u < UR; ++u) {
          const int i = tl + TP * u;
          gp[u] = i < n ? G[i + p * ldg] : 0.0;
          gq[u] = i < n ? G[i + q * ldg] : 0.0;
          if (HOIST) {
            jp[HOIST ? u : 0] = i < n ? J[i + p * ldj] : 0.0;
            jq[HOIST ? u : 0] = i < n ? J[i + q * ldj] : 0.0;
          }
        }
        double al = 0.0, be = 0.0, ga = 0.0;
#pragma unroll
        for (int u = 0; u < UR; ++u) {
          al = fma(gp[u], gp[u], al); be = fma(gq[u], gq[u], be); ga = fma(gp[u], gq[u], ga);
        }
        al = group_sum_dpp<TP>(al); be = group_sum_dpp<TP>(be); ga = group_sum_dpp<TP>(ga);
        double cs, sn;
        if (jacobi_rotation(al, be, ga, floor2, tol, &cs, &sn)) {
#pragma unroll
          for (int u = 0; u < UR; ++u) {
            const int i = tl + TP * u;
            if (i < n) {
              G[i + p * ldg] = cs * gp[u] - sn * gq[u];
              G[i + q * ldg] = sn * gp[u] + cs * gq[u];
              const double a = HOIST ? jp[HOIST ? u : 0] : J[i + p * ldj];
              const double c = HOIST ? jq[HOIST ? u : 0] : J[i + q * ldj];
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
    double acc = 0.0;
    for (int i = 0; i < n; ++i) { const double g = G[i + l * ldg]; acc += g * g; }
    const double sv = sqrt(acc);
    sig[l] = (sv == sv && sv <= 1.7976931348623157e308) ? sv : -1.0;
  }
  wave_lds_sync();
  bool bad = false;
  if (l < n) {
    const double sj = sig[l];
    int rank = 0;
    for (int i = 0; i < n; ++i) {
      const double si = sig[i];
      rank += (si > sj) || (si == sj && i < l);
    }
    perm[rank] = l;
    bad = sj < 0.0;
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
template <int TP>
__device__ int jacobi_svd(double* G, int ldg, double* J, int ldj, int n, double* sig, int* perm, int* flag,
                          int max_sweeps) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int e = tid; e < n * n; e += nt) {
    const int r = e % n, c = e / n;
    J[r + c * ldj] = (r == c) ? 1.0 : 0.0;
  }
  const int np = (n + 1) & ~1;      // players, padded to even
  const int npairs = np / 2;
  const int pi = tid / TP, tl = tid % TP;
  const double tol = 2.220446049250313e-16 * sqrt((double)n);
  // Columns whose squared norm is below (eps ||G||_F)^2 are rounding noise relative to the matrix: they
  // are not rotated (their singular values stay below eps ||G||_F either way), which keeps noise-level
  // columns from being stirred sweep after sweep by the rotations of the large ones.
  __shared__ double jac_floor;
  for (int c = tid; c < n; c += nt) {
    double acc = 0.0;
    for (int i = 0; i < n; ++i) { const double g = G[i + c * ldg]; acc = fma(g, g, acc); }
    sig[c] = acc;
  }
  __syncthreads();
  if (tid == 0) {
    double fro2 = 0.0;
    for (int c = 0; c < n; ++c) fro2 += sig[c];
    jac_floor = 4.930380657631324e-32 * fro2;   // eps^2 ||G||_F^2
  }
  __syncthreads();
  const double floor2 = jac_floor;
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
          double al = 0.0, be = 0.0, ga = 0.0;
          for (int i = tl; i < n; i += TP) {
            const double gp = G[i + p * ldg], gq = G[i + q * ldg];
            al += gp * gp; be += gq * gq; ga += gp * gq;
          }
          al = group_sum_dpp<TP>(al); be = group_sum_dpp<TP>(be); ga = group_sum_dpp<TP>(ga);
          double cs, sn;
          if (jacobi_rotation(al, be, ga, floor2, tol, &cs, &sn)) {
            for (int i = tl; i < n; i += TP) {
              const double gp = G[i + p * ldg], gq = G[i + q * ldg];
              G[i + p * ldg] = cs * gp - sn * gq;
              G[i + q * ldg] = sn * gp + cs * gq;
              const double jp = J[i + p * ldj], jq = J[i + q * ldj];
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
    double acc = 0.0;
    for (int i = 0; i < n; ++i) { const double g = G[i + j * ldg]; acc += g * g; }
    const double sv = sqrt(acc);
    // a non-finite column norm is ranked last (as -1) so perm stays a permutation; the caller sees it
    // through the returned status
    sig[j] = (sv == sv && sv <= 1.7976931348623157e308) ? sv : -1.0;
  }
  __syncthreads();
  for (int j = tid; j < n; j += nt) {
    const double sj = sig[j];
    int rank = 0;
    for (int i = 0; i < n; ++i) {
      const double si = sig[i];
      rank += (si > sj) || (si == sj && i < j);
    }
    perm[rank] = j;
  }
  __syncthreads();
  for (int j = 0; j < n; ++j)
    if (sig[j] < 0.0) return -2;
  return sweeps;
}

// jacobi_svd for n <= TP * UR with (n + 1) / 2 * TP <= blockDim.x (TP = 2, 4, 8 or 16 lanes per column
// pair): each pair's rows tl, tl + TP, ... of both columns (and of J when UR <= 8) are read once per round
// into registers, the group sums are DPP exchanges, and the work group synchronises once per round.  Same
// ordering, rotation, noise floor and stopping rule as jacobi_svd.  Every thread of the work group calls it.
template <int TP, int UR>
__device__ int jacobi_svd_mw(double* G, int ldg, double* J, int ldj, int n, double* sig, int* perm,
                             int max_sweeps) {
  constexpr bool HOIST = UR <= 8;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int e = tid; e < n * n; e += nt) {
    const int r = e % n, c = e / n;
    J[r + c * ldj] = (r == c) ? 1.0 : 0.0;
  }
  __shared__ double jac_floor_mw;
  for (int c = tid; c < n; c += nt) {
    double acc = 0.0;
    for (int i = 0; i < n; ++i) { const double g = G[i + c * ldg]; acc = fma(g, g, acc); }
    sig[c] = acc;
  }
  __syncthreads();
  if (tid == 0) {
    double fro2 = 0.0;
    for (int c = 0; c < n; ++c) fro2 += sig[c];
    jac_floor_mw = 4.930380657631324e-32 * fro2;   // eps^2 ||G||_F^2
  }
  __syncthreads();
  const double floor2 = jac_floor_mw;
  const int np = (n + 1) & ~1, npairs = np / 2, m = np - 1;
  const int pi = tid / TP, tl = tid % TP;
  const double tol = 2.220446049250313e-16 * sqrt((double)n);
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
        double gp[UR], gq[UR], jp[HOIST ? UR : 1], jq[HOIST ? UR : 1];
#pragma unroll
        for (int u = 0; u < UR; ++u) {
          const int i = tl + TP * u;
          gp[u] = i < n ? G[i + p * ldg] : 0.0;
          gq[u] = i < n ? G[i + q * ldg] : 0.0;
          if (HOIST) {
            jp[HOIST ? u : 0] = i < n ? J[i + p * ldj] : 0.0;
            jq[HOIST ? u : 0] = i < n ? J[i + q * ldj] : 0.0;
          }
        }
        double al = 0.0, be = 0.0, ga = 0.0;
#pragma unroll
        for (int u = 0; u < UR; ++u) {
          al = fma(gp[u], gp[u], al); be = fma(gq[u], gq[u], be); ga = fma(gp[u], gq[u], ga);
        }
        al = group_sum_dpp<TP>(al); be = group_sum_dpp<TP>(be); ga = group_sum_dpp<TP>(ga);
        double cs, sn;
        if (jacobi_rotation(al, be, ga, floor2, tol, &cs, &sn)) {
#pragma unroll
          for (int u = 0; u < UR; ++u) {
            const int i = tl + TP * u;
            if (i < n) {
              G[i + p * ldg] = cs * gp[u] - sn * gq[u];
              G[i + q * ldg] = sn * gp[u] + cs * gq[u];
              const double a = HOIST ? jp[HOIST ? u : 0] : J[i + p * ldj];
              const double c = HOIST ? jq[HOIST ? u : 0] : J[i + q * ldj];
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
    double acc = 0.0;
    for (int i = 0; i < n; ++i) { const double g = G[i + j * ldg]; acc += g * g; }
    const double sv = sqrt(acc);
    sig[j] = (sv == sv && sv <= 1.7976931348623157e308) ? sv : -1.0;
  }
  __syncthreads();
  for (int j = tid; j < n; j += nt) {
    const double sj = sig[j];
    int rank = 0;
    for (int i = 0; i < n; ++i) {
      const double si = sig[i];
      rank += (si > sj) || (si == sj && i < j);
    }
    perm[rank] = j;
  }
  __syncthreads();   // a full barrier with its memory fence: callers read sig and perm next
  for (int j = 0; j < n; ++j)
    if (sig[j] < 0.0) return -2;
  return sweeps;
}

// jacobi_svd through its faster forms: one wave (jacobi_svd_w0; 16, 8 or 4 lanes per column pair) when
// n <= 32, in which case G, J, sig and perm must be in LDS; with 512 threads, 16 lanes per pair up to
// n = 64 and 8 lanes per pair up to n = 128 (jacobi_svd_mw).  Every thread of the work group calls it; flag
// is an LDS int (it also carries the one-wave result to the other waves).
template <int TP>
__device__ int jacobi_svd_auto(double* G, int ldg, double* J, int ldj, int n, double* sig, int* perm, int* flag,
                               int max_sweeps) {
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
