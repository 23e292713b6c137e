// Right-looking TLR Cholesky on one GPU, one launch per phase of step k.
//
// Reference semantics (codes/17_tlr_cholesky_pilot.py:171-205, 230-240, 259-266):
//   L_kk = chol(A_kk)                         (non-PD pivot -> completed=False, failed_at=k)
//   V_ik <- L_kk^{-1} V_ik, i > k             (panel; U_ik unchanged)
//   A_ii -= (U_ik W) U_ik^T, W = V_ik^T V_ik  (diagonal low-rank update)
//   (U_ij, V_ij) <- recompress([U_ij, -U_ik W], [V_ij, U_jk]), W = V_ik^T V_jk, k < j < i
//   z = L^{-1} y by tile forward substitution; log L = -(z^T z + 2 sum log diag L_kk + N log 2 pi)/2.
// The forward solve is folded into the factorization (right-looking): z_k = L_kk^{-1} rhs_k at step k,
// then rhs_i -= U_ik (V_ik^T z_k) for i > k; it uses the same final factors as the reference's
// left-looking solve.  Updates with r_ik = 0 or r_jk = 0 are skipped, as in the reference.
#include <type_traits>

#include "common.hpp"
#include "panel.hpp"
#include "round.hpp"
#include "tlr_kernels.hpp"

namespace tlr {

constexpr int NB = 32;   // inner panel width of the diagonal-tile factorization
typedef double dbl4 __attribute__((ext_vector_type(4)));

// status[ST_FAILED] holds the first failing diagonal tile (INT_MAX when none).
__device__ __forceinline__ bool failed(const int* status) {
  return __hip_atomic_load(&status[ST_FAILED], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0x7fffffff;
}

// ------------------------------------------------------------------------------------------ POTRF
// One work group (512 threads) factors the b x b diagonal tile k in place (lower), writes the inverses
// of its 32 x 32 diagonal blocks to Dinv and sum(log diag L_kk) (tile rows only) to logdet_part[k].
__global__ __launch_bounds__(512) void k_potrf(SlotArgs sa, int k) {
  const int slot = blockIdx.z;
  const SlotPtrs P = sa.slot(slot);
  if (failed(P.status)) return;
  const int b = sa.b;
  const int bk = (int)min<int64_t>(b, sa.n - (int64_t)k * b);
  double* A = P.D + (int64_t)k * b * b;
  double* Dv = P.Dinv + (int64_t)k * sa.npan * NB * NB;
  __shared__ double Lp[NB][512];
  __shared__ double Lb[NB][NB + 1];
  __shared__ double Li[NB][NB + 1];
  __shared__ int s_fail;
  const int tid = threadIdx.x, w = wave_id(), l = lane_id();
  double logacc = 0.0;
  for (int p0 = 0; p0 < b; p0 += NB) {
    const int wd = min(NB, b - p0);
    if (w == 0) {
      double a[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) a[c] = (l < wd && c <= l) ? A[(p0 + l) + (int64_t)(p0 + c) * b] : 0.0;
      int fail = 0;
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        if (j < wd) {
          const double d = readlane_d(a[j], j);   // j is a compile-time lane after unrolling
          if (!(d > 0.0)) fail = 1;
          const double ljj = sqrt(d);
          if (l == j) a[j] = ljj;
          else if (l > j) a[j] = a[j] / ljj;
          const double lj = a[j];
#pragma unroll
          for (int c = j + 1; c < NB; ++c) {
            const double lcj = readlane_d(lj, c);
            if (l >= c) a[c] -= lj * lcj;
          }
        }
      }
      double lg = 0.0;
      if (l < wd) {
#pragma unroll
        for (int c = 0; c < NB; ++c) {
          if (c <= l) {
            Lb[l][c] = a[c];
            A[(p0 + l) + (int64_t)(p0 + c) * b] = a[c];
          } else {
            Lb[l][c] = 0.0;
          }
        }
#pragma unroll
        for (int c = 0; c < NB; ++c)
          if (c == l && p0 + l < bk) lg = log(a[c]);
      }
      lg = wave_sum(lg);
      if (l == 0) { logacc += lg; s_fail = fail; }
    }
    __syncthreads();
    if (s_fail) {
      if (tid == 0) atomicMin(&P.status[ST_FAILED], k);
      return;
    }
    // inverse of the diagonal block: lane m of wave 0 solves for column m
    if (w == 0 && l < wd) {
      double x[NB];
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        double v = 0.0;
        if (r < wd && r >= l) {
          double s = (r == l) ? 1.0 : 0.0;
#pragma unroll
          for (int c = 0; c < r; ++c) s -= Lb[r][c] * x[c];
          v = s / Lb[r][r];
        }
        x[r] = v;
      }
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        Li[r][l] = x[r];
        Dv[(p0 / NB) * NB * NB + r + l * NB] = (r < wd) ? x[r] : 0.0;
      }
    }
    __syncthreads();
    // panel below: L_i,P = A_i,P Linv^T
    const int i = p0 + wd + tid;
    if (i < b) {
      double av[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) av[c] = (c < wd) ? A[i + (int64_t)(p0 + c) * b] : 0.0;
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m <= c; ++m) s += av[m] * Li[c][m];
        if (c < wd) {
          A[i + (int64_t)(p0 + c) * b] = s;
          Lp[c][i] = s;
        } else {
          Lp[c][i] = 0.0;
        }
      }
    }
    __syncthreads();
    // trailing update A_22 -= L_21 L_21^T of the lower triangle of rows/cols >= p0 + wd on the matrix cores:
    // 16 x 16 blocks (I >= J), four per wave at a time (their A entries are loaded before the products so
    // that the loads' latency overlaps the MFMA chains).  The MFMA (v_mfma_f64_16x16x4f64) forms
    // D[x][y] = sum_c L[t0 + 16 J + x][c] L[t0 + 16 I + y][c], 4 panel columns per instruction, so item v of
    // lane l is entry (t0 + 16 I + l % 16, t0 + 16 J + l / 16 + 4 v): 16 consecutive rows of a column per 16
    // lanes.  The chained products add in the panel-column order of a sequential FMA loop.
    const int t0 = p0 + wd;
    const int m = b - t0;
    if (m > 0) {
      constexpr int GB = 4;
      const int n16 = (m + 15) / 16;
      const int nblk16 = n16 * (n16 + 1) / 2;
      const int lr = l & 15, lq = l >> 4;
      for (int e0 = GB * w; e0 < nblk16; e0 += GB * 8) {
        int ri[GB], rj[GB], cj[GB];
        double cv[GB][4];
#pragma unroll
        for (int g = 0; g < GB; ++g) {
          const int e = e0 + g;
          ri[g] = b;   // blocks past the end: every entry masked
          rj[g] = b;
          cj[g] = b;
          if (e < nblk16) {
            int I = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
            while ((I + 1) * (I + 2) / 2 <= e) ++I;
            while (I * (I + 1) / 2 > e) --I;
            const int J = e - I * (I + 1) / 2;
            ri[g] = t0 + 16 * I + lr;
            rj[g] = t0 + 16 * J + lr;
            cj[g] = t0 + 16 * J + lq;
          }
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int jc = cj[g] + 4 * v;
            cv[g][v] = (ri[g] < b && jc <= ri[g]) ? A[ri[g] + (int64_t)jc * b] : 0.0;
          }
        }
        dbl4 acc[GB];
#pragma unroll
        for (int g = 0; g < GB; ++g) acc[g] = (dbl4){0.0, 0.0, 0.0, 0.0};
        for (int s4 = 0; s4 < wd; s4 += 4) {
          const int c = s4 + lq;
#pragma unroll
          for (int g = 0; g < GB; ++g) {
            const double av = (rj[g] < b && c < wd) ? Lp[c][rj[g]] : 0.0;
            const double bv = (ri[g] < b && c < wd) ? Lp[c][ri[g]] : 0.0;
            acc[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[g], 0, 0, 0);
          }
        }
#pragma unroll
        for (int g = 0; g < GB; ++g)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int jc = cj[g] + 4 * v;
            if (ri[g] < b && jc <= ri[g]) A[ri[g] + (int64_t)jc * b] = cv[g][v] - acc[g][v];
          }
      }
    }
    __syncthreads();
  }
  if (tid == 0) P.logdet_part[k] = logacc;
}

// ----------------------------------------------------------------------------------------- TRTRI
// Explicit inverse of L_kk (lower, b x b) into the slot's Linv, one work group per 32-column block q:
// L X = E_q (the identity's columns of block q) by forward substitution over the 32-row panels p >= q,
// X_p = Dinv_p (E_q,p - sum_{q <= t < p} L_pt X_t), with the inverted diagonal blocks from POTRF.  Rows
// above block q are zero and are not written (Linv is only read on and below the diagonal).
__global__ __launch_bounds__(256) void k_trtri(SlotArgs sa, int k) {
  const int slot = blockIdx.z;
  const SlotPtrs P = sa.slot(slot);
  if (failed(P.status)) return;
  const int b = sa.b;
  const int c0 = NB * blockIdx.x;
  const double* L = P.D + (int64_t)k * b * b;
  const double* Dv = P.Dinv + (int64_t)k * sa.npan * NB * NB;
  double* X = P.Linv + (int64_t)c0 * b;
  const int cw = min(NB, b - c0);
  __shared__ double Lb[NB][NB + 1];
  __shared__ double Xb[NB][NB + 1];
  __shared__ double tmp[NB][NB + 1];
  __shared__ double Dl[NB][NB + 1];
  const int tid = threadIdx.x;
  // thread -> (row ll, columns cg, cg + 8, cg + 16, cg + 24)
  const int ll = tid % NB, cg = tid / NB;
  for (int p0 = c0; p0 < b; p0 += NB) {
    const int wd = min(NB, b - p0);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q0 = c0; q0 < p0; q0 += NB) {
      for (int e = tid; e < NB * NB; e += 256) {
        const int rr = e % NB, qq = e / NB;
        Lb[rr][qq] = (rr < wd) ? L[(p0 + rr) + (int64_t)(q0 + qq) * b] : 0.0;
        Xb[rr][qq] = (qq < cw) ? X[(q0 + rr) + (int64_t)qq * b] : 0.0;
      }
      __syncthreads();
#pragma unroll 4
      for (int qq = 0; qq < NB; ++qq) {
        const double lv = Lb[ll][qq];
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] = fma(lv, Xb[qq][cg + 8 * u], acc[u]);
      }
      __syncthreads();
    }
    for (int e = tid; e < NB * NB; e += 256) Dl[e % NB][e / NB] = Dv[(p0 / NB) * NB * NB + e];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int c = cg + 8 * u;
      tmp[ll][c] = (ll < wd) ? ((p0 + ll == c0 + c) ? 1.0 : 0.0) - acc[u] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int c = cg + 8 * u;
      if (c < cw && ll < wd) {
        double s = 0.0;
        for (int m = 0; m <= ll; ++m) s = fma(Dl[ll][m], tmp[m][c], s);
        X[(p0 + ll) + (int64_t)c * b] = s;
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------ TRSM
// Item 0: rhs_k -> z_k = L_kk^{-1} rhs_k.  Item m >= 1: tile i = k + m, V_ik <- L_kk^{-1} V_ik, then
// W_i = V_ik^T V_ik (r x r) and X_i = U_ik W_i (b x r) for the diagonal update.  The solve is the product
// with the explicit inverse from k_trtri, 8 columns at a time staged through LDS; thread t forms rows t
// and brows - 1 - t so that every thread runs the same number of inner steps.
constexpr int TRSM_LDS = 6144;   // doubles: 512 rows x 8 columns, or W_i (64 x 64) + 32 rows of U_ik
__global__ __launch_bounds__(256) void k_trsm(SlotArgs sa, int k) {
  const int slot = blockIdx.z;
  const SlotPtrs P = sa.slot(slot);
  if (failed(P.status)) return;
  const int b = sa.b, rc = sa.rank_cap;
  const int item = blockIdx.x;
  const double* Li = P.Linv;
  double* X;
  int r;
  int64_t tik = -1;
  if (item == 0) {
    X = P.rhs + (int64_t)k * b;
    r = 1;
  } else {
    const int i = k + item;
    tik = tri_index(i, k);
    r = P.rank[tik];
    X = P.V + tik * sa.fstride;
  }
  if (r <= 0 || r > rc) {
    if (r != 0) atomicMax(&P.status[ST_GUARD], 1);
    return;
  }
  const int bk = (int)min<int64_t>(b, sa.n - (int64_t)k * b);
  const int brows = (item == 0) ? bk : b;   // rhs holds exactly the tile's rows
  __shared__ __align__(16) double Xs[TRSM_LDS];
  const int tid = threadIdx.x;
  const int half = (brows + 1) / 2;
  for (int cb = 0; cb < r; cb += 8) {
    const int cw = min(8, r - cb);
    for (int e = tid; e < brows * 8; e += 256) {
      const int q = e >> 3, c = e & 7;
      Xs[e] = (c < cw) ? X[q + (int64_t)(cb + c) * b] : 0.0;
    }
    __syncthreads();
    if (tid < half) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int p = h ? brows - 1 - tid : tid;
        if (h && p == tid) break;
        double acc[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[c] = 0.0;
        const double* lrow = Li + p;
        for (int q = 0; q <= p; ++q) {
          const double lv = lrow[(int64_t)q * b];
          const double2* xq = reinterpret_cast<const double2*>(Xs + 8 * q);
#pragma unroll
          for (int c2 = 0; c2 < 4; ++c2) {
            const double2 xv = xq[c2];
            acc[2 * c2] = fma(lv, xv.x, acc[2 * c2]);
            acc[2 * c2 + 1] = fma(lv, xv.y, acc[2 * c2 + 1]);
          }
        }
#pragma unroll
        for (int c = 0; c < 8; ++c)
          if (c < cw) X[p + (int64_t)(cb + c) * b] = acc[c];
      }
    }
    __syncthreads();
  }
  if (item == 0) return;
  // W_i = V^T V and X_i = U W
  double* Wi = P.Wd + (int64_t)(k + item) * rc * rc;
  double* Xi = P.Xd + (int64_t)(k + item) * b * rc;
  double* U = P.U + tik * sa.fstride;
  double* Wsh = Xs;                  // r x r
  double* Ub = Xs + 64 * 64;         // 32 rows x r
  const int w = wave_id(), l = lane_id();
  if (sa.mixed) {
    // mixed rule: storage precision of L_ik under the full budget (17:183-184), before W_i and X_i use it;
    // X_i's slot is the Gram workspace (b x rank_cap >= 3 r^2)
    __shared__ double red[8];
    const int bi = (int)min<int64_t>(b, sa.n - (int64_t)(k + item) * b);
    const int pr = 3 * r * r <= b * rc ? round_pair_wg(U, b, bi, X, b, bk, r, sa.thr[2 * slot + 1], Xi, red) : 64;
    if (tid == 0) P.prec[tik] = pr;
  }
  for (int e = w; e < r * r; e += 4) {
    const int a = e % r, c = e / r;
    if (a > c) continue;
    double s = 0.0;
    for (int p = l; p < b; p += 64) s = fma(X[p + (int64_t)a * b], X[p + (int64_t)c * b], s);
    s = wave_sum(s);
    if (l == 0) { Wsh[a + c * r] = s; Wsh[c + a * r] = s; Wi[a + c * r] = s; Wi[c + a * r] = s; }
  }
  __syncthreads();
  if (tid == 0) {   // tnorm: ||U_ik||_F (stored with V orthonormal) times ||V_ik||_F = sqrt(tr W_i) after the solve
    double tr = 0.0;
    for (int a = 0; a < r; ++a) tr += Wsh[a + a * r];
    P.tnorm[tik] *= sqrt(tr);
  }
  // X_i = U W, rows staged through LDS 32 at a time
  for (int r0 = 0; r0 < b; r0 += 32) {
    for (int e = tid; e < 32 * r; e += 256) {
      const int rr = e % 32, a = e / 32, p = r0 + rr;
      Ub[a * 32 + rr] = (p < b) ? U[p + (int64_t)a * b] : 0.0;
    }
    __syncthreads();
    for (int e = tid; e < 32 * r; e += 256) {
      const int rr = e % 32, c = e / 32, p = r0 + rr;
      double s = 0.0;
      for (int a = 0; a < r; ++a) s = fma(Ub[a * 32 + rr], Wsh[a + c * r], s);
      if (p < b) Xi[p + (int64_t)c * b] = s;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------ diagonal update
// A_ii -= X_i U_ik^T on the lower 64 x 64 blocks of tile i = k + 1 + blockIdx.y; block 0 also applies
// the forward-solve update rhs_i -= U_ik (V_ik^T z_k).
__global__ __launch_bounds__(256) void k_diag_update(SlotArgs sa, int k) {
  const int slot = blockIdx.z;
  const SlotPtrs P = sa.slot(slot);
  if (failed(P.status)) return;
  const int b = sa.b;
  const int i = k + 1 + blockIdx.y;
  const int64_t tik = tri_index(i, k);
  const int r = P.rank[tik];
  if (r <= 0 || r > sa.rank_cap) {
    if (r != 0) atomicMax(&P.status[ST_GUARD], 1);
    return;
  }
  const double* U = P.U + tik * sa.fstride;
  const double* Xi = P.Xd + (int64_t)i * b * sa.rank_cap;
  double* A = P.D + (int64_t)i * b * b;
  const int tid = threadIdx.x;
  // columns staged 16 at a time (16 KB of LDS, so that several work groups share a CU)
  __shared__ double Xs[16][64];
  __shared__ double Us[16][64];
  __shared__ double zv[64];
  // 64 x 64 blocks (I >= J) of the lower triangle, blockIdx.x, blockIdx.x + gridDim.x, ...; thread (tr, tc)
  // owns rows pb + tr + 16 a and columns qb + 4 tc + q (a, q < 4): for each (a, q) a wave touches 4 columns
  // x 16 consecutive rows (128-byte runs).  The next block's 4 x 4 entries of A are loaded before the
  // current block's products, so that their latency overlaps them.
  const int tr = tid % 16, tc = tid / 16;
  const int nbt = (b + 63) / 64, nblk = nbt * (nbt + 1) / 2;
  auto block_of = [](int e, int* pb, int* qb) {
    int I = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
    while ((I + 1) * (I + 2) / 2 <= e) ++I;
    while (I * (I + 1) / 2 > e) --I;
    *pb = 64 * I;
    *qb = 64 * (e - I * (I + 1) / 2);
  };
  auto load_a = [&](int pb, int qb, double (&v)[4][4]) {
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int p = pb + tr + 16 * a;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int qq = qb + tc * 4 + q;
        v[a][q] = (p < b && qq < b && qq <= p) ? A[p + (int64_t)qq * b] : 0.0;
      }
    }
  };
  int pb = 0, qb = 0;
  double av[4][4];
  if ((int)blockIdx.x < nblk) {
    block_of(blockIdx.x, &pb, &qb);
    load_a(pb, qb, av);
  }
  for (int e = blockIdx.x; e < nblk; e += gridDim.x) {
    const int en = e + (int)gridDim.x;
    int pbn = 0, qbn = 0;
    double avn[4][4];
    if (en < nblk) {
      block_of(en, &pbn, &qbn);
      load_a(pbn, qbn, avn);
    }
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[a][c] = 0.0;
    for (int c0 = 0; c0 < r; c0 += 16) {
      const int cn = min(16, r - c0);
      if (c0 > 0) __syncthreads();
      for (int e2 = tid; e2 < 64 * cn; e2 += 256) {
        const int rr = e2 % 64, c = e2 / 64;
        Xs[c][rr] = (pb + rr < b) ? Xi[pb + rr + (int64_t)(c0 + c) * b] : 0.0;
        Us[c][rr] = (qb + rr < b) ? U[qb + rr + (int64_t)(c0 + c) * b] : 0.0;
      }
      __syncthreads();
      for (int c = 0; c < cn; ++c) {
        double xv[4], uv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) xv[a] = Xs[c][tr + 16 * a];
#pragma unroll
        for (int q = 0; q < 4; ++q) uv[q] = Us[c][tc * 4 + q];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int q = 0; q < 4; ++q) acc[a][q] = fma(xv[a], uv[q], acc[a][q]);
      }
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int p = pb + tr + 16 * a;
      if (p >= b) continue;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int qq = qb + tc * 4 + q;
        if (qq < b && qq <= p) A[p + (int64_t)qq * b] = av[a][q] - acc[a][q];
      }
    }
    __syncthreads();   // Xs and Us are restaged for the next block
    pb = pbn;
    qb = qbn;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int q = 0; q < 4; ++q) av[a][q] = avn[a][q];
  }
  if (blockIdx.x != 0) return;
  // forward-solve update of rhs_i
  const double* V = P.V + tik * sa.fstride;
  const double* z = P.rhs + (int64_t)k * b;
  const int bk = (int)min<int64_t>(b, sa.n - (int64_t)k * b);
  const int bi = (int)min<int64_t>(b, sa.n - (int64_t)i * b);
  const int w = wave_id(), l = lane_id();
  for (int c = w; c < r; c += 4) {
    double s = 0.0;
    for (int p = l; p < bk; p += 64) s = fma(V[p + (int64_t)c * b], z[p], s);
    s = wave_sum(s);
    if (l == 0) zv[c] = s;
  }
  __syncthreads();
  double* rhs = P.rhs + (int64_t)i * b;
  for (int p = tid; p < bi; p += 256) {
    double s = 0.0;
    for (int c = 0; c < r; ++c) s = fma(U[p + (int64_t)c * b], zv[c], s);
    rhs[p] -= s;
  }
}

// ------------------------------------------------------------------------------ recompression
// One work group per pair (i, j), k < j < i, with r_ik, r_jk > 0 (17:187-205).  The reference forms
// Unew = [U_ij, -U_ik W] and Vnew = [V_ij, U_jk] (W = V_ik^T V_jk), takes QR of both, the SVD of
// R_u R_v^T and truncates by the rule (17:119-133).  The same factorization is built here from the
// structure those outputs already have: V_ij has orthonormal columns and U_ij orthogonal columns with
// norms sigma_a, so with Q1 = U_ij diag(1/sigma) and X = -U_ik W,
//   Vnew = [V_ij, Q2v] [[I, Cv], [0, R2v]],  Cv = V_ij^T U_jk,      U_jk - V_ij Cv = Q2v R2v
//   Unew = [Q1,   Q2u] [[D, Cu], [0, R2u]],  Cu = Q1^T X,           X - Q1 Cu      = Q2u R2u
// (block Gram-Schmidt applied twice, Householder QR of the r_jk new columns only).  The core
// R_u R_v^T (w x w, w = r_ij + r_jk) goes to the one-sided Jacobi SVD; the outputs are formed row block
// by row block through LDS, in place.  The panel holds the r_jk new columns in NW waves x CPW column
// slots: r_jk <= 8 (nine pairs in ten) in 4 waves x 2 slots with a 36 KB LDS footprint, so that four
// work groups share a CU; r_jk <= 16 in 4 x 4 (two per CU); r_jk <= 32 in 8 x 4 and r_jk <= 64 in 8 x 8
// (one per CU, cores up to 64 x 64 in LDS).  The pairs of each class come from a work list.
template <int NW, int CPW>
struct RecompLayout {
  static constexpr int R2MAX = CPW * NW;
  static constexpr int WMAX = 64 + R2MAX;
  // cores up to WJ x WJ run in LDS (larger ones in the work group's scratch)
  // The 8-wave classes hold a CU's whole register file, so nothing else shares their CU: they take the
  // LDS as well (cores up to 64 x 64); the 4-wave classes keep to 36-56 KB (cores up to 32 x 32).
  static constexpr int WJ = NW >= 8 ? 64 : 32;
  static constexpr bool RR_LDS = R2MAX <= 32;
  static constexpr bool WC_LDS = true;
  // front: W and the CGS coefficients (rik x r2 and r1 x r2, 64 x R2MAX each), later the output staging
  // (RB x w, RB = 64 .. 8 rows as w allows)
  static constexpr int FRONT = 2 * 64 * R2MAX > 1024 ? 2 * 64 * R2MAX : 1024;
  static constexpr int RR = RR_LDS ? WMAX * R2MAX : 0;
  static constexpr int OFF_RV = FRONT, OFF_RU = OFF_RV + RR, OFF_SC = OFF_RU + RR, OFF_SG = OFF_SC + 64,
                       OFF_G = OFF_SG + 64, OFF_J = OFF_G + WJ * WJ, TOTAL = OFF_J + WJ * WJ;
  static constexpr int JTP = NW >= 8 ? 8 : 4;   // threads per Jacobi column pair (64 pairs: w <= 128)
};

__device__ __forceinline__ int recomp_class(int r2) { return r2 <= 8 ? 0 : (r2 <= 16 ? 1 : (r2 <= 32 ? 2 : 3)); }

// Append pair q of slot `slot` to list `level` of step k (panel: the pair updates the next panel column):
// level 0 the FP64 redo list, 1 the FP32 redo list, 2 the reduced-first list.
__device__ __forceinline__ void push_redo(const SlotArgs& sa, int k, int panel, int level, int q, int slot) {
  int* cnt = sa.wlr_count + WL_STRIDE * ((k * 2 + panel) * REDO_LEVELS + level);
  const int pos = atomicAdd(cnt, 1);
  if (pos < sa.wlr_cap) sa.wlr[(int64_t)(panel * REDO_LEVELS + level) * sa.wlr_cap + pos] = make_int2(q, slot);
  else atomicMax(&sa.slot(slot).status[ST_GUARD], 3);
}

// Work lists of step k: one thread per (pair, slot); pairs with r_ik, r_jk > 0 go to the list of their
// class (by r_jk, and by the core width r_ij + r_jk for the LDS), the heavy ones (wide cores: r_ij + r_jk above the class's threshold) from the front and the rest
// from the back, so that the persistent grid starts the long pairs first.  Ranks outside [0, rank_cap]
// trip the guard instead.  Pairs (i, k+1), which update the next panel column, go to separate panel lists
// so that they can run before the rest and let the next step's POTRF and TRSM start early.  The order
// inside a list depends on the atomics; every pair is computed independently of the others, so results
// do not.
__global__ __launch_bounds__(256) void k_recomp_list(SlotArgs sa, int k, int npairs) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  const int slot = blockIdx.y;
  if (q >= npairs) return;
  const SlotPtrs P = sa.slot(slot);
  if (failed(P.status)) return;
  int a_, b_;
  tri_decode((int64_t)q, &a_, &b_);   // pairs (a, b'), 0 <= b' < a; index 0 is (1, 0)
  const int i = a_ + k + 1, j = b_ + k + 1;
  const int r1 = P.rank[tri_index(i, j)], rik = P.rank[tri_index(i, k)], r2 = P.rank[tri_index(j, k)];
  if (r1 < 0 || r1 > sa.rank_cap || rik < 0 || rik > sa.rank_cap || r2 < 0 || r2 > sa.rank_cap) {
    atomicMax(&P.status[ST_GUARD], 1);
    return;
  }
  if (rik == 0 || r2 == 0) return;
  if (sa.prec_mode != PM_FP64) {
    // reduced precision: a pair goes to the reduced-first list only if it can pass the kernel's guard, judged
    // with an upper bound of the core's scale (tnorm_ij + tnorm_ik tnorm_jk >= max(||core||_F,
    // ||U_ik W||_F ||U_jk||_F)); the others run in FP64 on the class lists.  A replay follows its record.
    const int rep = sa.replay[slot];
    bool reduced;
    if (rep >= 0) {
      reduced = sa.rep_rpl[(int64_t)rep * sa.nlog + sa.rlog_off[k] + q] != PM_FP64;
    } else if (sa.rule_fixed) {
      // budget thr_rel ||core||_F against guard eps scale >= guard eps ||core||_F
      reduced = sa.thr[2 * slot] >= sa.guard * EPS_FP32;
    } else {
      const double scale = P.tnorm[tri_index(i, j)] + P.tnorm[tri_index(i, k)] * P.tnorm[tri_index(j, k)];
      reduced = sa.thr[2 * slot + 1] >= sa.guard * EPS_FP32 * (1.001 * scale);
    }
    if (reduced) {
      push_redo(sa, k, j == k + 1 ? 1 : 0, 2, q, slot);
      return;
    }
  }
  // the 4-wave classes keep cores up to 32 x 32 in LDS, the 8-wave ones up to 64 x 64: a wider core goes
  // to an 8-wave class, whose panel (32 columns) also holds its r_jk <= 16 new columns
  const int c = max(recomp_class(r2), r1 + r2 > 32 ? 2 : 0);
  const int heavy_w[RCLASSES] = {16, 32, 48, 80};
  const int key = (j == k + 1 ? 8 : 0) + 2 * c + (r1 + r2 > heavy_w[c] ? 1 : 0);
  // one counter update per wave and list (panel or not, class, heavy or light): the wave's entries of a list
  // take consecutive positions after its leader's atomic add
  const int lane = lane_id();
  for (int kk = 0; kk < 16; ++kk) {
    const unsigned long long mask = __ballot(key == kk);
    if (mask == 0ull) continue;
    const int leader = __ffsll((long long)mask) - 1;
    const bool panel = kk >= 8, heavy = kk & 1;
    const int cls = (kk >> 1) & 3;
    int* cnt = (panel ? sa.wlp_count : sa.wl_count) + WL_STRIDE * (RCLASSES * k + cls) + (heavy ? 0 : 1);
    int base = 0;
    if (lane == leader) base = atomicAdd(cnt, __popcll(mask));
    base = __shfl(base, leader);
    if (key == kk) {
      const int pos = base + __popcll(mask & ((1ull << lane) - 1ull));
      const int64_t cap = panel ? sa.wlp_cap : sa.wl_cap;
      int2* list = (panel ? sa.wlp : sa.wl) + (int64_t)cls * cap;
      list[heavy ? pos : cap - 1 - pos] = make_int2(q, slot);
    }
  }
}

// Block Gram-Schmidt of the panel against the r1 columns of A (global, or the LDS copy), applied twice:
// coefficients (scaled by scl when given) to R (leading dimension ldr), their sum over both passes.
template <int RPL, int NW, int CPW, class T, class S>
__device__ __forceinline__ void cgs2(RegPanel<RPL, NW, CPW, T>& pan, const S* A, int64_t lda, int m, int r1, int r2,
                                     const T* scl, T* R, int ldr, T* Ct) {
  panel_coef(pan, A, lda, m, r1, r2, scl, R, ldr);
  __syncthreads();
  panel_axpy(pan, A, lda, m, r1, r2, scl, R, ldr, (T)-1);
  panel_coef(pan, A, lda, m, r1, r2, scl, Ct, r1);
  __syncthreads();
  panel_axpy(pan, A, lda, m, r1, r2, scl, Ct, r1, (T)-1);
  for (int e = threadIdx.x; e < r1 * r2; e += 64 * NW) R[(e % r1) + (e / r1) * ldr] += Ct[e];
}

// One pair (i, j) = pair index q of step k, for slot `slot`; scr is the work group's scratch.  PM is the
// arithmetic (working type T: double for PM_FP64, float otherwise; PM_BF16 rounds the operands of the two
// products W = V_ik^T V_jk and U_ik W to bfloat16).  panel: the pair came from a panel list (j = k + 1).
template <int RPL, int NW, int CPW, int PM>
__device__ __forceinline__ void recompress_pair(const SlotArgs& sa, int k, int q, int slot, double* scr, int panel) {
  using L = RecompLayout<NW, CPW>;
  using T = typename std::conditional<PM == PM_FP64, double, float>::type;
  constexpr bool BF = PM == PM_BF16;
  const SlotPtrs P = sa.slot(slot);
  const int b = sa.b;
  int a_, b_;
  tri_decode((int64_t)q, &a_, &b_);
  const int i = a_ + k + 1, j = b_ + k + 1;
  const int64_t tij = tri_index(i, j), tik = tri_index(i, k), tjk = tri_index(j, k);
  const int r1 = P.rank[tij], rik = P.rank[tik], r2 = P.rank[tjk];
  const int w = r1 + r2;
  const double thr_rel = sa.thr[2 * slot], thr_abs = sa.thr[2 * slot + 1];
  const int64_t lo = sa.rlog_off[k] + q;
  const int rep = sa.replay[slot];
  if constexpr (PM != PM_FP64) {
    // a replay keeps the arithmetic its record used for this pair (frozen precisions)
    if (rep >= 0) {
      const int pr = sa.rep_rpl[(int64_t)rep * sa.nlog + lo];
      if (pr >= 0 && pr < PM) {
        if (threadIdx.x == 0) push_redo(sa, k, panel, pr == PM_FP32 ? 1 : 0, q, slot);
        return;
      }
    }
  }
  extern __shared__ double lds_raw[];
  T* lds = reinterpret_cast<T*>(lds_raw);
  T* scrT = reinterpret_cast<T*>(scr);
  // rik x r2 (phases 1-3) and the r1 x r2 temporary (phases 2-3): LDS, or scratch past the other regions
  T* Wm = L::WC_LDS ? lds : scrT + (int64_t)b * 64 + 2 * 128 * 64 + 2 * 128 * 128;
  T* Ct = L::WC_LDS ? lds + 64 * L::R2MAX : Wm + 64 * 64;
  T* Sb = lds;                                // RB x w staging     (phase 5, aliases the front)
  T* Q2v = scrT;                              // b x r2
  T* RvR = L::RR_LDS ? lds + L::OFF_RV : scrT + (int64_t)b * 64;            // w x r2: [Cv; R2v]
  T* RuR = L::RR_LDS ? lds + L::OFF_RU : scrT + (int64_t)b * 64 + 128 * 64; // w x r2: [Cu; R2u]
  T* sc = lds + L::OFF_SC;                    // 1 / sigma_a of U_ij
  T* sg = lds + L::OFF_SG;                    // sigma_a
  const bool core_lds = w <= L::WJ;
  T* G = core_lds ? lds + L::OFF_G : scrT + (int64_t)b * 64 + 2 * 128 * 64;
  T* J = core_lds ? lds + L::OFF_J : G + 128 * 128;
  // Householder vectors of the panel QRs (phases 2-3 only): the core's LDS region, unused until phase 4
  static_assert(2 * 64 * RPL <= L::WJ * L::WJ, "reflector buffer must fit the core region");
  T* vbuf = lds + L::OFF_G;
  __shared__ T tau[8 * NW];
  __shared__ T sig[L::WMAX];
  __shared__ int perm[L::WMAX];
  __shared__ int flag, s_rank;
  __shared__ double s_tail;
  const int tid = threadIdx.x, NT = 64 * NW, wv = wave_id(), ln = lane_id();
  const int64_t fs = sa.fstride;
  double* Uij = P.U + tij * fs;
  double* Vij = P.V + tij * fs;
  const double* Uik = P.U + tik * fs;
  const double* Vik = P.V + tik * fs;
  const double* Ujk = P.U + tjk * fs;
  const double* Vjk = P.V + tjk * fs;
  RegPanel<RPL, NW, CPW, T> pan;
  // reduced precision: this thread's share of ||U_jk||_F^2 and ||U_ik W||_F^2 (the guard's update scale)
  T nu2 = (T)0, nx2 = (T)0;

  // ---- 1. W = V_ik^T V_jk
  pan.template load<BF>(Vjk, b, b, r2);
  panel_coef<BF>(pan, Vik, b, b, rik, r2, (const T*)nullptr, Wm, rik);
  // ---- 2. V side: [V_ij, U_jk] = [V_ij, Q2v] [[I, Cv], [0, R2v]]
  pan.load(Ujk, b, b, r2);
  if constexpr (PM != PM_FP64) {
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj)
#pragma unroll
      for (int t = 0; t < RPL; ++t) nu2 = fma(pan.a[jj][t], pan.a[jj][t], nu2);
  }
  if (r1 > 0) cgs2(pan, Vij, b, b, r1, r2, (const T*)nullptr, RvR, w, Ct);
  pan.geqr2(b, r2, vbuf, tau);
  pan.store_r(RvR + r1, w, r2);            // R2v in rows r1.. of the right block
  pan.org2r(b, r2, vbuf, tau);
  pan.store(Q2v, b, b, r2);
  // ---- 3. U side: [U_ij, -U_ik W] = [Q1, Q2u] [[D, Cu], [0, R2u]]
  for (int a = wv; a < r1; a += NW) {
    T d = (T)0;
    for (int p = ln; p < b; p += 64) { const T u = (T)Uij[p + (int64_t)a * b]; d = fma(u, u, d); }
    d = wave_sum(d);
    if (ln == 0) { const T sgm = tsqrt(d); sg[a] = sgm; sc[a] = sgm > (T)0 ? (T)1 / sgm : (T)0; }
  }
  __syncthreads();
  pan.zero();
  panel_axpy<BF>(pan, Uik, b, b, rik, r2, (const T*)nullptr, Wm, rik, (T)-1);
  if constexpr (PM != PM_FP64) {
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj)
#pragma unroll
      for (int t = 0; t < RPL; ++t) nx2 = fma(pan.a[jj][t], pan.a[jj][t], nx2);
  }
  if (r1 > 0) cgs2(pan, Uij, b, b, r1, r2, (const T*)sc, RuR, w, Ct);
  pan.geqr2(b, r2, vbuf, tau);
  pan.store_r(RuR + r1, w, r2);
  pan.org2r(b, r2, vbuf, tau);   // Q2u in registers from here to the outputs (ends with a barrier)
  // narrow cores in the 4-wave classes form their outputs row by row (phase 5), reading Q2u back from the
  // work group's scratch past Q2v (the region the 8-wave classes use for R_u, R_v; these classes keep
  // those in LDS)
  constexpr int XO = 16;
  const bool rowout = NW == 4 && w <= XO;
  T* Q2u = scrT + (int64_t)b * 64;
  if (rowout) pan.store(Q2u, b, b, r2);
  // ---- 4. core K = R_u R_v^T, stored transposed: the one-sided Jacobi runs on K^T (on these cores it
  // needs fewer sweeps than on K), giving K^T J = V S (columns orthogonal, norms S) and J = U; truncation
  T nk2 = (T)0;
  for (int e = tid; e < w * w; e += NT) {
    const int a = e % w, c = e / w;
    T s = (a == c && a < r1) ? sg[a] : (T)0;
    for (int m = 0; m < r2; ++m) s = fma(RuR[a + m * w], RvR[c + m * w], s);
    G[c + a * w] = s;
    if constexpr (PM != PM_FP64) nk2 = fma(s, s, nk2);
  }
  if constexpr (PM != PM_FP64) {
    // the guard: ||core||_F and the update scale ||U_ik W||_F ||U_jk||_F, summed over the work group in
    // double
    __shared__ double gsum[3 * NW];
    __shared__ int s_go;
    const double a0 = wave_sum((double)nk2), a1 = wave_sum((double)nx2), a2 = wave_sum((double)nu2);
    if (ln == 0) { gsum[wv] = a0; gsum[NW + wv] = a1; gsum[2 * NW + wv] = a2; }
    __syncthreads();
    if (tid == 0) {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0;
      for (int u = 0; u < NW; ++u) { s0 += gsum[u]; s1 += gsum[NW + u]; s2 += gsum[2 * NW + u]; }
      const double nk = sqrt(s0), scale = fmax(nk, sqrt(s1) * sqrt(s2));
      const double thr = sa.rule_fixed ? thr_rel * nk : thr_abs;
      int go = PM;   // arithmetic this pair gets: PM, or a redo level
      if (!(nk == nk && scale <= 1.7976931348623157e308)) go = PM_FP64;   // non-finite: FP64 redo
      else if (rep < 0) {
        if (PM == PM_BF16 && !(thr >= sa.guard * EPS_BF16 * scale)) go = PM_FP32;
        if (go == PM_FP32 && !(thr >= sa.guard * EPS_FP32 * scale)) go = PM_FP64;
      }
      if (go != PM) push_redo(sa, k, panel, go == PM_FP32 ? 1 : 0, q, slot);
      s_go = go;
    }
    __syncthreads();
    if (s_go != PM) return;
  } else {
    __syncthreads();
  }
  const int sweeps = jacobi_svd_auto<L::JTP>(G, w, J, w, w, sig, perm, &flag, 60);
  if (sweeps == -2) {   // non-finite core
    if constexpr (PM == PM_FP64) {   // report, keep the tile unchanged
      if (tid == 0) atomicMax(&P.status[ST_GUARD], 2);
    } else {                         // redo in FP64
      if (tid == 0) push_redo(sa, k, panel, 0, q, slot);
    }
    return;
  }
  if (tid == 0) {
    double tot = 0.0;
    for (int c = 0; c < w; ++c) tot += (double)sig[c] * (double)sig[c];
    const double thr = sa.rule_fixed ? thr_rel * sqrt(tot) : thr_abs;
    double acc = 0.0;
    int r = w;
    for (int c = w - 1; c >= 0; --c) {
      const double s = (double)sig[perm[c]];
      const double nacc = acc + s * s;
      if (sqrt(nacc) <= thr) { acc = nacc; r = c; } else break;
    }
    if (rep >= 0) r = min(sa.rep_rlog[(int64_t)rep * sa.nlog + lo], w);   // frozen ranks
    if (r > sa.rank_cap) { atomicMax(&P.status[ST_RANKCAP], 1); r = sa.rank_cap; }
    double kept = 0.0;   // ||U_ij||_F of the new factors (V_ij orthonormal)
    for (int c = 0; c < r; ++c) kept += (double)sig[perm[c]] * (double)sig[perm[c]];
    P.tnorm[tij] = sqrt(kept);
    P.rlog[lo] = r;
    P.rpl[lo] = PM;
    s_rank = r;
    s_tail = sqrt(acc);
  }
  __syncthreads();
  const int rn = s_rank;
  // ---- 5. outputs
  // kept column c: U-side coefficients J[:, c] s_c, V-side (K^T J)[:, c] / s_c (orthonormal), scaled as used.
  // Narrow cores (rowout): each thread forms whole rows, reading the row's w inputs once (every load of a
  // row in flight together) and writing its rn outputs in place, with no barrier between row blocks; the
  // sums run in the same order as in the staged form below, so the values are the same.
  if (rowout) {
    // rows p = tid + NT h (h < NR): the inputs of all of a thread's rows are loaded before the first sum
    // (one round of loads per side)
    constexpr int NR = (64 * RPL + NT - 1) / NT;
    T x[NR][XO];
    // V_ij <- [V_ij, Q2v] V_core[:, perm]
#pragma unroll
    for (int h = 0; h < NR; ++h) {
      const int p = tid + NT * h;
#pragma unroll
      for (int m = 0; m < XO; ++m)
        x[h][m] = !(p < b && m < w) ? (T)0 : (m < r1 ? (T)Vij[p + (int64_t)m * b] : Q2v[p + (int64_t)(m - r1) * b]);
    }
    for (int c = 0; c < rn; ++c) {
      const int pc = perm[c];
      const T* Vc = G + pc * w;
      const T sv = sig[pc];
#pragma unroll
      for (int h = 0; h < NR; ++h) {
        const int p = tid + NT * h;
        T s = (T)0;
#pragma unroll
        for (int m = 0; m < XO; ++m)
          if (m < w) s = fma(x[h][m], Vc[m], s);
        if (p < b) Vij[p + (int64_t)c * b] = sv > (T)0 ? (double)(s / sv) : 0.0;
      }
    }
    // U_ij <- [Q1, Q2u] (U S)_core[:, perm],  Q1 = U_ij diag(1/sigma)
#pragma unroll
    for (int h = 0; h < NR; ++h) {
      const int p = tid + NT * h;
#pragma unroll
      for (int m = 0; m < XO; ++m)
        x[h][m] = !(p < b && m < w) ? (T)0
                                    : (m < r1 ? (T)Uij[p + (int64_t)m * b] * sc[m] : Q2u[p + (int64_t)(m - r1) * b]);
    }
    for (int c = 0; c < rn; ++c) {
      const int pc = perm[c];
      const T* Uc = J + pc * w;
      const T sv = sig[pc];
#pragma unroll
      for (int h = 0; h < NR; ++h) {
        const int p = tid + NT * h;
        T s = (T)0;
#pragma unroll
        for (int m = 0; m < XO; ++m)
          if (m < w) s = fma(x[h][m], Uc[m], s);
        if (p < b) Uij[p + (int64_t)c * b] = (double)(s * sv);
      }
    }
    __syncthreads();
  }
  // wider cores: row blocks of RB rows staged through LDS (in place); register row block t of the panel
  // covers rows [64 t, 64 t + 64), so t stays a compile-time index
  const int rbs = 64 * w <= L::FRONT ? 6 : (32 * w <= L::FRONT ? 5 : (16 * w <= L::FRONT ? 4 : 3));
  const int RB = 1 << rbs;
#pragma unroll
  for (int t = 0; t < RPL; ++t) {
    if (rowout || 64 * t >= b) break;
    const int rend = min(64 * t + 64, b);
    for (int r0 = 64 * t; r0 < rend; r0 += RB) {
      // V_ij <- [V_ij, Q2v] V_core[:, perm]
      for (int e = tid; e < RB * w; e += NT) {
        const int rr = e & (RB - 1), m = e >> rbs, p = r0 + rr;
        Sb[m * RB + rr] = (p < b) ? (m < r1 ? (T)Vij[p + (int64_t)m * b] : Q2v[p + (int64_t)(m - r1) * b]) : (T)0;
      }
      __syncthreads();
      for (int e = tid; e < RB * rn; e += NT) {
        const int rr = e & (RB - 1), c = e >> rbs, p = r0 + rr;
        const T* Vc = G + perm[c] * w;
        T s = (T)0;
        for (int m = 0; m < w; ++m) s = fma(Sb[m * RB + rr], Vc[m], s);
        const T sv = sig[perm[c]];
        if (p < b) Vij[p + (int64_t)c * b] = sv > (T)0 ? (double)(s / sv) : 0.0;
      }
      __syncthreads();
      // U_ij <- [Q1, Q2u] (U S)_core[:, perm],  Q1 = U_ij diag(1/sigma)
      for (int e = tid; e < RB * r1; e += NT) {
        const int rr = e & (RB - 1), m = e >> rbs, p = r0 + rr;
        Sb[m * RB + rr] = (p < b) ? (T)Uij[p + (int64_t)m * b] * sc[m] : (T)0;
      }
      const int rr_own = ln - (r0 - 64 * t);
#pragma unroll
      for (int jj = 0; jj < CPW; ++jj) {
        const int c = RegPanel<RPL, NW, CPW, T>::col_of(wv, jj);
        if (c < r2 && rr_own >= 0 && rr_own < RB) Sb[(r1 + c) * RB + rr_own] = pan.a[jj][t];
      }
      __syncthreads();
      for (int e = tid; e < RB * rn; e += NT) {
        const int rr = e & (RB - 1), c = e >> rbs, p = r0 + rr;
        const T* Uc = J + perm[c] * w;
        T s = (T)0;
        for (int m = 0; m < w; ++m) s = fma(Sb[m * RB + rr], Uc[m], s);
        if (p < b) Uij[p + (int64_t)c * b] = (double)(s * sig[perm[c]]);
      }
      __syncthreads();
    }
  }
  if constexpr (PM == PM_FP64) {
    if (sa.mixed) {
      // mixed rule: storage precision under the budget left after truncation (17:203-204); the work
      // group's scratch (Q2v is dead by now) holds the Gram matrices, sig[] the reductions
      const int bi = (int)min<int64_t>(b, sa.n - (int64_t)i * b), bj = (int)min<int64_t>(b, sa.n - (int64_t)j * b);
      round_pair_wg(Uij, b, bi, Vij, b, bj, rn, thr_abs - s_tail, scr, sig);
    }
  }
  if (tid == 0) P.rank[tij] = rn;
}

// Persistent grid over the class-cls work list of step k (grid = resident work groups of the class).
// The 4-wave classes ask for 2 waves per SIMD (two work groups per CU): their code needs about 200-240
// VGPRs, and a tighter target only adds scratch spills without reaching a third work group.
template <int NW, int CPW>
struct RecompWaves {
  static constexpr int value = NW * CPW <= 16 ? 2 : 1;
};

// Entries are claimed one at a time from the list's claim counter, in list order (heavy pairs first), so a
// work group that becomes resident late (its CU still held by another kernel) only takes what is left.
// src 0: the class-cls list of step k (panel or the rest); src 1 / 2: the FP64 / FP32 redo list of step k
// (pairs that a reduced-precision launch handed on; complete before this launch starts); src 3: the
// reduced-first list of step k (complete once the list kernel is done).
template <int RPL, int NW, int CPW, int PM>
__global__ __launch_bounds__(64 * NW, (RecompWaves<NW, CPW>::value)) void k_recompress(SlotArgs sa, int k, int cls,
                                                                                        int panel, int src) {
  int* cnts;
  const int2* list;
  int64_t cap;
  int nh, cnt;
  if (src == 0) {
    cnts = (panel ? sa.wlp_count : sa.wl_count) + WL_STRIDE * (RCLASSES * k + cls);
    cap = panel ? sa.wlp_cap : sa.wl_cap;
    nh = __builtin_amdgcn_readfirstlane(cnts[0]);
    cnt = nh + __builtin_amdgcn_readfirstlane(cnts[1]);
    list = (panel ? sa.wlp : sa.wl) + (int64_t)cls * cap;
  } else {
    const int lvl = src - 1;   // src 1: FP64 redo, 2: FP32 redo, 3: reduced-first
    cnts = sa.wlr_count + WL_STRIDE * ((k * 2 + panel) * REDO_LEVELS + lvl);
    cap = sa.wlr_cap;
    nh = (int)min<int64_t>(__builtin_amdgcn_readfirstlane(cnts[0]), cap);
    cnt = nh;
    list = sa.wlr + (int64_t)(panel * REDO_LEVELS + lvl) * cap;
  }
  double* scr = sa.scr_pool + (int64_t)blockIdx.x * sa.scr_stride;
  __shared__ int s_claim;
  for (;;) {
    if (threadIdx.x == 0) s_claim = atomicAdd(&cnts[2], 1);
    __syncthreads();
    const int q = __builtin_amdgcn_readfirstlane(s_claim);
    __syncthreads();
    if (q >= cnt) break;
    const int2 e = list[q < nh ? q : cap - 1 - (q - nh)];
    recompress_pair<RPL, NW, CPW, PM>(sa, k, __builtin_amdgcn_readfirstlane(e.x), __builtin_amdgcn_readfirstlane(e.y),
                                      scr, panel);
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------ final
// logdet = 2 sum_k logdet_part[k]; quad = z^T z; ll = -(quad + logdet + N log 2 pi) / 2.
// Fixed-order reductions (one work group per slot).
__global__ __launch_bounds__(256) void k_final(SlotArgs sa) {
  const int slot = blockIdx.x;
  const SlotPtrs P = sa.slot(slot);
  __shared__ double red[256];
  const int tid = threadIdx.x;
  double s = 0.0;
  for (int64_t p = tid; p < sa.n; p += 256) s = fma(P.rhs[p], P.rhs[p], s);
  red[tid] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) red[tid] += red[tid + h];
    __syncthreads();
  }
  if (tid == 0) {
    const double quad = red[0];
    double ld = 0.0;
    for (int k = 0; k < sa.nB; ++k) ld += P.logdet_part[k];
    ld *= 2.0;
    P.result[0] = -0.5 * (quad + ld + (double)sa.n * 1.8378770664093453);
    P.result[1] = ld;
    P.result[2] = quad;
  }
}

// ------------------------------------------------------------------------------------ launchers
hipError_t launch_potrf(const SlotArgs& sa, int k, int nslot, hipStream_t s) {
  hipLaunchKernelGGL(k_potrf, dim3(1, 1, nslot), dim3(512), 0, s, sa, k);
  return hipGetLastError();
}

hipError_t launch_trtri(const SlotArgs& sa, int k, int nslot, hipStream_t s) {
  hipLaunchKernelGGL(k_trtri, dim3(sa.npan, 1, nslot), dim3(256), 0, s, sa, k);
  return hipGetLastError();
}

hipError_t launch_trsm(const SlotArgs& sa, int k, int nslot, hipStream_t s) {
  hipLaunchKernelGGL(k_trsm, dim3(sa.nB - k, 1, nslot), dim3(256), 0, s, sa, k);
  return hipGetLastError();
}

hipError_t launch_diag_update(const SlotArgs& sa, int k, int nslot, hipStream_t s) {
  const int rows = sa.nB - k - 1;
  if (rows <= 0) return hipSuccess;
  // work groups per tile: each takes every diag_wg-th 64 x 64 block of the tile's lower triangle
  const int nb64 = (sa.b + 63) / 64;
  const int g = max(1, min(sa.diag_wg, nb64 * (nb64 + 1) / 2));
  hipLaunchKernelGGL(k_diag_update, dim3(g, rows, nslot), dim3(256), 0, s, sa, k);
  return hipGetLastError();
}

template <int PM>
using RecompT = typename std::conditional<PM == PM_FP64, double, float>::type;

// dynamic LDS of a recompression launch
template <int RPL, int NW, int CPW, int PM>
constexpr size_t recomp_lds_bytes() {
  return (size_t)RecompLayout<NW, CPW>::TOTAL * sizeof(RecompT<PM>);
}

template <int RPL, int NW, int CPW, int PM>
static hipError_t recomp_attr() {
  static bool attr_set = false;
  hipError_t e = hipSuccess;
  if (!attr_set) {
    e = hipFuncSetAttribute((const void*)k_recompress<RPL, NW, CPW, PM>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)recomp_lds_bytes<RPL, NW, CPW, PM>());
    attr_set = e == hipSuccess;
  }
  return e;
}

template <int RPL, int NW, int CPW, int PM>
static hipError_t launch_recomp_t(const SlotArgs& sa, int k, int cls, int panel, int grid, int src, hipStream_t s) {
  hipError_t e = recomp_attr<RPL, NW, CPW, PM>();
  if (e != hipSuccess) return e;
  const size_t lds = recomp_lds_bytes<RPL, NW, CPW, PM>();
  hipLaunchKernelGGL((k_recompress<RPL, NW, CPW, PM>), dim3(grid), dim3(64 * NW), lds, s, sa, k, cls, panel, src);
  return hipGetLastError();
}

template <int RPL, int NW, int CPW, int PM>
static hipError_t occ_t(int* out) {
  hipError_t e = recomp_attr<RPL, NW, CPW, PM>();
  if (e != hipSuccess) return e;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, (const void*)k_recompress<RPL, NW, CPW, PM>, 64 * NW,
                                                      recomp_lds_bytes<RPL, NW, CPW, PM>());
}

template <int PM>
static hipError_t occupancy_pm(int b, int blocks_per_cu[RCLASSES]) {
  hipError_t e;
  if (b <= 256) {
    if ((e = occ_t<4, 4, 2, PM>(blocks_per_cu + 0)) != hipSuccess) return e;
    if ((e = occ_t<4, 4, 4, PM>(blocks_per_cu + 1)) != hipSuccess) return e;
    if ((e = occ_t<4, 8, 4, PM>(blocks_per_cu + 2)) != hipSuccess) return e;
    return occ_t<4, 8, 8, PM>(blocks_per_cu + 3);
  }
  if ((e = occ_t<8, 4, 2, PM>(blocks_per_cu + 0)) != hipSuccess) return e;
  if ((e = occ_t<8, 4, 4, PM>(blocks_per_cu + 1)) != hipSuccess) return e;
  if ((e = occ_t<8, 8, 4, PM>(blocks_per_cu + 2)) != hipSuccess) return e;
  return occ_t<8, 8, 8, PM>(blocks_per_cu + 3);
}

hipError_t recompress_occupancy(int b, int pm, int blocks_per_cu[RCLASSES]) {
  if (pm == PM_FP32) return occupancy_pm<PM_FP32>(b, blocks_per_cu);
  if (pm == PM_BF16) return occupancy_pm<PM_BF16>(b, blocks_per_cu);
  return occupancy_pm<PM_FP64>(b, blocks_per_cu);
}

// Work lists of step k (one thread per pair and slot); false when step k has no pairs.
hipError_t launch_recomp_list(const SlotArgs& sa, int k, int nslot, hipStream_t s) {
  const int m = sa.nB - k - 1;            // tiles below the panel
  const int npairs = m * (m - 1) / 2;     // pairs (i, j), k < j < i
  if (npairs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_recomp_list, dim3((npairs + 255) / 256, nslot), dim3(256), 0, s, sa, k, npairs);
  return hipGetLastError();
}

template <int PM>
static hipError_t launch_class_pm(const SlotArgs& sa, int k, int cls, int panel, int grid, int src, hipStream_t s) {
  if (sa.b <= 256) {
    switch (cls) {
      case 0: return launch_recomp_t<4, 4, 2, PM>(sa, k, 0, panel, grid, src, s);
      case 1: return launch_recomp_t<4, 4, 4, PM>(sa, k, 1, panel, grid, src, s);
      case 2: return launch_recomp_t<4, 8, 4, PM>(sa, k, 2, panel, grid, src, s);
      default: return launch_recomp_t<4, 8, 8, PM>(sa, k, 3, panel, grid, src, s);
    }
  }
  switch (cls) {
    case 0: return launch_recomp_t<8, 4, 2, PM>(sa, k, 0, panel, grid, src, s);
    case 1: return launch_recomp_t<8, 4, 4, PM>(sa, k, 1, panel, grid, src, s);
    case 2: return launch_recomp_t<8, 8, 4, PM>(sa, k, 2, panel, grid, src, s);
    default: return launch_recomp_t<8, 8, 8, PM>(sa, k, 3, panel, grid, src, s);
  }
}

// One persistent launch of class cls for step k (grid work groups, each looping over the class list, the
// panel list when panel != 0, or a redo / reduced-first list (src)); scr is the launch's scratch pool, grid x
// scr_stride doubles.  Those lists may hold pairs of any class: they run on the widest class (cls 3).
hipError_t launch_recomp_class(const SlotArgs& sa_in, int k, int cls, int panel, int grid, double* scr, hipStream_t s,
                               int pm, int src) {
  SlotArgs sa = sa_in;
  sa.scr_pool = scr;
  if (src != 0) cls = 3;
  if (pm == PM_FP32) return launch_class_pm<PM_FP32>(sa, k, cls, panel, grid, src, s);
  if (pm == PM_BF16) return launch_class_pm<PM_BF16>(sa, k, cls, panel, grid, src, s);
  return launch_class_pm<PM_FP64>(sa, k, cls, panel, grid, src, s);
}

hipError_t launch_final(const SlotArgs& sa, int nslot, hipStream_t s) {
  hipLaunchKernelGGL(k_final, dim3(nslot), dim3(256), 0, s, sa);
  return hipGetLastError();
}

}  // namespace tlr
