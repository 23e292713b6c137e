// Tile solves with a completed TLR factor and kriging prediction.
//
// Reference semantics (codes/17_tlr_cholesky_pilot.py):
//   tlr_forward  17:230-240  Z = L^{-1} B:  Z_k = L_kk^{-1} (B_k - sum_{j<k} U_kj (V_kj^T Z_j))
//   tlr_backward 17:243-256  X = L^{-T} Z:  X_k = L_kk^{-T} (Z_k - sum_{i>k} V_ik (U_ik^T X_i))
//   tlr_predict  17:269-276  mu = Kp L^{-T} L^{-1} y, var = sigma2 (1 + eta) - colsum((L^{-1} Kp^T)^2)
// Both solves are run right-looking here (a tile's contribution is subtracted as soon as it is known),
// which uses the same factor blocks as the reference's left-looking loops.  The right-hand sides form an
// N x m column-major block (leading dimension N); every kernel works on 32-column groups.
#include "common.hpp"
#include "tlr_kernels.hpp"

namespace tlr {

constexpr int SB = 32;   // panel rows of the diagonal solves (the factor's stored inverse blocks)
constexpr int SC = 32;   // right-hand-side columns per work group

// X_k <- L_kk^{-1} X_k (trans = 0) or L_kk^{-T} X_k (trans = 1), columns [32 blockIdx.x, +32).
__global__ __launch_bounds__(256) void k_solve_diag(SlotArgs sa, int slot, int k, double* __restrict__ X,
                                                    int64_t ldx, int m, int trans) {
  const SlotPtrs P = sa.slot(slot);
  const int b = sa.b;
  const int bk = (int)min<int64_t>(b, sa.n - (int64_t)k * b);
  const double* L = P.D + (int64_t)k * b * b;
  const double* Dv = P.Dinv + (int64_t)k * sa.npan * SB * SB;
  double* Xk = X + (int64_t)k * b;
  const int c0 = blockIdx.x * SC;
  const int nc = min(SC, m - c0);
  __shared__ double Lb[SB][SB + 1];
  __shared__ double Xb[SB][SC + 1];
  __shared__ double tmp[SB][SC + 1];
  __shared__ double Dl[SB][SB + 1];
  const int tid = threadIdx.x;
  const int ll = tid % SB, cg = tid / SB;   // 8 column groups of 4 columns
  const int npan = (bk + SB - 1) / SB;
  for (int step = 0; step < npan; ++step) {
    const int P0 = trans ? npan - 1 - step : step;
    const int p0 = P0 * SB;
    const int wd = min(SB, bk - p0);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    // forward: rows p0.. minus L[p0.., 0:p0] X[0:p0]; backward: minus L[p0+wd.., p0..]^T X[p0+wd..]
    const int q_begin = trans ? p0 + wd : 0;
    const int q_end = trans ? bk : p0;
    for (int q0 = q_begin; q0 < q_end; q0 += SB) {
      const int qw = min(SB, q_end - q0);
      for (int e = tid; e < SB * SB; e += 256) {
        const int a = e % SB, c = e / SB;
        double v = 0.0;
        if (!trans) { if (a < wd && c < qw) v = L[(p0 + a) + (int64_t)(q0 + c) * b]; }   // Lb[a][c] = L[p0+a][q0+c]
        else { if (a < qw && c < wd) v = L[(q0 + a) + (int64_t)(p0 + c) * b]; }          // Lb[a][c] = L[q0+a][p0+c]
        Lb[a][c] = v;
      }
      for (int e = tid; e < SB * SC; e += 256) {
        const int qq = e % SB, c = e / SB;
        Xb[qq][c] = (qq < qw && c < nc) ? Xk[(q0 + qq) + (int64_t)(c0 + c) * ldx] : 0.0;
      }
      __syncthreads();
#pragma unroll 4
      for (int qq = 0; qq < SB; ++qq) {
        const double lv = trans ? Lb[qq][ll] : Lb[ll][qq];
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] = fma(lv, Xb[qq][cg + 8 * u], acc[u]);
      }
      __syncthreads();
    }
    for (int e = tid; e < SB * SB; e += 256) Dl[e % SB][e / SB] = Dv[P0 * SB * SB + e];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int c = cg + 8 * u;
      tmp[ll][c] = (ll < wd && c < nc) ? Xk[(p0 + ll) + (int64_t)(c0 + c) * ldx] - acc[u] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int c = cg + 8 * u;
      if (ll < wd && c < nc) {
        double s = 0.0;
        if (!trans) { for (int q = 0; q <= ll; ++q) s = fma(Dl[ll][q], tmp[q][c], s); }
        else { for (int q = ll; q < wd; ++q) s = fma(Dl[q][ll], tmp[q][c], s); }
        Xk[(p0 + ll) + (int64_t)(c0 + c) * ldx] = s;
      }
    }
    __syncthreads();
  }
}

// Low-rank updates after X_k is known.  dir = 0 (forward): X_i -= U_ik (V_ik^T X_k) for i > k, tile
// i = k + 1 + blockIdx.y.  dir = 1 (backward): X_j -= V_kj (U_kj^T X_k) for j < k, tile j = blockIdx.y.
__global__ __launch_bounds__(256) void k_solve_update(SlotArgs sa, int slot, int k, double* __restrict__ X,
                                                      int64_t ldx, int m, int dir) {
  const SlotPtrs P = sa.slot(slot);
  const int b = sa.b;
  const int other = dir == 0 ? k + 1 + blockIdx.y : blockIdx.y;
  const int64_t t = dir == 0 ? tri_index(other, k) : tri_index(k, other);
  const int r = P.rank[t];
  if (r <= 0 || r > sa.rank_cap) return;
  // the factor block L_ik = U V^T: rows of U are the sites of the larger tile index
  const double* U = P.U + t * sa.fstride;
  const double* V = P.V + t * sa.fstride;
  const double* A = dir == 0 ? V : U;   // projected against X_k
  const double* Bm = dir == 0 ? U : V;  // expanded into X_other
  const int nk = (int)min<int64_t>(b, sa.n - (int64_t)k * b);
  const int no = (int)min<int64_t>(b, sa.n - (int64_t)other * b);
  const double* Xk = X + (int64_t)k * b;
  double* Xo = X + (int64_t)other * b;
  const int c0 = blockIdx.x * SC;
  const int nc = min(SC, m - c0);
  __shared__ double C[64][SC + 1];
  __shared__ double St[64][SC + 1];   // staged rows of X_k / of A
  const int tid = threadIdx.x;
  // C = A^T X_k (r x nc): thread owns (a, c) pairs; rows staged 64 at a time
  for (int e = tid; e < 64 * SC; e += 256) C[e / SC][e % SC] = 0.0;
  __shared__ double Ab[64][64 + 1];
  for (int r0 = 0; r0 < nk; r0 += 64) {
    const int rw = min(64, nk - r0);
    for (int e = tid; e < 64 * SC; e += 256) {
      const int rr = e % 64, c = e / 64;
      St[rr][c] = (rr < rw && c < nc) ? Xk[(r0 + rr) + (int64_t)(c0 + c) * ldx] : 0.0;
    }
    for (int e = tid; e < 64 * r; e += 256) {
      const int rr = e % 64, a = e / 64;
      Ab[rr][a] = (rr < rw) ? A[(r0 + rr) + (int64_t)a * b] : 0.0;
    }
    __syncthreads();
    for (int e = tid; e < r * SC; e += 256) {
      const int a = e / SC, c = e % SC;
      double s = 0.0;
      for (int rr = 0; rr < rw; ++rr) s = fma(Ab[rr][a], St[rr][c], s);
      C[a][c] += s;
    }
    __syncthreads();
  }
  // X_other -= Bm C
  for (int r0 = 0; r0 < no; r0 += 64) {
    const int rw = min(64, no - r0);
    for (int e = tid; e < 64 * r; e += 256) {
      const int rr = e % 64, a = e / 64;
      Ab[rr][a] = (rr < rw) ? Bm[(r0 + rr) + (int64_t)a * b] : 0.0;
    }
    __syncthreads();
    for (int e = tid; e < 64 * SC; e += 256) {
      const int rr = e % 64, c = e / 64;
      if (rr < rw && c < nc) {
        double s = 0.0;
        for (int a = 0; a < r; ++a) s = fma(Ab[rr][a], C[a][c], s);
        Xo[(r0 + rr) + (int64_t)(c0 + c) * ldx] -= s;
      }
    }
    __syncthreads();
  }
}

// Xc[i + j ldx] = sigma2 matern32(S_i, Sp_j): the transposed cross covariance Kp^T (N x m).
__global__ void k_gen_cross(const double* __restrict__ sx, const double* __restrict__ sy, int64_t n,
                            const double* __restrict__ px, const double* __restrict__ py, int m, MaternParams mp,
                            double* __restrict__ Xc, int64_t ldx) {
  const int j = blockIdx.y;
  const double qx = px[j], qy = py[j];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    Xc[i + (int64_t)j * ldx] = matern32(mp, sx[i], sy[i], qx, qy);
}

// out[j] = sum_i X[i + j ldx] * (w ? w[i] : X[i + j ldx]) over i < n: one work group per column, fixed order.
__global__ __launch_bounds__(256) void k_col_dot(const double* __restrict__ X, int64_t ldx, int64_t n,
                                                 const double* __restrict__ w, double* __restrict__ out) {
  const int j = blockIdx.x;
  const double* x = X + (int64_t)j * ldx;
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) s = fma(x[i], w ? w[i] : x[i], s);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = red[0];
}

// ------------------------------------------------------------------------------------ launchers
hipError_t launch_solve(const SlotArgs& sa, int slot, double* X, int64_t ldx, int m, int trans, hipStream_t s) {
  const int ncb = (m + SC - 1) / SC;
  if (!trans) {
    for (int k = 0; k < sa.nB; ++k) {
      hipLaunchKernelGGL(k_solve_diag, dim3(ncb), dim3(256), 0, s, sa, slot, k, X, ldx, m, 0);
      if (k + 1 < sa.nB)
        hipLaunchKernelGGL(k_solve_update, dim3(ncb, sa.nB - k - 1), dim3(256), 0, s, sa, slot, k, X, ldx, m, 0);
    }
  } else {
    for (int k = sa.nB - 1; k >= 0; --k) {
      hipLaunchKernelGGL(k_solve_diag, dim3(ncb), dim3(256), 0, s, sa, slot, k, X, ldx, m, 1);
      if (k > 0) hipLaunchKernelGGL(k_solve_update, dim3(ncb, k), dim3(256), 0, s, sa, slot, k, X, ldx, m, 1);
    }
  }
  return hipGetLastError();
}

hipError_t launch_gen_cross(const double* sx, const double* sy, int64_t n, const double* px, const double* py, int m,
                            MaternParams mp, double* Xc, int64_t ldx, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_cross, dim3((unsigned)((n + 255) / 256 < 256 ? (n + 255) / 256 : 256), m), dim3(256), 0, s,
                     sx, sy, n, px, py, m, mp, Xc, ldx);
  return hipGetLastError();
}

hipError_t launch_col_dot(const double* X, int64_t ldx, int64_t n, int m, const double* w, double* out,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_col_dot, dim3(m), dim3(256), 0, s, X, ldx, n, w, out);
  return hipGetLastError();
}

}  // namespace tlr
