// Generation and initial compression of the Matern-3/2 covariance tiles.
//
// Reference semantics (codes/17_tlr_cholesky_pilot.py:155-166, 107-116): every off-diagonal tile
// T = sigma2 * matern32(dist(S_i, S_j), r) is truncated to the smallest rank r with
// ||T - T_r||_F <= thr (thr = eps ||T||_F for the fixed rule, tau b sigma2 eta / N for the adaptive one),
// stored as (U = U_r s_r, V = V_r).
//
// MI355X design: T is never stored.  Its entries are regenerated from the 2 x b site coordinates in
// every product pass (an exp and a sqrt per entry, no HBM traffic).  The subspace comes from alternating
// products with T and T^T (a randomized range finder with power iteration, Halko-Martinsson-Tropp), each
// followed by a register-resident Householder QR.  The rank decision uses an FP64-verified tail:
//   tail[r]^2 = ||T - T P||_F^2 (computed directly, entry by entry) + sum_{i>=r} sigma_i(core)^2,
// which differs from the exact SVD tail only by the subspace defect; a tile whose rank leaves less than
// `pmin` spare sketch columns is redone with a wider sketch.
#include "common.hpp"
#include "panel.hpp"
#include "round.hpp"
#include "tlr_kernels.hpp"

namespace tlr {

// ------------------------------------------------------------------------------------------ diagonal
// D_i = sigma2 matern32(S_i, S_i) + sigma2 eta I, padded to b x b with the identity beyond the tile.
__global__ void k_gen_diag(const double* __restrict__ sx, const double* __restrict__ sy, int64_t n, int b,
                           const MaternParams* __restrict__ mpp, double* __restrict__ D) {
  const MaternParams mp = *mpp;
  const int i = blockIdx.y;
  const int64_t i0 = (int64_t)i * b;
  const int bi = (int)min<int64_t>(b, n - i0);
  double* Di = D + (int64_t)i * b * b;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < (int64_t)b * b;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int p = (int)(e % b), q = (int)(e / b);
    double v;
    if (p < bi && q < bi) {
      v = matern32(mp, sx[i0 + p], sy[i0 + p], sx[i0 + q], sy[i0 + q]);
      if (p == q) v += mp.nugget;
    } else {
      v = (p == q) ? 1.0 : 0.0;
    }
    Di[e] = v;
  }
}

// --------------------------------------------------------------------------------- Matern products
// out_t (b x k) = op(T_t) X_t for every active tile t.  op = 0: T (rows: sites of tile i, inner index:
// sites of tile j); op = 1: T^T (the roles of the two site sets swap).  X_t has leading dimension b;
// xstride = 0 shares one X (the Gaussian test matrix) across tiles.
//
// On the matrix cores (v_mfma_f64_16x16x4f64), computing out^T = X^T op(T)^T so that the accumulators hold
// 16 consecutive rows of one output column per 16 lanes (stores in 128-byte runs).  Work group: 4 waves,
// each 32 rows x 16 NC columns (2 x NC accumulator blocks of 16 x 16); the work group covers 128 rows.  Per
// step of 4 inner indices a lane generates the 2 kernel entries of its operand (rows p0 + lane % 16 and
// p0 + 16 + lane % 16, inner index lane / 16) and reads NC operands of X from LDS; the 2 NC MFMAs reuse each
// generated entry for 16 NC columns.  On gfx950 the FP64 VALU work of the entries and the FP64 MFMAs do not
// overlap (the time per step is close to the sum of the two), so padding columns cost as much as real ones:
// the column blocks are sized to the width (48 columns: NC = 3, not 64 columns of which 16 are padding).
// Each output element is the same chain of MFMA steps over the inner index whatever NC is; on 3160-tile
// tests the outputs equal those of a sequential FMA loop over the inner index bit for bit.
constexpr int MM_ROWS = 128, MM_QC = 64;

// Work group -> (tile t, block blk) for 1-D grids of `per` blocks per tile.  The dispatcher deals work groups
// to the eight XCDs in turn (id mod 8), so the blocks of a tile are placed 8 apart in id order: every block of
// tile t runs on XCD t mod 8 and the operands the blocks share are fetched from HBM once, into that XCD's L2.
// Grid: ceil(ntile / 8) * 8 * per work groups; returns false past the last tile.
__device__ __forceinline__ bool xcd_tile_block(int64_t ntile, int per, int64_t* t, int* blk) {
  const int64_t id = blockIdx.x;
  const int64_t g = id / (8 * (int64_t)per);
  const int loc = (int)(id - g * 8 * per);
  *t = g * 8 + (loc & 7);
  *blk = loc >> 3;
  return *t < ntile;
}
typedef double dbl4 __attribute__((ext_vector_type(4)));

template <int NC>
__global__ __launch_bounds__(256, 4) void k_matern_product(const double* __restrict__ sx, const double* __restrict__ sy,
                                                           int64_t n, int b,
                                                           const MaternParams* __restrict__ mpp, int op,
                                                           const double* __restrict__ X, int64_t xstride,
                                                           double* __restrict__ out, int64_t ostride, int k,
                                                           const int* __restrict__ active, int64_t ntile) {
  // odd row pitch: the staging writes (64 consecutive inner indices of one column per wave) are conflict-free;
  // an operand read spans two rows per half-wave, which share banks at most two ways
  constexpr int COLS = 16 * NC, LD = COLS + 1;
  const int gx = (b + MM_ROWS - 1) / MM_ROWS, gy = (k + COLS - 1) / COLS;
  int64_t t;
  int blk;
  if (!xcd_tile_block(ntile, gx * gy, &t, &blk)) return;
  if (active && !active[t]) return;
  const int rblk = blk % gx, cblk = blk / gx;
  const MaternParams mp = *mpp;
  int i, j;
  tri_decode(t, &i, &j);
  if (op) { const int x = i; i = j; j = x; }
  const int64_t r0g = (int64_t)i * b, q0g = (int64_t)j * b;
  const int mr = (int)min<int64_t>(b, n - r0g), mq = (int)min<int64_t>(b, n - q0g);
  const int cb = cblk * COLS;
  const int tid = threadIdx.x, wv = tid >> 6, ln = tid & 63, lr = ln & 15, lq = ln >> 4;
  const int p0 = rblk * MM_ROWS + 32 * wv;   // this wave's first row
  __shared__ double Xs[MM_QC * LD];                // X chunk, [inner][column]
  __shared__ double qxs[MM_QC], qys[MM_QC];
  // the lane's two rows (sites); rows past the tile generate zeros
  const int pa = p0 + lr, pb = p0 + 16 + lr;
  const bool oka = pa < mr, okb = pb < mr;
  const double ax = oka ? sx[r0g + pa] : 0.0, ay = oka ? sy[r0g + pa] : 0.0;
  const double bx = okb ? sx[r0g + pb] : 0.0, by = okb ? sy[r0g + pb] : 0.0;
  const double* Xt = X + t * xstride;
  dbl4 acc[2][NC];
#pragma unroll
  for (int f = 0; f < 2; ++f)
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[f][c] = (dbl4){0.0, 0.0, 0.0, 0.0};
  for (int q0 = 0; q0 < mq; q0 += MM_QC) {
    __syncthreads();   // the previous chunk's reads are done
    for (int e = tid; e < MM_QC * COLS; e += 256) {
      const int qq = e % MM_QC, cc = e / MM_QC;
      const int q = q0 + qq, c = cb + cc;
      Xs[qq * LD + cc] = (q < mq && c < k) ? Xt[q + (int64_t)c * b] : 0.0;
    }
    if (tid < MM_QC) {
      const int q = q0 + tid;
      qxs[tid] = (q < mq) ? sx[q0g + q] : 0.0;
      qys[tid] = (q < mq) ? sy[q0g + q] : 0.0;
    }
    __syncthreads();
    const int qn = min(MM_QC, mq - q0);
    for (int s = 0; s < qn; s += 4) {
      const int qq = s + lq;
      const bool okq = qq < qn;
      const double qx = qxs[qq], qy = qys[qq];
      const double ta = (oka && okq) ? matern32(mp, ax, ay, qx, qy) : 0.0;
      const double tb = (okb && okq) ? matern32(mp, bx, by, qx, qy) : 0.0;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const double xv = Xs[qq * LD + 16 * c + lr];   // X^T[col 16 c + lane % 16][inner lane / 16]
        acc[0][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(xv, ta, acc[0][c], 0, 0, 0);
        acc[1][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(xv, tb, acc[1][c], 0, 0, 0);
      }
    }
  }
  // accumulator (f, c), item v: column cb + 16 c + lane / 16 + 4 v, row p0 + 16 f + lane % 16 (the MFMA's
  // result layout: item v of lane l holds D[l / 16 + 4 v][l % 16])
  double* Ot = out + t * ostride;
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    const int p = p0 + 16 * f + lr;
    if (p < b) {
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int col = cb + 16 * c + lq + 4 * v;
          if (col < k) Ot[p + (int64_t)col * b] = (p < mr) ? acc[f][c][v] : 0.0;
        }
    }
  }
}

// ------------------------------------------------------------------------------ orthonormalisation
// In place: M_t (b x k) -> Q_t (explicit, b x k) of its Householder QR; optionally R_t (k x k, col-major)
// stored to R (transposed when r_transposed).
template <int RPL, int NW, int CPW>
__global__ __launch_bounds__(64 * NW) void k_panel_qr(double* __restrict__ M, int64_t mstride, int b, int k,
                                                        double* __restrict__ R, int64_t rstride, int r_transposed,
                                                        const int* __restrict__ active) {
  const int64_t t = blockIdx.x;
  if (active && !active[t]) return;
  __shared__ double vbuf[2 * 64 * RPL];
  __shared__ double tau[NW * CPW];
  RegPanel<RPL, NW, CPW> P;
  double* Mt = M + t * mstride;
  // four columns per wave (the 33-64-column form): the reflector applications sum the wave's dot products
  // together (0.5-0.8 % on the log L + gradient step at N = 40,000); with eight per wave the compiler puts
  // the narrow forms' panel in scratch, so those keep one reduction per column
  constexpr bool MV = CPW == 4;
  P.load(Mt, b, b, k);
  if constexpr (MV) P.geqr2_mv(b, k, vbuf, tau);
  else P.geqr2(b, k, vbuf, tau);
  if (R) {
    if (r_transposed) P.store_rt(R + t * rstride, k, k);
    else P.store_r(R + t * rstride, k, k);
  }
  if constexpr (MV) P.org2r_mv(b, k, vbuf, tau);
  else P.org2r(b, k, vbuf, tau);
  P.store(Mt, b, b, k);
}

// ------------------------------------------------------------------------------------ residual
// For tile t with T ~ A B^T (A: b x k over the row sites of tile i, B: b x k over the sites of tile j),
// partial sums over 64 x 64 blocks of ||T - A B^T||_F^2 and ||T||_F^2, written to
// part[(t * nblk + blk) * 2 + {0, 1}].  Deterministic: the rank kernel adds them in block order.
// A B^T on the matrix cores (v_mfma_f64_16x16x4f64): each wave forms a 32 x 32 quarter of the block as
// 2 x 2 accumulator tiles, its operands read straight from the tile's panels (16-row runs of 128 bytes; the
// 64 blocks of a tile run on one XCD and share its L2, so there is no staging and no barrier in the loop);
// the entries of T are then generated in the accumulators' layout (item v of lane l: row 16 f + lane / 16 +
// 4 v, column 16 g + lane % 16 of the quarter) and the squared differences summed per lane, then over the
// wave and over the four waves in a fixed order.
__global__ __launch_bounds__(256) void k_residual(const double* __restrict__ sx, const double* __restrict__ sy,
                                                  int64_t n, int b, const MaternParams* __restrict__ mpp,
                                                  const double* __restrict__ A, const double* __restrict__ B,
                                                  int64_t stride, int k, double* __restrict__ part,
                                                  const int* __restrict__ active, int64_t ntile) {
  const int nbb = (b + 63) / 64;
  int64_t t;
  int blk;
  if (!xcd_tile_block(ntile, nbb * nbb, &t, &blk)) return;
  if (active && !active[t]) return;
  const MaternParams mp = *mpp;
  int i, j;
  tri_decode(t, &i, &j);
  const int64_t r0g = (int64_t)i * b, q0g = (int64_t)j * b;
  const int mr = (int)min<int64_t>(b, n - r0g), mq = (int)min<int64_t>(b, n - q0g);
  const int pb = (blk % nbb) * 64, qb = (blk / nbb) * 64;
  __shared__ double ps[2][64], qs[2][64];   // coordinates of the block's row and column sites
  __shared__ double red[2][4];
  const int tid = threadIdx.x, wv = tid >> 6, ln = tid & 63, lr = ln & 15, lq = ln >> 4;
  const int pw = (wv & 1) * 32, qw = (wv >> 1) * 32;   // the wave's quarter
  if (tid < 64) {
    const int p = pb + tid;
    ps[0][tid] = p < mr ? sx[r0g + p] : 0.0;
    ps[1][tid] = p < mr ? sy[r0g + p] : 0.0;
  } else if (tid < 128) {
    const int q = qb + tid - 64;
    qs[0][tid - 64] = q < mq ? sx[q0g + q] : 0.0;
    qs[1][tid - 64] = q < mq ? sy[q0g + q] : 0.0;
  }
  __syncthreads();
  const double* At = A + t * stride;
  const double* Bt = B + t * stride;
  dbl4 acc[2][2];
#pragma unroll
  for (int f = 0; f < 2; ++f)
#pragma unroll
    for (int g = 0; g < 2; ++g) acc[f][g] = (dbl4){0.0, 0.0, 0.0, 0.0};
  const int pa = pb + pw + lr, qa = qb + qw + lr;   // rows of the lane's operands (and + 16)
  const bool va0 = pa < b, va1 = pa + 16 < b, vb0 = qa < b, vb1 = qa + 16 < b;
#pragma unroll 4
  for (int s0 = 0; s0 < k; s0 += 4) {
    const int c = s0 + lq;
    const bool okc = c < k;   // columns past k add exact zeros
    const int64_t co = (int64_t)c * b;
    const double a0 = (okc && va0) ? At[pa + co] : 0.0, a1 = (okc && va1) ? At[pa + 16 + co] : 0.0;
    const double b0 = (okc && vb0) ? Bt[qa + co] : 0.0, b1 = (okc && vb1) ? Bt[qa + 16 + co] : 0.0;
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
  }
  double res2 = 0.0, fro2 = 0.0;
#pragma unroll
  for (int f = 0; f < 2; ++f)
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int ql = qw + 16 * g + lr;
      const bool okq = qb + ql < mq;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int pl = pw + 16 * f + lq + 4 * v;
        if (okq && pb + pl < mr) {
          const double tv = matern32(mp, ps[0][pl], ps[1][pl], qs[0][ql], qs[1][ql]);
          const double d = tv - acc[f][g][v];
          res2 += d * d;
          fro2 += tv * tv;
        }
      }
    }
  res2 = wave_sum(res2);
  fro2 = wave_sum(fro2);
  if (lane_id() == 0) { red[0][wave_id()] = res2; red[1][wave_id()] = fro2; }
  __syncthreads();
  if (tid == 0) {
    part[(t * nbb * nbb + blk) * 2 + 0] = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
    part[(t * nbb * nbb + blk) * 2 + 1] = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
  }
}

// ------------------------------------------------------------------- dense analysis of the rule's tiles
// Sigma(theta) of all N sites, N x N column-major (leading dimension ldk): sigma2 matern32 + sigma2 eta on the
// diagonal, entry for entry the values the diagonal tiles and the Matern products use (codes/11:61-65).
__global__ __launch_bounds__(256) void k_dense_cov(const double* __restrict__ sx, const double* __restrict__ sy,
                                                   int64_t n, MaternParams mp, double* __restrict__ K, int64_t ldk) {
  const int64_t q = blockIdx.x;                           // column
  for (int64_t p = (int64_t)blockIdx.y * 256 + threadIdx.x; p < n; p += (int64_t)gridDim.y * 256) {
    double v = matern32(mp, sx[p], sy[p], sx[q], sy[q]);
    if (p == q) v += mp.nugget;
    K[p + q * ldk] = v;
  }
}

// The assembled approximation's difference Delta = Sigma^TLR - Sigma on the off-diagonal tiles (codes/11:74-109:
// every tile (i, j), i > j, replaced by its rank-r truncation U V^T): Delta_ij = U_ij V_ij^T - T_ij written to D
// at (i, j) and transposed at (j, i) (column-major, leading dimension ldd; diagonal tiles untouched), and
// ||Delta_ij||_F^2 to err2[t].  One work group per tile, entries in a fixed order (reproducible sums).
__global__ __launch_bounds__(256) void k_assemble_delta(const double* __restrict__ sx, const double* __restrict__ sy,
                                                        int64_t n, int b, const MaternParams* __restrict__ mpp,
                                                        const double* __restrict__ U, const double* __restrict__ V,
                                                        int64_t fstride, const int* __restrict__ rank,
                                                        double* __restrict__ D, int64_t ldd,
                                                        double* __restrict__ err2) {
  const int64_t t = blockIdx.x;
  const MaternParams mp = *mpp;
  int i, j;
  tri_decode(t, &i, &j);
  const int64_t r0 = (int64_t)i * b, q0 = (int64_t)j * b;
  const int mr = (int)min<int64_t>(b, n - r0), mq = (int)min<int64_t>(b, n - q0);
  const int r = rank[t];
  const double* Ut = U + t * fstride;
  const double* Vt = V + t * fstride;
  __shared__ double red[4];
  double acc = 0.0;
  for (int64_t e = threadIdx.x; e < (int64_t)mr * mq; e += 256) {
    const int p = (int)(e % mr), q = (int)(e / mr);
    double s = 0.0;
    for (int c = 0; c < r; ++c) s = fma(Ut[p + (int64_t)c * b], Vt[q + (int64_t)c * b], s);
    const double d = s - matern32(mp, sx[r0 + p], sy[r0 + p], sx[q0 + q], sy[q0 + q]);
    D[(r0 + p) + (q0 + q) * ldd] = d;
    D[(q0 + q) + (r0 + p) * ldd] = d;
    acc = fma(d, d, acc);
  }
  acc = wave_sum(acc);
  if (lane_id() == 0) red[wave_id()] = acc;
  __syncthreads();
  if (threadIdx.x == 0) err2[t] = ((red[0] + red[1]) + red[2]) + red[3];
}

// ||A_ij||_F of every lower tile (i > j) of a dense N x N column-major matrix (leading dimension lda), tiles of
// b rows as tile_bounds(N, b), in tile order; one work group per tile, fixed-order sums.
__global__ __launch_bounds__(256) void k_tile_fro(const double* __restrict__ A, int64_t lda, int64_t n, int b,
                                                  double* __restrict__ out) {
  const int64_t t = blockIdx.x;
  int i, j;
  tri_decode(t, &i, &j);
  const int64_t r0 = (int64_t)i * b, q0 = (int64_t)j * b;
  const int mr = (int)min<int64_t>(b, n - r0), mq = (int)min<int64_t>(b, n - q0);
  __shared__ double red[4];
  double acc = 0.0;
  for (int64_t e = threadIdx.x; e < (int64_t)mr * mq; e += 256) {
    const double v = A[(r0 + e % mr) + (q0 + e / mr) * lda];
    acc = fma(v, v, acc);
  }
  acc = wave_sum(acc);
  if (lane_id() == 0) red[wave_id()] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[t] = sqrt(((red[0] + red[1]) + red[2]) + red[3]);
}

// Symmetric part of a square N x N matrix in place: mode 0 copies the lower triangle over the upper one, mode 1
// replaces both by their mean (A + A^T) / 2.  Entry (p, q), p > q, one thread each.
__global__ __launch_bounds__(256) void k_symmetrize(double* __restrict__ A, int64_t lda, int64_t n, int mode) {
  const int64_t q = blockIdx.y;
  for (int64_t p = q + 1 + (int64_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (int64_t)gridDim.x * 256) {
    const double lo = A[p + q * lda];
    if (mode == 0) {
      A[q + p * lda] = lo;
    } else {
      const double m = 0.5 * (lo + A[q + p * lda]);
      A[p + q * lda] = m;
      A[q + p * lda] = m;
    }
  }
}

// ------------------------------------------------------------------------------ rank decision
__global__ __launch_bounds__(256) void k_rank_decide(double* __restrict__ Rm, int64_t rstride,
                                                     double* Qrow, const double* __restrict__ Qcol,
                                                     int64_t qstride, const double* __restrict__ part, int b, int k,
                                                     int kmax, int pmin, int rule_fixed,
                                                     const double* __restrict__ thr2,
                                                     double* __restrict__ U, double* __restrict__ V, int64_t fstride,
                                                     int rank_cap, int* __restrict__ rank, int* __restrict__ need,
                                                     int* __restrict__ retry_count, int* __restrict__ status,
                                                     const int* __restrict__ forced, int mixed,
                                                     double* __restrict__ tnorm) {
  const int64_t t = blockIdx.x;
  if (!need[t]) return;
  const double thr_rel = thr2[0], thr_abs = thr2[1];
  extern __shared__ double lds[];
  double* G = lds;                // k x k
  double* J = G + k * k;          // k x k
  double* sig = J + k * k;        // k
  double* tail = sig + k;         // k + 1
  int* perm = (int*)(tail + k + 1);
  __shared__ int flag;
  __shared__ int s_rank;
  __shared__ double s_res2, s_fro2;
  __shared__ double s_red[8];
  const int tid = threadIdx.x;
  const double* Rt = Rm + t * rstride;
  for (int e = tid; e < k * k; e += blockDim.x) G[e] = Rt[e];
  if (tid == 0) {
    // residual and Frobenius norm from k_residual's partial sums (a forced rank needs neither, and the
    // residual pass is not run for it)
    const int nbb = (b + 63) / 64;
    double r2 = 0.0, f2 = 0.0;
    if (!forced)
      for (int e = 0; e < nbb * nbb; ++e) {
        r2 += part[(t * nbb * nbb + e) * 2 + 0];
        f2 += part[(t * nbb * nbb + e) * 2 + 1];
      }
    s_res2 = r2; s_fro2 = f2;
  }
  __syncthreads();
  // k = 64: 32 column pairs of 8 lanes fill the 256 threads (register-held rows); k = 96: the general form
  const int jr = k <= 64 ? jacobi_svd_mw<8, 8>(G, k, J, k, k, sig, perm, 60)
                          : jacobi_svd<4>(G, k, J, k, k, sig, perm, &flag, 60);
  if (jr == -2) {   // non-finite tile: reported, not stored
    if (tid == 0) { atomicMax(&status[ST_GUARD], 2); need[t] = 0; }
    return;
  }
  if (tid == 0) {
    // tail[r] = sqrt(res2 + sum_{i >= r} sig_perm[i]^2), r = 0..k
    double acc = s_res2;
    tail[k] = sqrt(acc);
    for (int r = k - 1; r >= 0; --r) {
      const double s = sig[perm[r]];
      acc += s * s;
      tail[r] = sqrt(acc);
    }
    const double thr = rule_fixed ? thr_rel * sqrt(s_fro2) : thr_abs;
    int r = -1;
    if (forced) {
      // replay of a recorded evaluation: the rank is given, the sketch must still leave pmin spare columns
      r = forced[t] <= k ? forced[t] : -1;
    } else {
      for (int c = 0; c <= k; ++c)
        if (tail[c] <= thr) { r = c; break; }
    }
    int accept;
    if (r < 0) accept = 0;                              // residual alone exceeds the budget
    else accept = (r + pmin <= k) || (k >= kmax);
    if (r < 0 && k >= kmax) { atomicMax(&status[0], 1); r = k; accept = 1; }
    if (accept && r > rank_cap) { atomicMax(&status[0], 1); r = rank_cap; }
    s_rank = accept ? r : -1;
  }
  __syncthreads();
  const int r = s_rank;
  if (r < 0) {
    if (tid == 0) {
      need[t] = 1;
      if (retry_count) atomicAdd(retry_count, 1);
    }
    return;
  }
  double* Qr = Qrow + t * qstride;
  const double* Qc = Qcol + t * qstride;
  double* Ut = U + t * fstride;
  double* Vt = V + t * fstride;
  // U = Qrow G[:, perm[:r]], V = Qcol J[:, perm[:r]]: a thread owns a row p and keeps CB output columns in
  // registers, so each Q panel is read once per CB columns (not once per column); the sum over l runs in the
  // same order as a column-by-column product
  constexpr int CB = 8;
  for (int c0 = 0; c0 < r; c0 += CB) {
    const int nc = min(CB, r - c0);   // uniform over the work group
    int pc[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) pc[j] = perm[c0 + (j < nc ? j : 0)];
    for (int p = tid; p < b; p += blockDim.x) {
      double u[CB], v[CB];
#pragma unroll
      for (int j = 0; j < CB; ++j) u[j] = v[j] = 0.0;
      for (int l = 0; l < k; ++l) {
        const double qr = Qr[p + (int64_t)l * b], qc = Qc[p + (int64_t)l * b];
#pragma unroll
        for (int j = 0; j < CB; ++j) {
          if (j < nc) {
            u[j] = fma(qr, G[l + pc[j] * k], u[j]);
            v[j] = fma(qc, J[l + pc[j] * k], v[j]);
          }
        }
      }
#pragma unroll
      for (int j = 0; j < CB; ++j) {
        if (j < nc) {
          Ut[p + (int64_t)(c0 + j) * b] = u[j];
          Vt[p + (int64_t)(c0 + j) * b] = v[j];
        }
      }
    }
  }
  if (mixed) {
    // storage precision under the budget left after truncation (17:163-164), for every pair as there; the
    // tile's row basis (b x kcap doubles, read for the last time by the loop above) gives the Gram workspace:
    // 3 r^2 doubles fit for every rank up to 64 when b >= 128 (only smaller tiles keep the ranks above
    // sqrt(b kcap / 3) in FP64)
    __syncthreads();
    if (3 * (int64_t)r * r <= qstride) round_pair_wg(Ut, b, b, Vt, b, b, r, thr_abs - tail[r], Qr, s_red);
  }
  if (tid == 0) {
    // ||U||_F of the stored factors (V orthonormal): the kept singular values
    double kept = 0.0;
    for (int c = 0; c < r; ++c) kept += sig[perm[c]] * sig[perm[c]];
    tnorm[t] = sqrt(kept);
    rank[t] = r;
    need[t] = 0;
  }
}

__global__ void k_fill_int(int* p, int64_t count, int value) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (int64_t)gridDim.x * blockDim.x)
    p[e] = value;
}

__global__ void k_copy_int(int* dst, const int* src, int64_t count) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (int64_t)gridDim.x * blockDim.x)
    dst[e] = src[e];
}

// Per-slot state at the start of an evaluation, slots 0 .. m-1 in one launch (grid.y = slot): status words
// zeroed with ST_FAILED = INT_MAX, rhs = y, the arithmetic log = -1 (no recompression run) and, for the mixed
// rule, every tile's storage precision = 64.  One launch instead of four or five per slot: with 80 slots the
// per-call enqueue cost of the small fills and copies had left the GPU idle for milliseconds.
__global__ void k_slot_init(const SlotPtrs* __restrict__ slots, const double* __restrict__ y, int64_t n,
                            int64_t ntile, int64_t nlog, int mixed) {
  const SlotPtrs p = slots[blockIdx.y];
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, step = (int64_t)gridDim.x * blockDim.x;
  if (t0 < ST_COUNT) p.status[t0] = t0 == ST_FAILED ? 0x7fffffff : 0;
  for (int64_t e = t0; e < n; e += step) p.rhs[e] = y[e];
  for (int64_t e = t0; e < nlog; e += step) p.rpl[e] = -1;
  if (mixed)
    for (int64_t e = t0; e < ntile; e += step) p.prec[e] = 64;
}

// Results of slots 0 .. m-1 gathered for one copy to the host (grid.y = slot s): the status words, the
// result (log L, log-det, quadratic form, spare) and, when rk is given, the initial and final ranks of every
// tile (rk + 2 s ntile and rk + (2 s + 1) ntile).
__global__ void k_slot_gather(const SlotPtrs* __restrict__ slots, int* __restrict__ st, double* __restrict__ res,
                              int* __restrict__ rk, int64_t ntile) {
  const int s = blockIdx.y;
  const SlotPtrs p = slots[s];
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, step = (int64_t)gridDim.x * blockDim.x;
  if (t0 < ST_COUNT) st[s * ST_COUNT + t0] = p.status[t0];
  if (t0 < 4) res[s * 4 + t0] = p.result[t0];
  if (rk)
    for (int64_t e = t0; e < ntile; e += step) {
      rk[2 * s * ntile + e] = p.rank0[e];
      rk[(2 * s + 1) * ntile + e] = p.rank[e];
    }
}

// Replay records of slots 0 .. m-1 in one launch (grid.y = slot s): the initial ranks (ntile) and the rank
// and arithmetic logs (nlog each) of slot s into record s.
__global__ void k_slot_records(const SlotPtrs* __restrict__ slots, int* __restrict__ rank0, int* __restrict__ rlog,
                               int* __restrict__ rpl, int64_t ntile, int64_t nlog) {
  const int s = blockIdx.y;
  const SlotPtrs p = slots[s];
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = t0; e < ntile; e += step) rank0[s * ntile + e] = p.rank0[e];
  for (int64_t e = t0; e < nlog; e += step) {
    rlog[s * nlog + e] = p.rlog[e];
    rpl[s * nlog + e] = p.rpl[e];
  }
}

// Compact copy of the initially compressed factors of one slot: tile t's first r0[t] columns of U and V
// (b rows each) live at stash + 2 b off[t] (U) and stash + 2 b off[t] + b r0[t] (V), off = exclusive prefix
// sum of r0.  dir 0: slot -> stash; dir 1: stash -> slot, and rank[t] = rank0[t] = r0[t].
__global__ void k_stash(double* __restrict__ U, double* __restrict__ V, int64_t fstride, int b,
                        const int* __restrict__ r0, const int64_t* __restrict__ off, double* __restrict__ stash,
                        int* __restrict__ rank, int* __restrict__ rank0, int dir, int64_t t0) {
  const int64_t t = t0 + blockIdx.y;
  const int r = r0[t];
  double* st = stash + 2 * (int64_t)b * off[t];
  double* Ut = U + t * fstride;
  double* Vt = V + t * fstride;
  const int64_t cnt = (int64_t)b * r;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * blockDim.x) {
    if (dir == 0) {
      st[e] = Ut[e];
      st[cnt + e] = Vt[e];
    } else {
      Ut[e] = st[e];
      Vt[e] = st[cnt + e];
    }
  }
  if (dir == 1 && blockIdx.x == 0 && threadIdx.x == 0) { rank[t] = r; rank0[t] = r; }
}

// Warm-started sketch of a replay with given ranks: X_t = [V_t of the recorded evaluation (its r0_t stored
// columns), the test matrix's columns r0_t .. k-1], so that the first product T X_t already spans the
// dominant range of T up to the change in theta (the right factor V_t spans the recorded tile's row space).
__global__ void k_warm_sketch(const double* __restrict__ stash, const int64_t* __restrict__ off,
                             const int* __restrict__ r0, const double* __restrict__ omega, double* __restrict__ X,
                             int64_t xstride, int b, int k, const int* __restrict__ active, int64_t t0) {
  const int64_t t = t0 + blockIdx.y;
  if (active && !active[t]) return;
  const int r = min(r0[t], k);
  const int64_t cnt = (int64_t)b * r0[t];
  const double* Vs = stash + 2 * (int64_t)b * off[t] + cnt;   // V part of the record's tile t (b x r0)
  double* Xt = X + t * xstride;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < (int64_t)b * k;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e / b);
    Xt[e] = c < r ? Vs[e] : omega[e];
  }
}

// ------------------------------------------------------------------------------------ launchers
static int nblk(int64_t a, int64_t d) { return (int)((a + d - 1) / d); }
// Grids whose y or z dimension runs over tiles are launched in chunks of at most this many tiles (the
// limit of those grid dimensions), each with its first tile as an argument.
constexpr int64_t TILE_CHUNK = 65535;

hipError_t launch_stash(double* U, double* V, int64_t fstride, int b, const int* r0, const int64_t* off,
                        double* stash, int* rank, int* rank0, int64_t ntile, int dir, hipStream_t s) {
  for (int64_t t0 = 0; t0 < ntile; t0 += TILE_CHUNK) {
    const int64_t nt = ntile - t0 < TILE_CHUNK ? ntile - t0 : TILE_CHUNK;
    hipLaunchKernelGGL(k_stash, dim3(4, (unsigned)nt), dim3(256), 0, s, U, V, fstride, b, r0, off, stash, rank, rank0,
                       dir, t0);
  }
  return hipGetLastError();
}

hipError_t launch_warm_sketch(const double* stash, const int64_t* off, const int* r0, const double* omega, double* X,
                              int64_t xstride, int b, int k, const int* active, int64_t ntile, hipStream_t s) {
  for (int64_t t0 = 0; t0 < ntile; t0 += TILE_CHUNK) {
    const int64_t nt = ntile - t0 < TILE_CHUNK ? ntile - t0 : TILE_CHUNK;
    hipLaunchKernelGGL(k_warm_sketch, dim3(nblk((int64_t)b * k, 256 * 8), (unsigned)nt), dim3(256), 0, s, stash, off,
                       r0, omega, X, xstride, b, k, active, t0);
  }
  return hipGetLastError();
}

hipError_t launch_fill_int(int* p, int64_t count, int value, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_int, dim3(min(nblk(count, 256), 1024)), dim3(256), 0, s, p, count, value);
  return hipGetLastError();
}

hipError_t launch_copy_int(int* dst, const int* src, int64_t count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_int, dim3(min(nblk(count, 256), 1024)), dim3(256), 0, s, dst, src, count);
  return hipGetLastError();
}

hipError_t launch_slot_init(const SlotPtrs* slots, int m, const double* y, int64_t n, int64_t ntile, int64_t nlog,
                            int mixed, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  int64_t most = n > nlog ? n : nlog;
  if (ntile > most) most = ntile;
  hipLaunchKernelGGL(k_slot_init, dim3(min(nblk(most, 256), 64), (unsigned)m), dim3(256), 0, s, slots, y, n, ntile,
                     nlog, mixed);
  return hipGetLastError();
}

hipError_t launch_slot_gather(const SlotPtrs* slots, int m, int* st, double* res, int* rk, int64_t ntile,
                              hipStream_t s) {
  if (m <= 0) return hipSuccess;
  const int64_t most = rk && ntile > 4 ? ntile : 4;
  hipLaunchKernelGGL(k_slot_gather, dim3(min(nblk(most, 256), 64), (unsigned)m), dim3(256), 0, s, slots, st, res, rk,
                     ntile);
  return hipGetLastError();
}

hipError_t launch_slot_records(const SlotPtrs* slots, int m, int* rank0, int* rlog, int* rpl, int64_t ntile,
                               int64_t nlog, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  const int64_t most = ntile > nlog ? ntile : nlog;
  if (most <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_slot_records, dim3(min(nblk(most, 256), 64), (unsigned)m), dim3(256), 0, s, slots, rank0,
                     rlog, rpl, ntile, nlog);
  return hipGetLastError();
}

hipError_t launch_gen_diag(const double* sx, const double* sy, int64_t n, int b, int nB, const MaternParams* mp,
                           double* D, hipStream_t s) {
  dim3 grid(nblk((int64_t)b * b, 256 * 8), nB);
  hipLaunchKernelGGL(k_gen_diag, grid, dim3(256), 0, s, sx, sy, n, b, mp, D);
  return hipGetLastError();
}

hipError_t launch_matern_product(const double* sx, const double* sy, int64_t n, int b, int64_t ntile,
                                 const MaternParams* mp,
                                 int op, const double* X, int64_t xstride, double* out, int64_t ostride, int k,
                                 const int* active, hipStream_t s) {
  if (ntile <= 0) return hipSuccess;
  const int64_t groups = (ntile + 7) / 8;
  // column blocks of 16 per work group: the fewest padded columns, ties to the wider group (fewer
  // regenerations of the tile); 48 and 96 columns take 3 blocks, 64 and 128 take 4
  const int nc = k <= 32 ? 2 : (nblk(k, 48) * 48 < nblk(k, 64) * 64 ? 3 : 4);
  const unsigned grid = (unsigned)(groups * 8 * nblk(b, MM_ROWS) * nblk(k, 16 * nc));
  if (nc == 2)
    hipLaunchKernelGGL(k_matern_product<2>, dim3(grid), dim3(256), 0, s, sx, sy, n, b, mp, op, X, xstride, out,
                       ostride, k, active, ntile);
  else if (nc == 3)
    hipLaunchKernelGGL(k_matern_product<3>, dim3(grid), dim3(256), 0, s, sx, sy, n, b, mp, op, X, xstride, out,
                       ostride, k, active, ntile);
  else
    hipLaunchKernelGGL(k_matern_product<4>, dim3(grid), dim3(256), 0, s, sx, sy, n, b, mp, op, X, xstride, out,
                       ostride, k, active, ntile);
  return hipGetLastError();
}

template <int RPL, int NW, int CPW = 8>
static void launch_qr_t(double* M, int64_t mstride, int b, int k, double* R, int64_t rstride, int rt,
                        const int* active, int64_t ntile, hipStream_t s) {
  hipLaunchKernelGGL((k_panel_qr<RPL, NW, CPW>), dim3((unsigned)ntile), dim3(64 * NW), 0, s, M, mstride, b, k, R,
                     rstride, rt, active);
}

hipError_t launch_panel_qr(double* M, int64_t mstride, int b, int k, double* R, int64_t rstride, int r_transposed,
                           const int* active, int64_t ntile, hipStream_t s) {
  if (b > 512 || k > 128 || k > b) return hipErrorInvalidValue;
  const int nw = (k + 7) / 8;
  if (b <= 256) {
    if (nw <= 4) launch_qr_t<4, 4>(M, mstride, b, k, R, rstride, r_transposed, active, ntile, s);
    else if (nw <= 8) launch_qr_t<4, 8>(M, mstride, b, k, R, rstride, r_transposed, active, ntile, s);
    else if (nw <= 12) launch_qr_t<4, 12>(M, mstride, b, k, R, rstride, r_transposed, active, ntile, s);
    else launch_qr_t<4, 16>(M, mstride, b, k, R, rstride, r_transposed, active, ntile, s);
  } else {
    // 64 columns: 16 waves of 4 columns (a quarter of the register panel per wave of the 8 x 8 form, so that
    // four waves share each SIMD and the per-column reflector work is split over twice as many waves)
    if (nw <= 4) launch_qr_t<8, 4>(M, mstride, b, k, R, rstride, r_transposed, active, ntile, s);
    else if (nw <= 8) launch_qr_t<8, 16, 4>(M, mstride, b, k, R, rstride, r_transposed, active, ntile, s);
    else if (nw <= 12) launch_qr_t<8, 12>(M, mstride, b, k, R, rstride, r_transposed, active, ntile, s);
    else launch_qr_t<8, 16>(M, mstride, b, k, R, rstride, r_transposed, active, ntile, s);
  }
  return hipGetLastError();
}

hipError_t launch_residual(const double* sx, const double* sy, int64_t n, int b, int64_t ntile, const MaternParams* mp,
                           const double* A, const double* B, int64_t stride, int k, double* part, const int* active,
                           hipStream_t s) {
  const int nbb = nblk(b, 64);
  if (ntile <= 0) return hipSuccess;
  const int64_t groups = (ntile + 7) / 8;
  hipLaunchKernelGGL(k_residual, dim3((unsigned)(groups * 8 * nbb * nbb)), dim3(256), 0, s, sx, sy, n, b, mp, A, B,
                     stride, k, part, active, ntile);
  return hipGetLastError();
}

size_t rank_decide_lds(int k) {
  return (size_t)(2 * k * k + 2 * k + 1) * sizeof(double) + (size_t)k * sizeof(int);
}

hipError_t launch_rank_decide(double* Rm, int64_t rstride, double* Qrow, const double* Qcol,
                              int64_t qstride, const double* part, int b, int k, int kmax, int pmin, int rule_fixed,
                              const double* thr2, double* U, double* V, int64_t fstride, int rank_cap,
                              int* rank, int* need, int* retry_count, int* status, const int* forced, int mixed,
                              double* tnorm, int64_t ntile, hipStream_t s) {
  const size_t lds = rank_decide_lds(k);
  if (lds > 158 * 1024) return hipErrorInvalidValue;
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute((const void*)k_rank_decide, hipFuncAttributeMaxDynamicSharedMemorySize, 158 * 1024);
    attr_set = true;
  }
  hipLaunchKernelGGL(k_rank_decide, dim3((unsigned)ntile), dim3(256), lds, s, Rm, rstride, Qrow, Qcol, qstride, part,
                     b, k, kmax, pmin, rule_fixed, thr2, U, V, fstride, rank_cap, rank, need, retry_count,
                     status, forced, mixed, tnorm);
  return hipGetLastError();
}

hipError_t launch_dense_cov(const double* sx, const double* sy, int64_t n, MaternParams mp, double* K, int64_t ldk,
                            hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int rows = nblk(n, 256) < 1024 ? nblk(n, 256) : 1024;
  hipLaunchKernelGGL(k_dense_cov, dim3((unsigned)n, (unsigned)rows), dim3(256), 0, s, sx, sy, n, mp, K, ldk);
  return hipGetLastError();
}

hipError_t launch_assemble_delta(const double* sx, const double* sy, int64_t n, int b, const MaternParams* mp,
                                 const double* U, const double* V, int64_t fstride, const int* rank, double* D,
                                 int64_t ldd, double* err2, int64_t ntile, hipStream_t s) {
  if (ntile <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_assemble_delta, dim3((unsigned)ntile), dim3(256), 0, s, sx, sy, n, b, mp, U, V, fstride, rank,
                     D, ldd, err2);
  return hipGetLastError();
}

hipError_t launch_tile_fro(const double* A, int64_t lda, int64_t n, int b, double* out, int64_t ntile, hipStream_t s) {
  if (ntile <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tile_fro, dim3((unsigned)ntile), dim3(256), 0, s, A, lda, n, b, out);
  return hipGetLastError();
}

hipError_t launch_symmetrize(double* A, int64_t lda, int64_t n, int mode, hipStream_t s) {
  if (n <= 1 || n > 65535) return n <= 1 ? hipSuccess : hipErrorInvalidValue;
  const int rows = nblk(n, 256) < 256 ? nblk(n, 256) : 256;
  hipLaunchKernelGGL(k_symmetrize, dim3((unsigned)rows, (unsigned)n), dim3(256), 0, s, A, lda, n, mode);
  return hipGetLastError();
}

}  // namespace tlr
