// Kernel launchers and the device-side layout of one likelihood evaluation ("slot").
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "common.hpp"

namespace tlr {

// status words: rank capacity exceeded; first non-PD tile (INT_MAX: none); a kernel guard tripped
// (rank or core out of range: the result is unusable and reported as an error)
enum { ST_RANKCAP = 0, ST_FAILED = 1, ST_GUARD = 2, ST_COUNT = 4 };

// width classes of the recompression, by r_jk: 1-8, 9-16, 17-32, 33-64
constexpr int RCLASSES = 4;
// counters per (step, class) of a work list: heavy entries, light entries, entries claimed, unused
constexpr int WL_STRIDE = 4;
// per-step lists beside the class lists: FP64 redo, FP32 redo, reduced-first
constexpr int REDO_LEVELS = 3;

// Device buffers of one slot (one theta being evaluated).  Layouts (all FP64 column-major):
//   D      nB tiles of b x b: diagonal tiles, factored in place into L_kk (lower)
//   Dinv   nB x npan blocks of 32 x 32: inverses of the diagonal blocks of L_kk
//   Linv   b x b: the whole L_kk^{-1} of the step being factored (TRSM by product)
//   U, V   ntile slots of b x rank_cap: off-diagonal factors, tile (i,j) at tri_index(i,j)
//   rank   ntile ints (current rank), rank0 the ranks after the initial compression
//   Wd, Xd nB x (rank_cap^2) and nB x (b x rank_cap): W_i = V_ik^T V_ik and U_ik W_i of step k
//   rhs    N: y on entry, z = L^{-1} y on exit
struct SlotPtrs {
  double* D;
  double* Dinv;
  double* Linv;       // b x b: L_kk^{-1} of the current step (lower)
  double* U;
  double* V;
  int* rank;
  int* rank0;
  double* Wd;
  double* Xd;
  double* rhs;
  double* logdet_part;
  int* status;
  double* result;   // [loglik, logdet, quad]
  int* rlog;          // rank chosen by every recompression, at rlog_off[k] + pair index
  int* rpl;           // arithmetic of every recompression (PM_FP64, PM_FP32, PM_BF16; -1 not run), as rlog
  int* prec;          // storage precision (8, 16, 32, 64) of every factor tile, mixed rule
  double* tnorm;      // per off-diagonal tile: a bound of ||U||_F ||V||_2 of its current factors (the initial
                      // compression and every recompression store ||U||_F with V orthonormal; TRSM scales it
                      // by ||V||_F), from which the list routes reduced-precision pairs
};

// Arithmetic of the factorization's recompressions (the products W = V_ik^T V_jk and U_ik W, the projections,
// Householder QRs, the Jacobi SVD of the core and the rebuilt factors): FP64; FP32; bfloat16 operands of the
// two products with FP32 accumulation and FP32 for the rest.  POTRF, TRSM, the diagonal updates, the
// initial compression and the reductions stay FP64 in every mode.
enum { PM_FP64 = 0, PM_FP32 = 1, PM_BF16 = 2 };
// A reduced-precision recompression goes ahead only when its budget is at least `guard` unit roundoffs of
// its working precision times the scale of the core (the larger of ||core||_F and the bound ||U_ik W||_F
// ||U_jk||_F of the update it absorbs); otherwise the pair goes to a redo list of a higher precision.
// The work lists apply the same test beforehand with an upper bound of that scale from the tiles' norms
// (tnorm_ij + tnorm_ik tnorm_jk): pairs that cannot pass go straight to the FP64 class lists, the others to
// the reduced-first list, whose kernel still checks the exact scale.
constexpr double GUARD_DEFAULT = 16384.0;
constexpr double EPS_FP32 = 1.1920928955078125e-07, EPS_BF16 = 3.90625e-03;   // 2^-23, 2^-8

struct SlotArgs {
  int64_t n;
  int b, nB, npan, rank_cap;
  int64_t fstride;      // b * rank_cap
  int64_t scr_stride;   // doubles of recompression scratch per resident work group
  double* scr_pool;     // scratch of the running recompression class (shared by all slots)
  int2* wl;             // recompression work lists: RCLASSES lists of wl_cap (pair, slot) entries
  int* wl_count;        // nB x RCLASSES x WL_STRIDE counters (zeroed at the start of every evaluation)
  int64_t wl_cap;
  int2* wlp;            // the same for the pairs (i, k+1) that update the next panel column (run first)
  int* wlp_count;
  int64_t wlp_cap;
  int2* wlr;            // (panel, level) lists: level 0 redo in FP64, 1 redo in FP32, 2 the pairs routed to
                        // the reduced precision at list time (REDO_LEVELS lists of wlr_cap per panel flag)
  int* wlr_count;       // nB x 2 (panel) x REDO_LEVELS x WL_STRIDE counters
  int64_t wlr_cap;
  int prec_mode;        // PM_FP64, PM_FP32 or PM_BF16 (arithmetic of the class launches)
  double guard;         // unit roundoffs of headroom a reduced-precision pair needs (GUARD_DEFAULT)
  int diag_wg;          // work groups per tile of the diagonal update
  const int* rep_rpl;   // records: arithmetic of each recompression, nlog per record (frozen in replays)
  int rule_fixed;       // 1: fixed relative tolerance, 0: adaptive absolute budget
  int mixed;            // 1: adaptive budget plus per-pair storage precision (rounded factors)
  const double* thr;    // per slot: (thr_rel, thr_abs)
  const int* replay;    // per slot: -1, or the record whose ranks this evaluation reuses (frozen ranks)
  const int* rep_rlog;  // records: recompression ranks, nlog per record
  const int64_t* rlog_off;   // nB offsets of step k in a rank log
  int64_t nlog;
  const SlotPtrs* slots;
  __device__ __forceinline__ SlotPtrs slot(int s) const { return slots[s]; }
};

// compress.hip
// mp: device pointer to the slot's Matern parameters (read by the kernels, so that a captured launch sees the
// current theta); thr2: device pointer to the slot's (thr_rel, thr_abs)
hipError_t launch_gen_diag(const double* sx, const double* sy, int64_t n, int b, int nB, const MaternParams* mp,
                           double* D, hipStream_t s);
hipError_t launch_matern_product(const double* sx, const double* sy, int64_t n, int b, int64_t ntile,
                                 const MaternParams* mp,
                                 int op, const double* X, int64_t xstride, double* out, int64_t ostride, int k,
                                 const int* active, hipStream_t s);
hipError_t launch_panel_qr(double* M, int64_t mstride, int b, int k, double* R, int64_t rstride, int r_transposed,
                           const int* active, int64_t ntile, hipStream_t s);
hipError_t launch_residual(const double* sx, const double* sy, int64_t n, int b, int64_t ntile, const MaternParams* mp,
                           const double* A, const double* B, int64_t stride, int k, double* part, const int* active,
                           hipStream_t s);
size_t rank_decide_lds(int k);
hipError_t launch_rank_decide(double* Rm, int64_t rstride, double* Qrow, const double* Qcol,
                              int64_t qstride, const double* part, int b, int k, int kmax, int pmin, int rule_fixed,
                              const double* thr2, double* U, double* V, int64_t fstride, int rank_cap,
                              int* rank, int* need, int* retry_count, int* status, const int* forced, int mixed,
                              double* tnorm, int64_t ntile, hipStream_t s);
hipError_t launch_warm_sketch(const double* stash, const int64_t* off, const int* r0, const double* omega, double* X,
                              int64_t xstride, int b, int k, const int* active, int64_t ntile, hipStream_t s);
hipError_t launch_fill_int(int* p, int64_t count, int value, hipStream_t s);
hipError_t launch_stash(double* U, double* V, int64_t fstride, int b, const int* r0, const int64_t* off,
                        double* stash, int* rank, int* rank0, int64_t ntile, int dir, hipStream_t s);
hipError_t launch_copy_int(int* dst, const int* src, int64_t count, hipStream_t s);
// Start-of-evaluation state of slots 0 .. m-1 (status, rhs = y, arithmetic log, mixed-rule precisions).
hipError_t launch_slot_init(const SlotPtrs* slots, int m, const double* y, int64_t n, int64_t ntile, int64_t nlog,
                            int mixed, hipStream_t s);
// Status words (ST_COUNT per slot), results (4 per slot) and, if rk, initial and final ranks (2 ntile per slot)
// of slots 0 .. m-1, for one copy to the host.
hipError_t launch_slot_gather(const SlotPtrs* slots, int m, int* st, double* res, int* rk, int64_t ntile,
                              hipStream_t s);
// Replay records (initial ranks, rank log, arithmetic log) of slots 0 .. m-1 into records 0 .. m-1.
hipError_t launch_slot_records(const SlotPtrs* slots, int m, int* rank0, int* rlog, int* rpl, int64_t ntile,
                               int64_t nlog, hipStream_t s);
// dense analysis of the rule's tiles: Sigma(theta) (N x N) and Delta = Sigma^TLR - Sigma from the compressed
// factors, with ||Delta_ij||_F^2 per tile
hipError_t launch_dense_cov(const double* sx, const double* sy, int64_t n, MaternParams mp, double* K, int64_t ldk,
                            hipStream_t s);
hipError_t launch_assemble_delta(const double* sx, const double* sy, int64_t n, int b, const MaternParams* mp,
                                 const double* U, const double* V, int64_t fstride, const int* rank, double* D,
                                 int64_t ldd, double* err2, int64_t ntile, hipStream_t s);
// ||A_ij||_F of the lower tiles of a dense matrix; its symmetric part in place (mode 0: lower -> upper, 1: mean)
hipError_t launch_tile_fro(const double* A, int64_t lda, int64_t n, int b, double* out, int64_t ntile, hipStream_t s);
hipError_t launch_symmetrize(double* A, int64_t lda, int64_t n, int mode, hipStream_t s);

// factor.hip
hipError_t launch_potrf(const SlotArgs& sa, int k, int nslot, hipStream_t s);
hipError_t launch_trtri(const SlotArgs& sa, int k, int nslot, hipStream_t s);
hipError_t launch_trsm(const SlotArgs& sa, int k, int nslot, hipStream_t s);
hipError_t launch_diag_update(const SlotArgs& sa, int k, int nslot, hipStream_t s);
hipError_t launch_recomp_list(const SlotArgs& sa, int k, int nslot, hipStream_t s);
// pm: arithmetic (PM_*); src: 0 the class's work list, 1 the FP64 redo list, 2 the FP32 redo list, 3 the
// reduced-first list
hipError_t launch_recomp_class(const SlotArgs& sa, int k, int cls, int panel, int grid, double* scr, hipStream_t s,
                               int pm = PM_FP64, int src = 0);
// resident work groups per CU of each recompression class in arithmetic pm (b selects the row-panel variant)
hipError_t recompress_occupancy(int b, int pm, int blocks_per_cu[RCLASSES]);
hipError_t launch_final(const SlotArgs& sa, int nslot, hipStream_t s);

// solve.hip
hipError_t launch_solve(const SlotArgs& sa, int slot, double* X, int64_t ldx, int m, int trans, hipStream_t s);
hipError_t launch_gen_cross(const double* sx, const double* sy, int64_t n, const double* px, const double* py, int m,
                            MaternParams mp, double* Xc, int64_t ldx, hipStream_t s);
hipError_t launch_col_dot(const double* X, int64_t ldx, int64_t n, int m, const double* w, double* out,
                          hipStream_t s);

}  // namespace tlr

// tlr_api.cpp: the calling thread's tlr_last_error message; returns code.
int tlr_set_error(int code, const char* msg);
// tlr_api.cpp: tlr_loglik on a context carved from caller-owned device memory [base, base + bytes) (null
// stream, persistent grids of at most grid_cap work groups, no device allocation).
int tlr_loglik_in_arena(void* base, size_t bytes, int grid_cap, const double* S, const double* y, int64_t n,
                        int32_t b, const double theta[3], int32_t rule, double value, int32_t precision,
                        double out3[3], int32_t* ranks_sigma, int32_t* ranks_factor, int32_t* failed_at);
