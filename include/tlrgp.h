/*
 * tlrgp: C ABI of the MI355X adaptive-rank tile low-rank (TLR) Gaussian-process log-likelihood.
 *
 * The reference has no FFI: its boundary is the Python call
 *   tlr_loglik(S, y, theta, bounds, rule, value) -> (ll | None, (L_diag, L_off) | None, info)
 * in codes/17_tlr_cholesky_pilot.py:259-266 (factorization 17:137-215, forward solve 17:230-240), wrapped
 * for samplers by Lik.__call__ in codes/38_tlr_cholesky_mcmc.py:43-56 (failure -> -inf) and for
 * optimizers by neg_loglik_tlr in 17:279-290 (failure -> 1e12).  The entry points below are what a
 * ctypes binding of that call needs; the Python facade in adaptive-tlr-pilots_amd/tlrgp restores the
 * reference signatures on top of them.
 *
 * Conventions
 *  - S is float64, shape (N, 2), row-major (x0, y0, x1, y1, ...), Morton-ordered by the caller; y is
 *    float64 (N).  Both are host pointers and are copied at context creation; the caller keeps ownership.
 *  - Tiles follow tile_bounds(N, b) of codes/11_tlr_rank_rules.py:69-71: nB = ceil(N / b) tiles, the last
 *    one possibly short; the rank rule uses the maximum tile size b (17:143).
 *  - theta = (sigma2, range, eta); the covariance is sigma2 * matern32(h, range) + sigma2 * eta * I.
 *  - rule TLR_RULE_FIXED: value = eps (relative Frobenius tolerance per tile, 17:114, 17:131);
 *    rule TLR_RULE_ADAPTIVE: value = tau, absolute budget tau * b * sigma2 * eta / N (17:150);
 *    rule TLR_RULE_MIXED: the adaptive budget, and every factor pair stored in the lowest of FP8-E4M3,
 *    FP16, FP32, FP64 whose rounding error fits the budget left after truncation (17:37-82, 146-147,
 *    163-164, 183-184, 203-204); arithmetic stays FP64.
 *  - Rank arrays are dense nB x nB int32, row-major, entry [i * nB + j] for the tile (i, j), i > j;
 *    other entries are 0.  ranks_sigma: after the initial compression; ranks_factor: of the factor L.
 *  - Return values: TLR_OK when the factorization completed; TLR_NOT_PD when a diagonal tile was not
 *    positive definite (failed_at receives the first failing tile k, as info["failed_at"] in 17:175);
 *    a negative value on error, with a message from tlr_last_error().
 *  - One context per host thread; calls on one context are serialised by the caller.
 */
#ifndef TLRGP_H
#define TLRGP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TLR_OK 0
#define TLR_NOT_PD 1
/* Per-theta status of the batch entry points (status[s]) when a tile of that theta needed more columns than
 * the rank capacity of 64: its outputs are NaN and the batch's other theta are evaluated as usual.  A call
 * with one theta returns TLR_ERR_RANKCAP instead.  No counterpart in the reference, which has no capacity. */
#define TLR_RANK_CAPPED 2
#define TLR_ERR_ARG (-1)
#define TLR_ERR_HIP (-2)
#define TLR_ERR_RANKCAP (-3)
#define TLR_ERR_NODEVICE (-4)

#define TLR_RULE_FIXED 0
#define TLR_RULE_ADAPTIVE 1
#define TLR_RULE_MIXED 2

/* Arithmetic of the factorization's off-diagonal low-rank work: the products W = V_ik^T V_jk and U_ik W of
 * every trailing update and the recompression of the updated tile (QR of both concatenations, SVD of the
 * core, rebuilt factors; 17:119-133, 187-205).  POTRF, TRSM, the diagonal updates, the initial compression
 * with its rank decision and the log-determinant stay FP64 in every mode.
 *  TLR_PREC_FP64  plain FP64;
 *  TLR_PREC_FP32  FP32 throughout the recompression;
 *  TLR_PREC_BF16  the two products with bfloat16 operands and FP32 accumulation, the rest in FP32.
 * A pair whose budget is below `guard` (16384 by default, TLR_TUNE_GUARD) unit roundoffs of the working
 * precision times the core's scale (the larger of ||core||_F and ||U_ik W||_F ||U_jk||_F) is recompressed at
 * the next higher precision instead (BF16 -> FP32 -> FP64); the work lists apply the test first with an
 * upper bound of that scale from per-tile norms and send the pairs that cannot pass straight to FP64.  The frozen-rank replays of the gradient keep each pair's arithmetic.  The storage rule
 * TLR_RULE_MIXED runs with TLR_PREC_FP64 only.  No counterpart in the reference (its arithmetic is FP64). */
#define TLR_PREC_FP64 0
#define TLR_PREC_FP32 1
#define TLR_PREC_BF16 2

typedef struct tlr_ctx tlr_ctx;

/* Upload sites and observations; allocate device workspace for up to max_batch simultaneous theta.
 * Replaces the per-call setup of tlr_cholesky (17:141-150) and the data held by Lik (38:40-42). */
int tlr_ctx_create(tlr_ctx** ctx, const double* S, const double* y, int64_t n, int32_t b, int32_t max_batch);

int tlr_ctx_destroy(tlr_ctx* ctx);

/* N, b and the number of tiles nB of a context. */
int tlr_ctx_dims(const tlr_ctx* ctx, int64_t* n, int32_t* b, int32_t* n_tiles);

/* Replace the observations y (N float64, host) of a context. */
int tlr_ctx_set_y(tlr_ctx* ctx, const double* y);

/* TLR log-likelihood at one theta: replaces tlr_loglik (17:259-266).
 * out3 receives (loglik, logdet, quad) with loglik = -(quad + logdet + N log 2 pi) / 2, logdet =
 * 2 sum_k sum log diag L_kk and quad = z^T z, z = L^{-1} y (17:263-265).  ranks_sigma / ranks_factor
 * (nB * nB each) may be NULL. */
int tlr_loglik(tlr_ctx* ctx, const double theta[3], int32_t rule, double value, int32_t precision, double out3[3],
               int32_t* ranks_sigma, int32_t* ranks_factor, int32_t* failed_at);

/* m theta at once (m <= max_batch), each with its own ranks; thetas is m x 3, out is m x 3, status and
 * failed_at are m-vectors (per-theta TLR_OK / TLR_NOT_PD).  ranks_* are m x nB x nB or NULL.
 * The batched form of repeated tlr_loglik calls (Lik.__call__ per proposal, 38:43-56). */
int tlr_loglik_batch(tlr_ctx* ctx, int32_t m, const double* thetas, int32_t rule, double value, int32_t precision,
                     double* out, int32_t* status, int32_t* failed_at, int32_t* ranks_sigma, int32_t* ranks_factor);

/* tlr_loglik_batch through a HIP graph: the first call of a configuration (m, rule, precision) runs as
 * tlr_loglik_batch and then captures the same device work (compression, every factorization step on the
 * context's streams, the reductions) into a graph; later calls with that configuration copy the new theta and
 * rule value into the context's staging buffers and launch the graph, one launch per batch instead of about
 * twenty per tile step.  Results equal tlr_loglik_batch's bit for bit.  Changing a tunable drops the graphs.
 * No counterpart in the reference (the batched sampler evaluation of 38:89-126, 19:170-203 is one process per
 * chain). */
int tlr_loglik_batch_graph(tlr_ctx* ctx, int32_t m, const double* thetas, int32_t rule, double value,
                           int32_t precision, double* out, int32_t* status, int32_t* failed_at, int32_t* ranks_sigma,
                           int32_t* ranks_factor);

/* Log-likelihood and its gradient in (log sigma2, log range, log eta) at frozen ranks: the ranks of every
 * tile after the initial compression and after every recompression are those chosen at theta.  The
 * sigma2 component is exact (0.5 * quad - N / 2: tiles and budgets scale together with sigma2); the range
 * and eta components are central differences in log theta with step h (1e-5 recommended), each
 * evaluation replaying the recorded ranks (the eta evaluations reuse theta's compressed tiles, the range
 * evaluations start their compression sketch from them).  grad is m x 3 (NaN when a perturbed evaluation
 * is not PD).
 * No counterpart in the reference, which has no gradient (its standard errors use finite differences of
 * log L with re-chosen ranks, codes/15_pilot_standard_errors.py:21-31). */
int tlr_loglik_grad_batch(tlr_ctx* ctx, int32_t m, const double* thetas, int32_t rule, double value, int32_t precision,
                          double h, double* out, double* grad, int32_t* status, int32_t* failed_at,
                          int32_t* ranks_sigma, int32_t* ranks_factor);
int tlr_loglik_grad(tlr_ctx* ctx, const double theta[3], int32_t rule, double value, int32_t precision, double h,
                    double out3[3], double grad3[3], int32_t* ranks_sigma, int32_t* ranks_factor, int32_t* failed_at);

/* Kriging prediction through the TLR factor at theta: replaces tlr_predict (17:269-276) with its solves
 * tlr_forward / tlr_backward (17:230-256).  The factor is the one the last evaluation of batch slot 0 left
 * resident when that was this theta, rule, value and precision (the factor handle of tlr_loglik, as the
 * reference passes (L_diag, L_off) from tlr_loglik to tlr_predict), else a new factorization at theta (out3
 * receives its loglik, logdet, quad; may be NULL).  Sp is n_pred x 2 row-major (host); mu and sd (n_pred, host) receive
 * Kp L^{-T} L^{-1} y and sqrt(max(sigma2 (1 + eta) - colsum((L^{-1} Kp^T)^2), 1e-12)).  Returns TLR_NOT_PD
 * when the factorization fails. */
int tlr_predict(tlr_ctx* ctx, const double theta[3], int32_t rule, double value, int32_t precision, const double* Sp,
                int64_t n_pred, double* mu, double* sd, double out3[3]);

/* Solve with the TLR factor at theta (resident or new, as for tlr_predict), in place on the host block B (N x m, column-major,
 * leading dimension ldb >= N): mode 0 B <- L^{-1} B (tlr_forward, 17:230-240), mode 1 B <- L^{-T} B
 * (tlr_backward, 17:243-256), mode 2 B <- (L L^T)^{-1} B.  out3 receives (loglik, logdet, quad) of the
 * factorization (may be NULL).  Returns TLR_NOT_PD when the factorization fails (B unchanged). */
int tlr_solve(tlr_ctx* ctx, const double theta[3], int32_t rule, double value, int32_t precision, double* B,
              int64_t ldb, int64_t m, int32_t mode, double out3[3]);

/* Factorizations a context has run (every theta of every batch, replays included): shows when tlr_predict /
 * tlr_solve reused a resident factor.  No counterpart in the reference. */
int tlr_ctx_stats(const tlr_ctx* ctx, int64_t* factorizations);

/* Timing of one kernel class with HIP events on the context's stream, for roofline reports (no
 * counterpart in the reference, which times whole evaluations with time.time(), 17:359-362).
 * kernel_class: 0 off, 1 Matern products, 2 residual, 3 panel QR, 5 one factorization step with lookahead
 * (the work lists and recompressions of step k with the POTRF, inverse, TRSM and diagonal update of step
 * k + 1 that overlap them, from the first launch to the join of the streams).  Setting it resets the
 * accumulators; tlr_ctx_get_timing returns the summed time, the number of timed launches (steps for class
 * 5), their algorithmic flops and their algorithmic HBM bytes (products: 2 b^2 k flops and 16 b k + 32 b
 * bytes per processed tile; class 5: per recompression pair 4 b r_ik r_jk + 8 b w^2 + 22 w^3 + 4 b w r'
 * flops and 16 b (r_ij + r_ik + r_jk + r') bytes, w = r_ij + r_jk, from the rank logs, plus POTRF and the
 * inverse 2 b^3 / 3, TRSM b^2 r and the diagonal update 2 b r^2 + 2 b^2 r per tile). */
int tlr_ctx_set_timing(tlr_ctx* ctx, int32_t kernel_class);
int tlr_ctx_get_timing(const tlr_ctx* ctx, double* total_ms, int64_t* launches, double* flops, double* bytes);

/* Storage precision (8, 16, 32 or 64 bits) of every factor tile of the last evaluation in batch slot
 * `slot` under TLR_RULE_MIXED: prec is nB x nB int32 row-major, entry [i * nB + j] for the tile (i, j),
 * 0 elsewhere.  Replaces PREC_OF and info["precision_mix_factor_tiles"] of 17:38, 210-214. */
int tlr_ctx_precisions(const tlr_ctx* ctx, int32_t slot, int32_t* prec);

/* Tunables of a context (no counterpart in the reference; set by tlrgp.autotune):
 *  TLR_TUNE_GUARD      unit roundoffs of headroom a reduced-precision recompression needs (default 16384):
 *                      budget >= guard * u * scale, u = 2^-23 (FP32) or 2^-8 (BF16);
 *  TLR_TUNE_DIAG_WG    work groups per tile of the diagonal update (default 9);
 *  TLR_TUNE_GRID_FRAC  persistent recompression grids as a fraction (0, 1] of the resident capacity;
 *  TLR_TUNE_SKETCH     columns of the first compression round's sketch (default 48; a multiple of 8 from 32
 *                      to the second round's width, min(96, b)): a rank is accepted only with 24 sketch
 *                      columns to spare, so a narrower first round sends more tiles to the second. */
#define TLR_TUNE_GUARD 0
#define TLR_TUNE_DIAG_WG 1
#define TLR_TUNE_GRID_FRAC 2
#define TLR_TUNE_SKETCH 3
/*  TLR_TUNE_PRODUCTS   products with T and T^T of a cold compression before its rank decision (default 4: two
 *                      power iterations; an integer from 2 to 6). */
#define TLR_TUNE_PRODUCTS 4
int tlr_ctx_set_tunable(tlr_ctx* ctx, int32_t which, double value);
int tlr_ctx_get_tunable(const tlr_ctx* ctx, int32_t which, double* value);

/* Recompressions of the last evaluation in batch slot `slot` by arithmetic: counts[0] FP64, counts[1] FP32,
 * counts[2] BF16 (pairs with r_ik = 0 or r_jk = 0 are not recompressed and not counted).  No counterpart in
 * the reference. */
int tlr_ctx_arith_counts(const tlr_ctx* ctx, int32_t slot, int64_t counts[3]);

/* tlr_loglik without a context and without any device allocation: no counterpart in the reference; for
 * callers that cannot call hipMalloc (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  Same arguments and results as tlr_loglik plus the field (S, y host pointers, N, b as
 * in tlr_ctx_create).  Its device workspace is a static array of TLR_NOALLOC_BYTES per GPU, allocated when
 * the library is loaded and held until the process ends; N <= TLR_NOALLOC_MAX_N and b <= TLR_NOALLOC_MAX_B
 * (larger problems return TLR_ERR_ARG).  Runs on the null stream and returns after a synchronisation; calls
 * on one GPU share the array, so this process serialises them, and calls that overlap on one GPU from
 * elsewhere would race. */
#define TLR_NOALLOC_BYTES (64ull << 20)
#define TLR_NOALLOC_MAX_N 2048
#define TLR_NOALLOC_MAX_B 512
int tlr_loglik_noalloc(const double* S, const double* y, int64_t n, int32_t b, const double theta[3], int32_t rule,
                       double value, int32_t precision, double out3[3], int32_t* ranks_sigma, int32_t* ranks_factor,
                       int32_t* failed_at);

/* Dense FP64 log-likelihood of m theta: out[3 s + {0, 1, 2}] = log L, log-determinant, quadratic form of the
 * full Sigma(theta) of the context's sites and observations (loglik_dense of codes/11_tlr_rank_rules.py:112-120,
 * by a Cholesky factorization), NaN when Sigma is not positive definite.  The exact likelihood beside the TLR
 * one, for the approximation's error; one N x N device array allocated and freed per call (N <= 46340),
 * rocSOLVER's Cholesky. */
int tlr_dense_loglik(tlr_ctx* ctx, int32_t m, const double* thetas, double* out);

/* The error terms of the assembled approximation Sigma^TLR (every off-diagonal tile replaced by the rule's
 * rank-r truncation, codes/11_tlr_rank_rules.py:74-109) that codes/36_likelihood_error_terms.py:75-136
 * decomposes, for nrules rules (rules[q] TLR_RULE_FIXED or TLR_RULE_ADAPTIVE, values[q] its value) at theta.
 * The rule's tiles are the initial compression of an evaluation at theta (batch slot 0, whose resident factor
 * is given up); Sigma, Delta = Sigma^TLR - Sigma and E = L^-1 Delta L^-T are dense N x N device arrays that
 * this call allocates and frees (five of them: N <= 46340).  Outputs (host):
 *   dense[2]             log L of Sigma, z^T z (z = L^-1 y);
 *   kinv_tiles[n_tiles]  ||(Sigma^-1)_ij||_F, cov_tiles[n_tiles] ||Sigma_ij||_F (tri_index order; optional);
 *   alpha_blocks[nB]     ||alpha_i|| of alpha = Sigma^-1 y per tile row (optional);
 *   terms[6 nrules]      per rule: log L of Sigma^TLR (NaN when it is not positive definite, and then the
 *                        rest NaN too), tr(Sigma^-1 Delta), alpha^T Delta alpha, ||E||_F, ||E||_2,
 *                        (Delta alpha)^T Sigma^-1 (Delta alpha);
 *   tile_err[nrules x n_tiles]  ||Delta_ij||_F (optional); ranks[nrules x nB x nB] as ranks_sigma (optional).
 * TLR_ERR_ARG when Sigma itself is not positive definite. */
int tlr_bound_terms(tlr_ctx* ctx, const double theta[3], int32_t nrules, const int32_t* rules, const double* values,
                    double dense[2], double* kinv_tiles, double* cov_tiles, double* alpha_blocks, double* terms,
                    double* tile_err, int32_t* ranks);

/* Message of the last error on the calling thread. */
const char* tlr_last_error(void);

/* Name of device 0 and its compute units, for reports. */
int tlr_device_info(char* buf, int32_t len);

#ifdef __cplusplus
}
#endif

#endif
