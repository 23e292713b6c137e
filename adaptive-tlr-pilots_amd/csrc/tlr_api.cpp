// C ABI of the TLR likelihood: context management and the per-evaluation launch schedule.
//
// Schedule of one batch of m theta (one slot per theta), on the context's stream (and two side streams):
//   per slot: diagonal tiles; initial compression of every off-diagonal tile (sketch width 64, then 96
//             for the tiles whose rank left too little oversampling)
//   for k = 0 .. nB-1 (all slots in each launch): POTRF(k) -> inverse of L_kk -> TRSM + forward solve(k) ->
//             diagonal updates(k) -> recompressions(k) (four width classes on three streams, with the next
//             step's POTRF, inverse, TRSM and diagonal update overlapped)
//   final reductions; one synchronisation and a small copy back.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/tlrgp.h"
#include "tlr_kernels.hpp"

using namespace tlr;

static thread_local std::string g_err;

static int set_err(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                              \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess)                                                                         \
      return set_err(TLR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));             \
  } while (0)

namespace {

constexpr int K1 = 48, K2 = 96;     // sketch widths of the two compression rounds (K1: default, tunable)
constexpr int K_WARM = 64;          // smallest first round of a warm-started (replay) compression
constexpr int PMIN = 24;            // spare sketch columns required to accept a rank
constexpr int NPRODUCT = 4;         // products with T, T^T, T, T^T before the rank decision (default, tunable)
constexpr int NCS = 3;              // streams that compress slots concurrently
constexpr int RANK_CAP = 64;        // column capacity of the U, V slots
constexpr int WCAP = 128;           // widest recompression concatenation

struct Slot {
  SlotPtrs p;
};

// A launch bracketed by HIP events on the context's stream (timing mode only).
struct TimedLaunch {
  hipEvent_t a, b;
  int slot, round;
  double flops_per_tile;   // algorithmic flops per processed tile
  double bytes_per_tile;   // algorithmic HBM bytes per processed tile
};

}  // namespace

// What the factor resident in a slot is: the evaluation at theta (rule, value, precision) with its result,
// or nothing (before the first evaluation, after an error, or after a frozen-rank replay, whose factor is not
// the plain one at its theta).
struct SlotState {
  bool valid = false;
  double theta[3] = {0.0, 0.0, 0.0};
  int rule = -1, precision = -1;
  double value = 0.0;
  double out3[3] = {0.0, 0.0, 0.0};
  int status = TLR_OK, failed_at = -1;
  bool matches(const double th[3], int r, double v, int p) const {
    return valid && th[0] == theta[0] && th[1] == theta[1] && th[2] == theta[2] && r == rule && v == value &&
           p == precision;
  }
};

struct tlr_ctx {
  int64_t n = 0;
  int b = 0, nB = 0, npan = 0, max_batch = 0;
  int64_t ntile = 0, npairs_max = 0;
  int kcap = 0;
  hipStream_t stream = nullptr;
  double *sx = nullptr, *sy = nullptr, *y = nullptr, *omega = nullptr;
  // compression workspace, one set per compression stream (the slots of a batch are compressed in turn on
  // the main stream and the two side streams)
  double *bufA[NCS] = {}, *bufB[NCS] = {}, *bufR[NCS] = {}, *part[NCS] = {};
  int* need[NCS] = {};
  int nset = 0;                        // workspace sets allocated: min(NCS, max_batch)
  std::vector<void*> allocs;
  std::vector<Slot> slots;
  std::vector<SlotState> sstate;       // per slot: the factor it holds
  // captured evaluations (tlr_loglik_batch_graph), by (m, rule, precision); dropped when a tunable changes
  std::map<std::tuple<int, int, int>, hipGraphExec_t> graphs;
  void drop_graphs() {
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
  }
  int64_t factorizations = 0;          // evaluations run (theta factorized), for reports
  SlotPtrs* d_slots = nullptr;
  int* d_gst = nullptr;                // gathered status words of a batch (ST_COUNT per slot)
  double* d_gres = nullptr;            // gathered results (4 per slot)
  int* d_grk = nullptr;                // gathered initial and final ranks (2 ntile per slot)
  int64_t scr_stride = 0;
  double* scr_pool = nullptr;          // recompression scratch, one scr_stride per resident work group
  double* scr_cls[RCLASSES] = {nullptr, nullptr, nullptr, nullptr};   // the part of each class
  hipStream_t cls_stream[RCLASSES] = {nullptr, nullptr, nullptr, nullptr};   // [0] = stream
  hipEvent_t ev_list = nullptr, ev_cls[RCLASSES] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t ev_pan[RCLASSES] = {nullptr, nullptr, nullptr, nullptr};   // side stream passed its panel pairs
  int2* wl = nullptr;                  // recompression work lists (RCLASSES x wl_cap)
  int* wl_count = nullptr;             // nB x RCLASSES x WL_STRIDE list counters
  int64_t wl_cap = 0;
  int2* wlp = nullptr;                 // panel-pair lists (pairs (i, k+1))
  int* wlp_count = nullptr;
  int64_t wlp_cap = 0;
  double* scr_pan[RCLASSES] = {nullptr, nullptr, nullptr, nullptr};   // scratch of the panel launches
  double* scr_help[2] = {nullptr, nullptr};   // scratch of the second 17-32 launches (panel, rest)
  // arena mode (the allocation-free entry point): every device buffer is carved from [arena, arena +
  // arena_cap), one stream (the null stream), no events, no pinned staging, persistent grids of at most
  // grid_cap work groups
  char* arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  bool arena_short = false;
  int grid_cap = 0;
  bool single_stream = false;
  std::vector<double> h_thr_host;
  std::vector<int> h_replay_host;
  int recomp_grid[3][RCLASSES] = {};   // persistent grid of each recompression class, per arithmetic (PM_*)
  int recomp_grid_max[3][RCLASSES] = {};   // resident capacity (the grids above never exceed it)
  double guard = GUARD_DEFAULT;        // tunables (tlr_ctx_set_tunable)
  int diag_wg = 9;
  int sketch1 = K1;                    // first-round sketch width (TLR_TUNE_SKETCH)
  int nproduct = NPRODUCT;             // products of a cold compression (TLR_TUNE_PRODUCTS)
  int2* wlr = nullptr;                 // redo lists (panel x level) and their counters
  int* wlr_count = nullptr;
  int* rep_rpl = nullptr;              // max_batch records of recompression arithmetic (nlog each)
  int* d_retry = nullptr;              // per slot: tiles redone with the wider sketch
  double* d_thr = nullptr;             // per slot (thr_rel, thr_abs) of the current batch
  MaternParams* d_mp = nullptr;        // per slot Matern parameters of the current batch (read by the kernels)
  MaternParams* h_mp = nullptr;        // pinned staging of d_mp
  std::vector<MaternParams> h_mp_host;
  double* h_thr = nullptr;             // pinned staging of d_thr
  int* d_replay = nullptr;             // per slot: record whose ranks are replayed, or -1
  int* h_replay = nullptr;             // pinned staging of d_replay
  int64_t nlog = 0;                    // recompressions per evaluation (rank-log length)
  int64_t* d_rlog_off = nullptr;       // nB offsets into a rank log
  int* rep_rank0 = nullptr;            // max_batch records of initial ranks (ntile each)
  int* rep_rlog = nullptr;             // max_batch records of recompression ranks (nlog each)
  // initially compressed factors of the replay records (compact: the first rank0 columns of each tile),
  // reused by perturbations that leave the off-diagonal tiles unchanged (eta)
  double* stash = nullptr;
  int64_t stash_cap = 0;               // doubles
  int64_t* d_stash_off = nullptr;      // max_batch x ntile column offsets into the stash
  // prediction workspace, grown on demand: N x m cross covariances, alpha, prediction sites, reductions
  double* pred_buf = nullptr;
  int64_t pred_cap = 0;
  // timing mode: launches of kernel class timing_kernel are bracketed by events
  int timing_kernel = 0;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<TimedLaunch> timed;
  double timed_ms = 0.0, timed_flops = 0.0, timed_bytes = 0.0;
  int64_t timed_launches = 0;
  ~tlr_ctx() {
    drop_graphs();
    if (pred_buf) (void)hipFree(pred_buf);
    if (stash) (void)hipFree(stash);
    if (h_thr && !arena) (void)hipHostFree(h_thr);
    if (h_mp && !arena) (void)hipHostFree(h_mp);
    if (h_replay && !arena) (void)hipHostFree(h_replay);
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    for (void* p : allocs) (void)hipFree(p);
    if (ev_list) (void)hipEventDestroy(ev_list);
    for (int q = 1; q < RCLASSES; ++q) {
      if (ev_cls[q]) (void)hipEventDestroy(ev_cls[q]);
      if (ev_pan[q]) (void)hipEventDestroy(ev_pan[q]);
      if (cls_stream[q]) (void)hipStreamDestroy(cls_stream[q]);
    }
    if (stream) (void)hipStreamDestroy(stream);
  }
  hipEvent_t next_event() {
    if (ev_used == ev_pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      ev_pool.push_back(e);
    }
    return ev_pool[ev_used++];
  }
  // Start / stop a timed launch of class id (no-op unless timing that class).
  bool tic(int id, TimedLaunch& t) {
    if (timing_kernel != id) return false;
    t.a = next_event(); t.b = next_event();
    return t.a && t.b && hipEventRecord(t.a, stream) == hipSuccess;
  }
  void toc(bool on, TimedLaunch& t) {
    if (!on) return;
    if (hipEventRecord(t.b, stream) == hipSuccess) timed.push_back(t);
  }
  template <class T>
  hipError_t alloc(T** p, size_t count) {
    if (arena) {   // carve, 256-byte aligned
      const size_t bytes = count * sizeof(T) > 0 ? count * sizeof(T) : 16;
      const size_t off = (arena_used + 255) & ~(size_t)255;
      if (off + bytes > arena_cap) {
        *p = nullptr;
        arena_short = true;
        return hipErrorOutOfMemory;
      }
      *p = (T*)(arena + off);
      arena_used = off + bytes;
      return hipSuccess;
    }
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T) > 0 ? count * sizeof(T) : 16);
    if (e == hipSuccess) allocs.push_back(q);
    *p = (T*)q;
    return e;
  }
  SlotArgs args() const {
    SlotArgs a;
    a.n = n; a.b = b; a.nB = nB; a.npan = npan; a.rank_cap = RANK_CAP;
    a.fstride = (int64_t)b * RANK_CAP;
    a.scr_stride = scr_stride;
    a.scr_pool = scr_pool;
    a.wl = wl;
    a.wl_count = wl_count;
    a.wl_cap = wl_cap;
    a.wlp = wlp;
    a.wlp_count = wlp_count;
    a.wlp_cap = wlp_cap;
    a.wlr = wlr;
    a.wlr_count = wlr_count;
    a.wlr_cap = wl_cap;
    a.prec_mode = PM_FP64;
    a.guard = guard;
    a.diag_wg = diag_wg;
    a.rep_rpl = rep_rpl;
    a.rule_fixed = 0;
    a.mixed = 0;
    a.thr = d_thr;
    a.replay = d_replay;
    a.rep_rlog = rep_rlog;
    a.rlog_off = d_rlog_off;
    a.nlog = nlog;
    a.slots = d_slots;
    return a;
  }
};

// Deterministic standard normal test matrix (splitmix64 + Box-Muller), column-major b x k.
static void gaussian_matrix(std::vector<double>& out, int rows, int cols, uint64_t seed) {
  out.resize((size_t)rows * cols);
  uint64_t s = seed;
  auto next = [&]() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  for (size_t e = 0; e < out.size(); e += 2) {
    const double u1 = next(), u2 = next();
    const double r = sqrt(-2.0 * log(u1)), a = 6.283185307179586 * u2;
    out[e] = r * cos(a);
    if (e + 1 < out.size()) out[e + 1] = r * sin(a);
  }
}

extern "C" {

const char* tlr_last_error(void) { return g_err.c_str(); }

int tlr_device_info(char* buf, int32_t len) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_err(TLR_ERR_NODEVICE, "no HIP device");
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, 0));
  snprintf(buf, (size_t)len, "%s (%s), %d CUs, %.1f GiB", prop.name, prop.gcnArchName, prop.multiProcessorCount,
           prop.totalGlobalMem / 1073741824.0);
  return TLR_OK;
}

}  // extern "C"

// Fill a fresh context (its arena fields set for arena mode); on failure the caller destroys it.
static int ctx_build(tlr_ctx* c, const double* S, const double* y, int64_t n, int32_t b, int32_t max_batch) {
  if (!S || !y || n < 2 || b < 2 || max_batch < 1) return set_err(TLR_ERR_ARG, "invalid argument");
  if (b > 512) return set_err(TLR_ERR_ARG, "tile size b > 512 is not supported");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return set_err(TLR_ERR_NODEVICE, "no HIP device");
  auto fail = [&](int code) { return code; };
  c->n = n; c->b = b;
  c->nB = (int)((n + b - 1) / b);
  c->npan = (b + 31) / 32;
  c->max_batch = max_batch;
  c->ntile = (int64_t)c->nB * (c->nB - 1) / 2;
  const int64_t m = c->nB - 1;
  c->npairs_max = m * (m - 1) / 2;
  c->kcap = K2 < b ? K2 : b;
  // pair indices of one step and the tile count stay within int range up to nB = 46,341
  if (c->nB > 46340) return fail(set_err(TLR_ERR_ARG, "more than 46340 tiles per side"));
  hipError_t e;
#define ACHK(expr)                                                                           \
  do {                                                                                       \
    e = (expr);                                                                              \
    if (e != hipSuccess) return fail(set_err(TLR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e))); \
  } while (0)
  if (c->arena) {
    c->stream = nullptr;   // the null stream
    c->single_stream = true;
  } else {
    ACHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  }
  // sites as structure of arrays
  std::vector<double> hx(n), hy(n);
  for (int64_t i = 0; i < n; ++i) { hx[i] = S[2 * i]; hy[i] = S[2 * i + 1]; }
  ACHK(c->alloc(&c->sx, n));
  ACHK(c->alloc(&c->sy, n));
  ACHK(c->alloc(&c->y, n));
  ACHK(hipMemcpy(c->sx, hx.data(), n * sizeof(double), hipMemcpyHostToDevice));
  ACHK(hipMemcpy(c->sy, hy.data(), n * sizeof(double), hipMemcpyHostToDevice));
  ACHK(hipMemcpy(c->y, y, n * sizeof(double), hipMemcpyHostToDevice));
  std::vector<double> om;
  gaussian_matrix(om, b, c->kcap, 0x5EED0F7A11ull);
  ACHK(c->alloc(&c->omega, om.size()));
  ACHK(hipMemcpy(c->omega, om.data(), om.size() * sizeof(double), hipMemcpyHostToDevice));
  const int64_t nt = c->ntile > 0 ? c->ntile : 1;
  const int nbb = (b + 63) / 64;
  // one compression workspace per slot that can be compressed concurrently (at most NCS)
  c->nset = max_batch < NCS ? max_batch : NCS;
  for (int q = 0; q < c->nset; ++q) {
    ACHK(c->alloc(&c->bufA[q], (size_t)nt * b * c->kcap));
    ACHK(c->alloc(&c->bufB[q], (size_t)nt * b * c->kcap));
    ACHK(c->alloc(&c->bufR[q], (size_t)nt * c->kcap * c->kcap));
    ACHK(c->alloc(&c->part[q], (size_t)nt * nbb * nbb * 2));
    ACHK(c->alloc(&c->need[q], (size_t)nt));
  }
  // recompression: a persistent grid per width class sized to the resident capacity of the device; each
  // resident work group owns scratch for Q2v (b x 64), the right blocks of R_u, R_v (128 x 64 each) and
  // the Jacobi core and rotations (128 x 128 each) when they do not fit in LDS, W and the projection
  // coefficients of the widest class (64 x 64 each)
  c->scr_stride = (int64_t)b * 64 + 2 * WCAP * 64 + 2 * WCAP * WCAP + 2 * 64 * 64;
  {
    int dev = 0, cus = 0, per_cu[RCLASSES];
    ACHK(hipGetDevice(&dev));
    ACHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    int gmax[RCLASSES] = {0, 0, 0, 0};   // the largest grid of each class over the arithmetic modes
    for (int pm = 0; pm < 3; ++pm) {
      ACHK(recompress_occupancy(b, pm, per_cu));
      for (int q = 0; q < RCLASSES; ++q) {
        c->recomp_grid[pm][q] = (per_cu[q] > 0 ? per_cu[q] : 1) * (cus > 0 ? cus : 1);
        if (c->grid_cap > 0 && c->recomp_grid[pm][q] > c->grid_cap) c->recomp_grid[pm][q] = c->grid_cap;
        c->recomp_grid_max[pm][q] = c->recomp_grid[pm][q];
        gmax[q] = c->recomp_grid[pm][q] > gmax[q] ? c->recomp_grid[pm][q] : gmax[q];
      }
    }
    int64_t gsum = 0;
    for (int q = 0; q < RCLASSES; ++q) gsum += gmax[q];
    // the classes of one step run concurrently (one stream each) and so do the panel launches, so each
    // class has its own scratch twice over, and the 17-32 class, launched on two streams, four times
    ACHK(c->alloc(&c->scr_pool, (size_t)(2 * gsum + 2 * gmax[2]) * c->scr_stride));
    int64_t off = 0;
    for (int q = 0; q < RCLASSES; ++q) {
      c->scr_cls[q] = c->scr_pool + off * c->scr_stride;
      off += gmax[q];
      c->scr_pan[q] = c->scr_pool + off * c->scr_stride;
      off += gmax[q];
    }
    for (int h = 0; h < 2; ++h) {
      c->scr_help[h] = c->scr_pool + off * c->scr_stride;
      off += gmax[2];
    }
    c->cls_stream[0] = c->stream;
    if (c->single_stream) {
      c->cls_stream[1] = c->cls_stream[2] = c->stream;
    } else {
      for (int q = 1; q <= 2; ++q) {
        ACHK(hipStreamCreateWithFlags(&c->cls_stream[q], hipStreamNonBlocking));
        ACHK(hipEventCreateWithFlags(&c->ev_cls[q], hipEventDisableTiming));
        ACHK(hipEventCreateWithFlags(&c->ev_pan[q], hipEventDisableTiming));
      }
      ACHK(hipEventCreateWithFlags(&c->ev_list, hipEventDisableTiming));
    }
    const int64_t npm = c->npairs_max > 0 ? c->npairs_max : 1;
    c->wl_cap = npm * max_batch;
    ACHK(c->alloc(&c->wl, (size_t)RCLASSES * c->wl_cap));
    ACHK(c->alloc(&c->wl_count, (size_t)WL_STRIDE * RCLASSES * c->nB));
    c->wlp_cap = (int64_t)(c->nB > 0 ? c->nB : 1) * max_batch;
    ACHK(c->alloc(&c->wlp, (size_t)RCLASSES * c->wlp_cap));
    ACHK(c->alloc(&c->wlp_count, (size_t)WL_STRIDE * RCLASSES * c->nB));
    // redo and reduced-first lists of the reduced-precision modes: (panel, level) lists of wl_cap entries,
    // counters per step
    ACHK(c->alloc(&c->wlr, (size_t)2 * REDO_LEVELS * c->wl_cap));
    ACHK(c->alloc(&c->wlr_count, (size_t)WL_STRIDE * 2 * REDO_LEVELS * c->nB));
  }
  c->slots.resize(max_batch);
  c->sstate.assign(max_batch, SlotState());
  for (int s = 0; s < max_batch; ++s) {
    SlotPtrs& p = c->slots[s].p;
    ACHK(c->alloc(&p.D, (size_t)c->nB * b * b));
    ACHK(c->alloc(&p.Dinv, (size_t)c->nB * c->npan * 32 * 32));
    ACHK(c->alloc(&p.Linv, (size_t)b * b));
    ACHK(c->alloc(&p.U, (size_t)nt * b * RANK_CAP));
    ACHK(c->alloc(&p.V, (size_t)nt * b * RANK_CAP));
    ACHK(c->alloc(&p.rank, (size_t)nt));
    ACHK(c->alloc(&p.rank0, (size_t)nt));
    ACHK(c->alloc(&p.Wd, (size_t)c->nB * RANK_CAP * RANK_CAP));
    ACHK(c->alloc(&p.Xd, (size_t)c->nB * b * RANK_CAP));
    ACHK(c->alloc(&p.rhs, (size_t)n));
    ACHK(c->alloc(&p.logdet_part, (size_t)c->nB));
    ACHK(c->alloc(&p.status, (size_t)ST_COUNT));
    ACHK(c->alloc(&p.result, (size_t)4));
    ACHK(c->alloc(&p.prec, (size_t)nt));
    ACHK(c->alloc(&p.tnorm, (size_t)nt));
  }
  // rank logs: offsets of each step's recompressions, one log per slot and max_batch replay records
  {
    std::vector<int64_t> off(c->nB);
    int64_t acc = 0;
    for (int k = 0; k < c->nB; ++k) {
      off[k] = acc;
      const int64_t mk = c->nB - k - 1;
      acc += mk * (mk - 1) / 2;
    }
    c->nlog = acc > 0 ? acc : 1;
    ACHK(c->alloc(&c->d_rlog_off, (size_t)c->nB));
    ACHK(hipMemcpy(c->d_rlog_off, off.data(), c->nB * sizeof(int64_t), hipMemcpyHostToDevice));
  }
  for (int s = 0; s < max_batch; ++s) {
    ACHK(c->alloc(&c->slots[s].p.rlog, (size_t)c->nlog));
    ACHK(c->alloc(&c->slots[s].p.rpl, (size_t)c->nlog));
  }
  ACHK(c->alloc(&c->rep_rank0, (size_t)max_batch * nt));
  ACHK(c->alloc(&c->rep_rlog, (size_t)max_batch * c->nlog));
  ACHK(c->alloc(&c->rep_rpl, (size_t)max_batch * c->nlog));
  ACHK(c->alloc(&c->d_stash_off, (size_t)max_batch * nt));
  ACHK(c->alloc(&c->d_replay, (size_t)max_batch));
  ACHK(c->alloc(&c->d_slots, (size_t)max_batch));
  ACHK(c->alloc(&c->d_gst, (size_t)max_batch * ST_COUNT));
  ACHK(c->alloc(&c->d_gres, (size_t)max_batch * 4));
  ACHK(c->alloc(&c->d_grk, (size_t)max_batch * 2 * (nt > 0 ? nt : 1)));
  ACHK(c->alloc(&c->d_retry, (size_t)max_batch));
  ACHK(c->alloc(&c->d_thr, (size_t)2 * max_batch));
  ACHK(c->alloc(&c->d_mp, (size_t)max_batch));
  if (c->arena) {   // pageable staging (the copies on the null stream complete before they return)
    c->h_replay_host.assign(max_batch, -1);
    c->h_thr_host.assign(2 * (size_t)max_batch, 0.0);
    c->h_mp_host.assign(max_batch, MaternParams{0.0, 0.0, 0.0});
    c->h_replay = c->h_replay_host.data();
    c->h_thr = c->h_thr_host.data();
    c->h_mp = c->h_mp_host.data();
  } else {
    ACHK(hipHostMalloc((void**)&c->h_replay, sizeof(int) * max_batch, hipHostMallocDefault));
    ACHK(hipHostMalloc((void**)&c->h_thr, sizeof(double) * 2 * max_batch, hipHostMallocDefault));
    ACHK(hipHostMalloc((void**)&c->h_mp, sizeof(MaternParams) * max_batch, hipHostMallocDefault));
  }
  std::vector<SlotPtrs> hp(max_batch);
  for (int s = 0; s < max_batch; ++s) hp[s] = c->slots[s].p;
  ACHK(hipMemcpy(c->d_slots, hp.data(), max_batch * sizeof(SlotPtrs), hipMemcpyHostToDevice));
#undef ACHK
  return TLR_OK;
}

extern "C" {

int tlr_ctx_create(tlr_ctx** out, const double* S, const double* y, int64_t n, int32_t b, int32_t max_batch) {
  if (!out) return set_err(TLR_ERR_ARG, "invalid argument");
  tlr_ctx* c = new tlr_ctx();
  const int rc = ctx_build(c, S, y, n, b, max_batch);
  if (rc != TLR_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return TLR_OK;
}

int tlr_ctx_destroy(tlr_ctx* ctx) {
  delete ctx;
  return TLR_OK;
}

int tlr_ctx_dims(const tlr_ctx* ctx, int64_t* n, int32_t* b, int32_t* n_tiles) {
  if (!ctx) return set_err(TLR_ERR_ARG, "null context");
  if (n) *n = ctx->n;
  if (b) *b = ctx->b;
  if (n_tiles) *n_tiles = ctx->nB;
  return TLR_OK;
}

int tlr_ctx_set_y(tlr_ctx* ctx, const double* y) {
  if (!ctx || !y) return set_err(TLR_ERR_ARG, "null argument");
  HIPCHK(hipMemcpy(ctx->y, y, ctx->n * sizeof(double), hipMemcpyHostToDevice));
  return TLR_OK;
}

// warm >= 0 (a replay whose ranks are given, forced != nullptr): the sketch starts from the right factors of
// stash record `warm` (the recorded evaluation at a nearby theta), and two products (T X, T^T Q) replace the
// four of a cold start.
// cs: the compression stream (and workspace set) this slot uses.  k1: width of the first round (0: the
// context's tunable; a warm start takes at least K_WARM columns: with two products instead of four its sketch
// needs the wider margin, or the error left beside the truncation, smooth in theta, enters the gradient's
// difference quotient).
static int compress_slot(tlr_ctx* c, int slot, const SlotPtrs& p, int rule_fixed, const int* forced, int mixed,
                         int warm, int cs, int k1 = 0) {
  const MaternParams* mp = c->d_mp + slot;
  const double* thr2 = c->d_thr + 2 * slot;
  hipStream_t s = c->cls_stream[cs];
  double* const bufA = c->bufA[cs];
  double* const bufB = c->bufB[cs];
  double* const bufR = c->bufR[cs];
  double* const part = c->part[cs];
  int* const need = c->need[cs];
  const int b = c->b;
  const int64_t nt = c->ntile;
  HIPCHK(launch_gen_diag(c->sx, c->sy, c->n, b, c->nB, mp, p.D, s));
  if (nt == 0) return TLR_OK;
  HIPCHK(launch_fill_int(need, nt, 1, s));
  const int64_t ms = (int64_t)b * c->kcap, rs = (int64_t)c->kcap * c->kcap;
  const int kmax = c->kcap;
  if (k1 <= 0) k1 = (warm >= 0 && forced && c->sketch1 < K_WARM) ? K_WARM : c->sketch1;
  const int widths[2] = {k1 < kmax ? k1 : kmax, kmax};
  const int rounds = widths[0] == widths[1] ? 1 : 2;
  for (int rd = 0; rd < rounds; ++rd) {
    const int k = widths[rd];
    // alternate products: the buffer holding the latest product and the orthonormal basis of the other side
    double* cur = bufA;   // receives T * X
    double* oth = bufB;   // receives T^T * X
    const double gemm_flops = 2.0 * b * (double)b * k;     // one product with a b x b tile
    const double qr_flops = 4.0 * b * (double)k * k;        // Householder QR + explicit Q
    // bytes per tile: a product reads the b x k operand and the tile's 2 x 2b coordinates and writes the
    // b x k result (the tile itself is regenerated, never stored); the residual reads two b x k panels; the
    // QR reads and writes the panel
    const double gemm_bytes = 16.0 * b * k + 32.0 * b, qr_bytes = 16.0 * b * (double)k;
    TimedLaunch tl{nullptr, nullptr, slot, rd, gemm_flops, gemm_bytes};
    const bool warm_start = warm >= 0 && forced;
    if (warm_start)
      HIPCHK(launch_warm_sketch(c->stash, c->d_stash_off + (int64_t)warm * nt, c->rep_rank0 + (int64_t)warm * nt,
                                c->omega, oth, ms, b, k, need, nt, s));
    bool on = c->tic(1, tl);
    if (warm_start)
      HIPCHK(launch_matern_product(c->sx, c->sy, c->n, b, nt, mp, 0, oth, ms, cur, ms, k, need, s));
    else
      HIPCHK(launch_matern_product(c->sx, c->sy, c->n, b, nt, mp, 0, c->omega, 0, cur, ms, k, need, s));
    c->toc(on, tl);
    int last_op = 0;
    const int nproduct = warm_start ? 2 : c->nproduct;
    for (int pidx = 1; pidx < nproduct; ++pidx) {
      double* src = (last_op == 0) ? cur : oth;
      double* dst = (last_op == 0) ? oth : cur;
      TimedLaunch tq{nullptr, nullptr, slot, rd, qr_flops, qr_bytes};
      on = c->tic(3, tq);
      HIPCHK(launch_panel_qr(src, ms, b, k, nullptr, 0, 0, need, nt, s));
      c->toc(on, tq);
      last_op ^= 1;
      on = c->tic(1, tl);
      HIPCHK(launch_matern_product(c->sx, c->sy, c->n, b, nt, mp, last_op, src, ms, dst, ms, k, need, s));
      c->toc(on, tl);
    }
    // last product M; the other buffer holds the orthonormal basis Qprev
    double* M = (last_op == 0) ? cur : oth;
    double* Qprev = (last_op == 0) ? oth : cur;
    // T ~ M Qprev^T (last op T) or Qprev M^T (last op T^T); residual before M is orthonormalised
    const double* Arow = (last_op == 0) ? M : Qprev;
    const double* Bcol = (last_op == 0) ? Qprev : M;
    if (!forced) {   // the residual only serves the rank choice; a replay's ranks are given
      TimedLaunch tr{nullptr, nullptr, slot, rd, gemm_flops, gemm_bytes};
      on = c->tic(2, tr);
      HIPCHK(launch_residual(c->sx, c->sy, c->n, b, nt, mp, Arow, Bcol, ms, k, part, need, s));
      c->toc(on, tr);
    }
    TimedLaunch tq{nullptr, nullptr, slot, rd, qr_flops, qr_bytes};
    on = c->tic(3, tq);
    HIPCHK(launch_panel_qr(M, ms, b, k, bufR, rs, last_op == 1 ? 1 : 0, need, nt, s));
    c->toc(on, tq);
    double* Qrow = (last_op == 0) ? M : Qprev;   // also the mixed rule's Gram workspace
    const double* Qcol = (last_op == 0) ? Qprev : M;
    HIPCHK(launch_rank_decide(bufR, rs, Qrow, Qcol, ms, part, b, k, kmax, PMIN, rule_fixed, thr2,
                              p.U, p.V, (int64_t)b * RANK_CAP, RANK_CAP, p.rank, need, c->d_retry + slot,
                              p.status, forced, mixed, p.tnorm, nt, s));
  }
  HIPCHK(launch_copy_int(p.rank0, p.rank, nt, s));
  return TLR_OK;
}

}  // extern "C"

// Step k after its TRSM and diagonal update, with one step of lookahead.  The work lists are built on the
// main stream; each class runs on its stream (the context has four hardware queues and the null stream
// takes one: the narrow class on the main stream, 17-32 on a second, 33-64 and then 9-16 on a third, the
// widest first so that the longest pairs start at once), first over its panel pairs (i, k+1),
// which update the next panel column, then over the rest.  POTRF of step k+1 needs only D_{k+1}, final
// after the diagonal update of step k, so the main stream runs it (and the explicit inverse of L_{k+1,k+1})
// first, while the side streams start their pairs; TRSM and the diagonal update of step k+1 follow once
// every stream has passed its panel pairs, and all
// streams are joined into the main stream.  Nothing the lookahead writes (D, rhs, the (., k+1) factors, W_i
// and U_ik W_i of step k+1) is touched by the pairs (i, j), j > k+1, that may still be running, and the
// lists of step k are complete before POTRF can record a failure.
// Redo launches of step k on the main stream (panel != 0: the panel pairs' lists): FP32 first under BF16
// (it may hand pairs on to FP64), then FP64.
static hipError_t launch_redo(tlr_ctx* c, const SlotArgs& sa, int k, int panel) {
  hipError_t e;
  double* scr = panel ? c->scr_pan[3] : c->scr_cls[3];
  if (sa.prec_mode == PM_BF16 &&
      (e = launch_recomp_class(sa, k, 3, panel, c->recomp_grid[PM_FP32][3], scr, c->stream, PM_FP32, 2)) != hipSuccess)
    return e;
  return launch_recomp_class(sa, k, 3, panel, c->recomp_grid[PM_FP64][3], scr, c->stream, PM_FP64, 1);
}

static hipError_t launch_step_lookahead(tlr_ctx* c, const SlotArgs& sa, int k, int nslot) {
  hipError_t e;
  TimedLaunch tt{nullptr, nullptr, -1, k, 0.0, 0.0};
  const bool on = c->tic(5, tt);
  // one stream (arena mode): every launch below runs in the order written, which respects each dependency
  const bool ev = !c->single_stream;
  if ((e = launch_recomp_list(sa, k, nslot, c->stream)) != hipSuccess) return e;
  if (ev && (e = hipEventRecord(c->ev_list, c->stream)) != hipSuccess) return e;
  // classes run by side streams 1 and 2, in order; the third entry of stream 2 is a second launch over the
  // 17-32 lists (claims are atomic, so two launches share a list), which takes what stream 1 has not
  // started by the time stream 2 is done with its own classes
  // In the reduced-precision modes the class lists hold the pairs routed to FP64 and run exactly as in plain
  // FP64; stream 2 then runs the reduced-first list of each pass in the working precision on the widest
  // class (after that class's own launch of the pass, whose scratch it reuses).
  static const int plan[2][3] = {{2, -1, -1}, {3, 1, 2}};
  const int pm = sa.prec_mode;
  const int* grid = c->recomp_grid[PM_FP64];
  for (int q = 1; q <= 2; ++q) {
    hipStream_t sq = c->cls_stream[q];
    if (ev && (e = hipStreamWaitEvent(sq, c->ev_list, 0)) != hipSuccess) return e;
    for (int pass = 1; pass >= 0; --pass) {   // panel pairs first, then the rest
      for (int u = 0; u < 3; ++u) {
        const int cls = plan[q - 1][u];
        if (cls < 0) continue;
        double* scr = u == 2 ? c->scr_help[pass] : (pass ? c->scr_pan[cls] : c->scr_cls[cls]);
        if ((e = launch_recomp_class(sa, k, cls, pass, grid[cls], scr, sq, PM_FP64)) != hipSuccess) return e;
      }
      if (q == 2 && pm != PM_FP64 &&
          (e = launch_recomp_class(sa, k, 3, pass, c->recomp_grid[pm][3], pass ? c->scr_pan[3] : c->scr_cls[3], sq,
                                   pm, 3)) != hipSuccess)
        return e;
      if (ev && pass == 1 && (e = hipEventRecord(c->ev_pan[q], sq)) != hipSuccess) return e;
    }
    if (ev && (e = hipEventRecord(c->ev_cls[q], sq)) != hipSuccess) return e;
  }
  if ((e = launch_potrf(sa, k + 1, nslot, c->stream)) != hipSuccess) return e;
  if ((e = launch_trtri(sa, k + 1, nslot, c->stream)) != hipSuccess) return e;
  if ((e = launch_recomp_class(sa, k, 0, 1, grid[0], c->scr_pan[0], c->stream, PM_FP64)) != hipSuccess) return e;
  if ((e = launch_recomp_class(sa, k, 0, 0, grid[0], c->scr_cls[0], c->stream, PM_FP64)) != hipSuccess) return e;
  for (int q = 1; q <= 2 && ev; ++q)
    if ((e = hipStreamWaitEvent(c->stream, c->ev_pan[q], 0)) != hipSuccess) return e;
  // reduced precision: the pairs the guard handed on, once every launch that could add to their lists is
  // done (BF16 -> FP32 -> FP64); the widest class takes any pair, on the scratch of the finished panel (or
  // rest) launches of that class (the reduced-first launch of stream 2 among them)
  if (pm != PM_FP64 && (e = launch_redo(c, sa, k, 1)) != hipSuccess) return e;
  if ((e = launch_trsm(sa, k + 1, nslot, c->stream)) != hipSuccess) return e;
  if ((e = launch_diag_update(sa, k + 1, nslot, c->stream)) != hipSuccess) return e;
  for (int q = 1; q <= 2 && ev; ++q)
    if ((e = hipStreamWaitEvent(c->stream, c->ev_cls[q], 0)) != hipSuccess) return e;
  if (pm != PM_FP64 && (e = launch_redo(c, sa, k, 0)) != hipSuccess) return e;
  c->toc(on, tt);
  return hipSuccess;
}

// Algorithmic work of the factorization steps that run with lookahead (k = 0 .. nB - 3, the launch group
// that timing class 5 brackets), replayed on the host from the initial ranks and the rank log.  Per step k:
//  * recompressions: pair (i, j), k < j < i, with r_ik, r_jk > 0 costs 4 b r_ik r_jk (W and U_ik W) +
//    8 b w^2 (QR of both concatenations) + 22 w^3 (core SVD) + 4 b w r' (rebuilt factors) flops, w = r_ij + r_jk,
//    r' the rank it was truncated to, and reads U, V of the tiles (i, j), (i, k), (j, k) and writes the new
//    (i, j) pair: 2 b (r_ij + r_ik + r_jk + r') doubles;
//  * the lookahead of step k + 1: POTRF b^3 / 3 and the explicit inverse b^3 / 3 flops (reading and writing
//    the tile: 2 b^2 doubles each); TRSM b^2 r per tile (i, k+1) (V read and written: 2 b r doubles); the
//    diagonal update 2 b r^2 + 2 b^2 r flops per tile (the lower triangle read and written: b (b + 1)
//    doubles).
// Steps from failed_at on are skipped (they did not run).
static void step_work(const tlr_ctx* c, const std::vector<int>& rank0, const std::vector<int>& rfin,
                      const std::vector<int>& rlog, int failed_at, double* flops, double* bytes) {
  std::vector<int> cur(rank0);
  const double b = (double)c->b;
  double fl = 0.0, by = 0.0;
  int64_t off = 0;
  const int kend = failed_at >= 0 && failed_at < c->nB ? failed_at - 1 : c->nB;
  for (int k = 0; k < kend && c->nB - k - 1 >= 2; ++k) {
    const int mk = c->nB - k - 1;
    int64_t q = 0;
    for (int a = 1; a < mk; ++a)
      for (int bb = 0; bb < a; ++bb, ++q) {
        const int i = a + k + 1, j = bb + k + 1;
        const int rik = rfin[tri_index(i, k)], rjk = rfin[tri_index(j, k)];
        if (rik == 0 || rjk == 0) continue;
        const int64_t tij = tri_index(i, j);
        const double w = (double)(cur[tij] + rjk), rn = (double)rlog[off + q];
        fl += 4.0 * b * rik * rjk + 8.0 * b * w * w + 22.0 * w * w * w + 4.0 * b * w * rn;
        by += 16.0 * b * (cur[tij] + rik + rjk + rn);
        cur[tij] = rlog[off + q];
      }
    off += (int64_t)mk * (mk - 1) / 2;
    const int k1 = k + 1;
    fl += 2.0 * b * b * b / 3.0;
    by += 4.0 * 8.0 * b * b;
    for (int i = k1 + 1; i < c->nB; ++i) {
      const double r = (double)rfin[tri_index(i, k1)];
      fl += b * b * r + 2.0 * b * r * r + 2.0 * b * b * r;
      by += 16.0 * b * r + (r > 0.0 ? 8.0 * b * (b + 1.0) : 0.0);
    }
  }
  *flops += fl;
  *bytes += by;
}

// Copy the initially compressed factors of slots 0..m-1 into stash records 0..m-1 (compact, offsets from
// each slot's initial ranks).  Called between the compression and the factorization.
static int stash_compressed(tlr_ctx* c, int m) {
  const int64_t nt = c->ntile;
  if (nt == 0) return TLR_OK;
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<int> r0(nt);
  std::vector<int64_t> off((size_t)m * nt);
  int64_t cols = 0;
  for (int s = 0; s < m; ++s) {
    HIPCHK(hipMemcpy(r0.data(), c->slots[s].p.rank0, nt * sizeof(int), hipMemcpyDeviceToHost));
    for (int64_t t = 0; t < nt; ++t) { off[(size_t)s * nt + t] = cols; cols += r0[t]; }
  }
  const int64_t need = 2 * (int64_t)c->b * (cols > 0 ? cols : 1);
  if (need > c->stash_cap) {
    if (c->stash) HIPCHK(hipFree(c->stash));
    c->stash = nullptr;
    c->stash_cap = 0;
    HIPCHK(hipMalloc(&c->stash, need * sizeof(double)));
    c->stash_cap = need;
  }
  HIPCHK(hipMemcpy(c->d_stash_off, off.data(), (size_t)m * nt * sizeof(int64_t), hipMemcpyHostToDevice));
  for (int s = 0; s < m; ++s) {
    const SlotPtrs& p = c->slots[s].p;
    HIPCHK(launch_stash(p.U, p.V, (int64_t)c->b * RANK_CAP, c->b, p.rank0, c->d_stash_off + (int64_t)s * nt, c->stash,
                        p.rank, p.rank0, nt, 0, c->stream));
  }
  return TLR_OK;
}

// One pass over m theta; replay[s] >= 0 makes slot s reuse the ranks of record replay[s] (frozen ranks);
// reuse[s] >= 0 makes it also take record reuse[s]'s stashed compressed factors instead of compressing
// (valid when theta differs from the record's only in eta, which leaves the off-diagonal tiles unchanged);
// stash != 0 stores every slot's compressed factors as records 0..m-1.
// graph: 0 enqueue and run; 1 capture the device work into *exec (nothing runs, no results); 2 run *exec (a
// capture of the same configuration) instead of enqueueing.  The host staging of theta (thresholds, Matern
// parameters, replay records) happens in every mode: a captured graph copies it from the pinned buffers.
static int run_batch_impl(tlr_ctx* c, int32_t m, const double* thetas, int32_t rule, double value,
                          int32_t precision, const int* replay, double* out, int32_t* status, int32_t* failed_at,
                          int32_t* ranks_sigma, int32_t* ranks_factor, const int* reuse, int stash, const int* warm,
                          int graph = 0, hipGraphExec_t* exec = nullptr) {
  if (!c || !thetas || !out || m < 1 || m > c->max_batch) return set_err(TLR_ERR_ARG, "invalid argument");
  if (rule != TLR_RULE_FIXED && rule != TLR_RULE_ADAPTIVE && rule != TLR_RULE_MIXED)
    return set_err(TLR_ERR_ARG, "unknown rule");
  if (rule == TLR_RULE_MIXED && replay) return set_err(TLR_ERR_ARG, "frozen-rank replay is not available for the mixed rule");
  if (precision != TLR_PREC_FP64 && precision != TLR_PREC_FP32 && precision != TLR_PREC_BF16)
    return set_err(TLR_ERR_ARG, "unsupported precision");
  if (rule == TLR_RULE_MIXED && precision != TLR_PREC_FP64)
    return set_err(TLR_ERR_ARG, "the mixed storage rule runs with FP64 arithmetic only");
  for (int s = 0; s < m; ++s) {
    const double* th = thetas + 3 * s;
    if (!(th[0] > 0.0) || !(th[1] > 0.0) || !(th[2] >= 0.0)) return set_err(TLR_ERR_ARG, "theta must be positive");
  }
  if (!(value > 0.0)) return set_err(TLR_ERR_ARG, "rule value must be positive");
  hipStream_t st = c->stream;
  const int rule_fixed = rule == TLR_RULE_FIXED;
  const int mixed = rule == TLR_RULE_MIXED;
  if (graph != 1) {   // a capture runs nothing: the slots keep their factors
    for (int s = 0; s < m; ++s) c->sstate[s].valid = false;
    c->factorizations += m;
  }
  c->ev_used = 0;
  c->timed.clear();
  // per-slot thresholds of the recompression (the adaptive budget depends on theta), Matern parameters and
  // replay records; the previous call ended with a synchronisation, so the pinned staging buffers are free
  const double bmax_ = (double)(c->n < c->b ? c->n : c->b);
  for (int s = 0; s < m; ++s) {
    const double* th = thetas + 3 * s;
    c->h_thr[2 * s] = rule == TLR_RULE_FIXED ? value : 0.0;
    c->h_thr[2 * s + 1] = rule == TLR_RULE_FIXED ? 0.0 : value * bmax_ * th[0] * th[2] / (double)c->n;
    c->h_mp[s].sigma2 = th[0];
    c->h_mp[s].inv_range_sqrt3 = sqrt(3.0) / th[1];
    c->h_mp[s].nugget = th[0] * th[2];
    c->h_replay[s] = replay ? replay[s] : -1;
  }
  auto enqueue = [&]() -> int {
  HIPCHK(launch_fill_int(c->d_retry, c->max_batch, 0, st));
  HIPCHK(launch_fill_int(c->wl_count, (int64_t)WL_STRIDE * RCLASSES * c->nB, 0, st));
  HIPCHK(launch_fill_int(c->wlp_count, (int64_t)WL_STRIDE * RCLASSES * c->nB, 0, st));
  HIPCHK(launch_fill_int(c->wlr_count, (int64_t)WL_STRIDE * 2 * REDO_LEVELS * c->nB, 0, st));
  HIPCHK(hipMemcpyAsync(c->d_thr, c->h_thr, sizeof(double) * 2 * m, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->d_mp, c->h_mp, sizeof(MaternParams) * m, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->d_replay, c->h_replay, sizeof(int) * m, hipMemcpyHostToDevice, st));
  // per-slot state on the main stream first; then the slots are compressed in turn on the main stream and
  // the two side streams (each with its own workspace), which overlap one slot's latency-bound kernels
  // (panel QR, rank decision) and launch tails with another's.  When a compression kernel is being timed
  // (its events are recorded on the main stream) every slot stays on the main stream.
  HIPCHK(launch_slot_init(c->d_slots, m, c->y, c->n, c->ntile, c->nlog, mixed, st));
  int ncs = (c->timing_kernel >= 1 && c->timing_kernel <= 3) ? 1 : c->nset;
  if (ncs > m) ncs = m;
  if (ncs > 1) {
    HIPCHK(hipEventRecord(c->ev_list, st));
    for (int q = 1; q < ncs; ++q) HIPCHK(hipStreamWaitEvent(c->cls_stream[q], c->ev_list, 0));
  }
  int ncomp = 0;   // slots compressed so far (stream of the next: ncomp % ncs)
  for (int s = 0; s < m; ++s) {
    const SlotPtrs& p = c->slots[s].p;
    const int* forced = (replay && replay[s] >= 0) ? c->rep_rank0 + (int64_t)replay[s] * c->ntile : nullptr;
    if (reuse && reuse[s] >= 0) {
      const int rec = reuse[s];
      HIPCHK(launch_gen_diag(c->sx, c->sy, c->n, c->b, c->nB, c->d_mp + s, p.D, st));
      HIPCHK(launch_stash(p.U, p.V, (int64_t)c->b * RANK_CAP, c->b, c->rep_rank0 + (int64_t)rec * c->ntile,
                          c->d_stash_off + (int64_t)rec * c->ntile, c->stash, p.rank, p.rank0, c->ntile, 1, st));
    } else {
      int rc = compress_slot(c, s, p, rule_fixed, forced, mixed, warm ? warm[s] : -1, ncomp % ncs);
      if (rc != TLR_OK) return rc;
      ++ncomp;
    }
  }
  if (ncs > 1)
    for (int q = 1; q < ncs; ++q) {   // join the side streams into the main stream
      HIPCHK(hipEventRecord(c->ev_cls[q], c->cls_stream[q]));
      HIPCHK(hipStreamWaitEvent(st, c->ev_cls[q], 0));
    }
  if (stash) {
    const int rc = stash_compressed(c, m);
    if (rc != TLR_OK) return rc;
  }
  SlotArgs sa = c->args();
  sa.rule_fixed = rule_fixed;
  sa.mixed = mixed;
  sa.prec_mode = precision == TLR_PREC_FP32 ? PM_FP32 : (precision == TLR_PREC_BF16 ? PM_BF16 : PM_FP64);
  HIPCHK(launch_potrf(sa, 0, m, st));
  HIPCHK(launch_trtri(sa, 0, m, st));
  HIPCHK(launch_trsm(sa, 0, m, st));
  HIPCHK(launch_diag_update(sa, 0, m, st));
  for (int k = 0; k + 1 < c->nB; ++k) {
    if (c->nB - k - 1 >= 2) {
      HIPCHK(launch_step_lookahead(c, sa, k, m));   // recompressions of step k with POTRF ... of step k + 1
    } else {
      HIPCHK(launch_potrf(sa, k + 1, m, st));
      HIPCHK(launch_trtri(sa, k + 1, m, st));
      HIPCHK(launch_trsm(sa, k + 1, m, st));
      HIPCHK(launch_diag_update(sa, k + 1, m, st));
    }
  }
  HIPCHK(launch_final(sa, m, st));
  return TLR_OK;
  };   // enqueue
  if (graph == 2) {
    HIPCHK(hipGraphLaunch(*exec, st));
  } else if (graph == 1) {
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc != TLR_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return set_err(TLR_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return set_err(TLR_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    return TLR_OK;
  } else {
    const int rc = enqueue();
    if (rc != TLR_OK) return rc;
  }
  // every slot's status and result (and ranks, if asked for) in one gather, read back with one copy each
  // rather than two to four synchronous copies per slot
  const bool want_ranks = (ranks_sigma || ranks_factor) && c->ntile > 0;
  HIPCHK(launch_slot_gather(c->d_slots, m, c->d_gst, c->d_gres, want_ranks ? c->d_grk : nullptr, c->ntile, st));
  HIPCHK(hipStreamSynchronize(st));
  std::vector<int> gst((size_t)m * ST_COUNT);
  std::vector<double> gres((size_t)m * 4);
  std::vector<int> grk(want_ranks ? (size_t)m * 2 * c->ntile : 0);
  HIPCHK(hipMemcpy(gst.data(), c->d_gst, gst.size() * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(gres.data(), c->d_gres, gres.size() * sizeof(double), hipMemcpyDeviceToHost));
  if (want_ranks) HIPCHK(hipMemcpy(grk.data(), c->d_grk, grk.size() * sizeof(int), hipMemcpyDeviceToHost));
  if (!c->timed.empty()) {
    std::vector<int> retry(c->max_batch);
    HIPCHK(hipMemcpy(retry.data(), c->d_retry, c->max_batch * sizeof(int), hipMemcpyDeviceToHost));
    for (const TimedLaunch& t : c->timed) {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, t.a, t.b));
      c->timed_ms += ms;
      c->timed_launches += 1;
      if (c->timing_kernel != 5) {
        const double tiles = t.round == 0 ? (double)c->ntile : (double)retry[t.slot];
        c->timed_flops += tiles * t.flops_per_tile;
        c->timed_bytes += tiles * t.bytes_per_tile;
      }
    }
    c->timed.clear();
    if (c->timing_kernel == 5 && c->ntile > 0) {   // recompression: flops from the rank logs
      std::vector<int> r0(c->ntile), rf(c->ntile), lg(c->nlog);
      for (int s = 0; s < m; ++s) {
        const SlotPtrs& p = c->slots[s].p;
        int stv[ST_COUNT];
        HIPCHK(hipMemcpy(stv, p.status, sizeof(stv), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(r0.data(), p.rank0, c->ntile * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(rf.data(), p.rank, c->ntile * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(lg.data(), p.rlog, c->nlog * sizeof(int), hipMemcpyDeviceToHost));
        const int fa = stv[ST_FAILED] == 0x7fffffff ? -1 : stv[ST_FAILED];
        step_work(c, r0, rf, lg, fa, &c->timed_flops, &c->timed_bytes);
      }
    }
  }
  int ret = TLR_OK;
  for (int s = 0; s < m; ++s) {
    const int* stv = gst.data() + (size_t)s * ST_COUNT;
    const double* res = gres.data() + (size_t)s * 4;
    if (stv[ST_GUARD]) return set_err(TLR_ERR_HIP, "a kernel guard tripped (rank or core out of range)");
    // a tile that needed more columns than the rank capacity: this theta has no result (the others of the
    // batch do); a one-theta call reports it as an error
    const bool capped = stv[ST_RANKCAP] != 0;
    const int fa = stv[ST_FAILED];
    const bool ok = !capped && fa == 0x7fffffff;
    out[3 * s + 0] = ok ? res[0] : NAN;
    out[3 * s + 1] = ok ? res[1] : NAN;
    out[3 * s + 2] = ok ? res[2] : NAN;
    if (status) status[s] = ok ? TLR_OK : (capped ? TLR_RANK_CAPPED : TLR_NOT_PD);
    if (failed_at) failed_at[s] = ok || capped ? -1 : fa;
    if (!ok && m == 1)
      ret = capped ? set_err(TLR_ERR_RANKCAP, "tile rank exceeded the capacity of 64 columns") : TLR_NOT_PD;
    if (!replay) {   // a plain evaluation: its factor stays resident in the slot until the next evaluation
      SlotState& ss = c->sstate[s];
      ss.valid = !capped;
      memcpy(ss.theta, thetas + 3 * s, sizeof(ss.theta));
      ss.rule = rule;
      ss.value = value;
      ss.precision = precision;
      memcpy(ss.out3, out + 3 * s, sizeof(ss.out3));
      ss.status = ok ? TLR_OK : TLR_NOT_PD;
      ss.failed_at = ok ? -1 : fa;
    }
    const int nB = c->nB;
    for (int which = 0; which < 2; ++which) {
      int32_t* dst = which == 0 ? ranks_sigma : ranks_factor;
      if (!dst) continue;
      int32_t* d = dst + (int64_t)s * nB * nB;
      memset(d, 0, sizeof(int32_t) * nB * nB);
      if (!want_ranks) continue;
      const int* rk = grk.data() + (size_t)(2 * s + which) * c->ntile;
      for (int i = 1; i < nB; ++i)
        for (int j = 0; j < i; ++j) d[i * nB + j] = rk[tri_index(i, j)];
    }
  }
  return ret;
}

// run_batch_impl, and on an error return the side streams and the main stream are drained, so that the next
// call finds no work in flight (it reuses the pinned staging buffers and the work lists).
static int run_batch(tlr_ctx* c, int32_t m, const double* thetas, int32_t rule, double value, int32_t precision,
                     const int* replay, double* out, int32_t* status, int32_t* failed_at, int32_t* ranks_sigma,
                     int32_t* ranks_factor, const int* reuse = nullptr, int stash = 0, const int* warm = nullptr,
                     int graph = 0, hipGraphExec_t* exec = nullptr) {
  const int rc = run_batch_impl(c, m, thetas, rule, value, precision, replay, out, status, failed_at, ranks_sigma,
                                ranks_factor, reuse, stash, warm, graph, exec);
  if (rc < 0 && c) {
    for (int q = 1; q < RCLASSES; ++q)
      if (c->cls_stream[q]) (void)hipStreamSynchronize(c->cls_stream[q]);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
  }
  return rc;
}

int tlr_ctx_set_timing(tlr_ctx* ctx, int32_t kernel_class) {
  if (!ctx || kernel_class < 0 || kernel_class > 5) return set_err(TLR_ERR_ARG, "invalid kernel class");
  ctx->timing_kernel = kernel_class;
  ctx->timed_ms = 0.0;
  ctx->timed_flops = 0.0;
  ctx->timed_bytes = 0.0;
  ctx->timed_launches = 0;
  return TLR_OK;
}

int tlr_ctx_get_timing(const tlr_ctx* ctx, double* total_ms, int64_t* launches, double* flops, double* bytes) {
  if (!ctx) return set_err(TLR_ERR_ARG, "null context");
  if (bytes) *bytes = ctx->timed_bytes;
  if (total_ms) *total_ms = ctx->timed_ms;
  if (launches) *launches = ctx->timed_launches;
  if (flops) *flops = ctx->timed_flops;
  return TLR_OK;
}

int tlr_ctx_precisions(const tlr_ctx* ctx, int32_t slot, int32_t* prec) {
  if (!ctx || !prec || slot < 0 || slot >= ctx->max_batch) return set_err(TLR_ERR_ARG, "invalid argument");
  const int nB = ctx->nB;
  memset(prec, 0, sizeof(int32_t) * nB * nB);
  if (ctx->ntile == 0) return TLR_OK;
  std::vector<int> pv(ctx->ntile);
  HIPCHK(hipMemcpy(pv.data(), ctx->slots[slot].p.prec, ctx->ntile * sizeof(int), hipMemcpyDeviceToHost));
  for (int i = 1; i < nB; ++i)
    for (int j = 0; j < i; ++j) prec[i * nB + j] = pv[tri_index(i, j)];
  return TLR_OK;
}

int tlr_set_error(int code, const char* msg) { return set_err(code, msg); }

// tlr_loglik on a context whose every device buffer is carved from [base, base + bytes) (device memory
// the caller owns), on the null stream, with persistent grids of at most grid_cap work groups; allocates
// no device memory.  Used by the allocation-free entry point.
int tlr_loglik_in_arena(void* base, size_t bytes, int grid_cap, const double* S, const double* y, int64_t n,
                        int32_t b, const double theta[3], int32_t rule, double value, int32_t precision,
                        double out3[3], int32_t* ranks_sigma, int32_t* ranks_factor, int32_t* failed_at) {
  if (!base || bytes == 0 || grid_cap < 1) return set_err(TLR_ERR_ARG, "invalid workspace");
  tlr_ctx c;
  c.arena = (char*)base;
  c.arena_cap = bytes;
  c.grid_cap = grid_cap;
  int rc = ctx_build(&c, S, y, n, b, 1);
  if (rc != TLR_OK) {
    if (c.arena_short)
      return set_err(TLR_ERR_ARG, "N=" + std::to_string(n) + ", b=" + std::to_string(b) + " needs more than the " +
                                      std::to_string(bytes >> 20) + " MiB of the static workspace");
    return rc;
  }
  return tlr_loglik(&c, theta, rule, value, precision, out3, ranks_sigma, ranks_factor, failed_at);
}

int tlr_ctx_set_tunable(tlr_ctx* ctx, int32_t which, double value) {
  if (!ctx) return set_err(TLR_ERR_ARG, "null context");
  ctx->drop_graphs();   // the tunables are baked into captured launches
  switch (which) {
    case TLR_TUNE_GUARD:
      if (!(value >= 1.0 && value <= 1e12)) return set_err(TLR_ERR_ARG, "guard must be in [1, 1e12]");
      ctx->guard = value;
      return TLR_OK;
    case TLR_TUNE_DIAG_WG:
      if (!(value >= 1.0 && value <= 64.0)) return set_err(TLR_ERR_ARG, "diag_wg must be in [1, 64]");
      ctx->diag_wg = (int)value;
      return TLR_OK;
    case TLR_TUNE_GRID_FRAC:
      // persistent recompression grids as a fraction of the resident capacity (scratch is sized for all of it)
      if (!(value > 0.0 && value <= 1.0)) return set_err(TLR_ERR_ARG, "grid fraction must be in (0, 1]");
      for (int pm = 0; pm < 3; ++pm)
        for (int q = 0; q < RCLASSES; ++q) {
          const int g = (int)(value * ctx->recomp_grid_max[pm][q]);
          ctx->recomp_grid[pm][q] = g > 0 ? g : 1;
        }
      return TLR_OK;
    case TLR_TUNE_SKETCH:
      // a multiple of 8 (the register panels' column groups) with room for PMIN spare columns
      if (!(value >= 32.0 && value <= (double)ctx->kcap) || (int)value != value || (int)value % 8 != 0)
        return set_err(TLR_ERR_ARG, "sketch width must be a multiple of 8 in [32, min(96, b)]");
      ctx->sketch1 = (int)value;
      return TLR_OK;
    case TLR_TUNE_PRODUCTS:
      if (!(value >= 2.0 && value <= 6.0) || (int)value != value)
        return set_err(TLR_ERR_ARG, "products must be an integer in [2, 6]");
      ctx->nproduct = (int)value;
      return TLR_OK;
    default:
      return set_err(TLR_ERR_ARG, "unknown tunable");
  }
}

int tlr_ctx_get_tunable(const tlr_ctx* ctx, int32_t which, double* value) {
  if (!ctx || !value) return set_err(TLR_ERR_ARG, "null argument");
  switch (which) {
    case TLR_TUNE_GUARD: *value = ctx->guard; return TLR_OK;
    case TLR_TUNE_DIAG_WG: *value = ctx->diag_wg; return TLR_OK;
    case TLR_TUNE_GRID_FRAC:
      *value = (double)ctx->recomp_grid[0][0] / (ctx->recomp_grid_max[0][0] > 0 ? ctx->recomp_grid_max[0][0] : 1);
      return TLR_OK;
    case TLR_TUNE_SKETCH: *value = ctx->sketch1 < ctx->kcap ? ctx->sketch1 : ctx->kcap; return TLR_OK;
    case TLR_TUNE_PRODUCTS: *value = ctx->nproduct; return TLR_OK;
    default:
      return set_err(TLR_ERR_ARG, "unknown tunable");
  }
}

int tlr_ctx_arith_counts(const tlr_ctx* ctx, int32_t slot, int64_t counts[3]) {
  if (!ctx || !counts || slot < 0 || slot >= ctx->max_batch) return set_err(TLR_ERR_ARG, "invalid argument");
  std::vector<int> pv(ctx->nlog);
  HIPCHK(hipMemcpy(pv.data(), ctx->slots[slot].p.rpl, ctx->nlog * sizeof(int), hipMemcpyDeviceToHost));
  counts[0] = counts[1] = counts[2] = 0;
  for (int v : pv)
    if (v >= 0 && v < 3) ++counts[v];
  return TLR_OK;
}

extern "C" {

int tlr_loglik_batch(tlr_ctx* c, int32_t m, const double* thetas, int32_t rule, double value, int32_t precision,
                     double* out, int32_t* status, int32_t* failed_at, int32_t* ranks_sigma, int32_t* ranks_factor) {
  return run_batch(c, m, thetas, rule, value, precision, nullptr, out, status, failed_at, ranks_sigma, ranks_factor);
}

int tlr_loglik_batch_graph(tlr_ctx* c, int32_t m, const double* thetas, int32_t rule, double value,
                           int32_t precision, double* out, int32_t* status, int32_t* failed_at, int32_t* ranks_sigma,
                           int32_t* ranks_factor) {
  if (!c) return set_err(TLR_ERR_ARG, "null context");
  if (c->single_stream) return set_err(TLR_ERR_ARG, "graph capture needs a context with its own streams");
  if (c->timing_kernel != 0) return set_err(TLR_ERR_ARG, "graph capture with kernel timing on");
  const auto key = std::make_tuple((int)m, (int)rule, (int)precision);
  auto it = c->graphs.find(key);
  if (it != c->graphs.end())
    return run_batch(c, m, thetas, rule, value, precision, nullptr, out, status, failed_at, ranks_sigma, ranks_factor,
                     nullptr, 0, nullptr, 2, &it->second);
  // first call of this configuration: run it as usual (every kernel attribute is set and every code object
  // loaded before a capture), then capture the same device work for the next calls
  const int rc = run_batch(c, m, thetas, rule, value, precision, nullptr, out, status, failed_at, ranks_sigma,
                           ranks_factor);
  if (rc < 0) return rc;
  hipGraphExec_t exec = nullptr;
  const int rc2 = run_batch(c, m, thetas, rule, value, precision, nullptr, out, status, failed_at, nullptr, nullptr,
                            nullptr, 0, nullptr, 1, &exec);
  if (rc2 < 0) return rc2;
  c->graphs[key] = exec;
  return rc;
}

int tlr_loglik_grad_batch(tlr_ctx* c, int32_t m, const double* thetas, int32_t rule, double value, int32_t precision,
                          double h, double* out, double* grad, int32_t* status, int32_t* failed_at,
                          int32_t* ranks_sigma, int32_t* ranks_factor) {
  if (!c || !grad || !(h > 0.0) || h > 0.1) return set_err(TLR_ERR_ARG, "invalid argument");
  if (rule == TLR_RULE_MIXED) return set_err(TLR_ERR_ARG, "the gradient is not available for the mixed rule");
  std::vector<int32_t> st(m > 0 ? m : 1);
  int rc = run_batch(c, m, thetas, rule, value, precision, nullptr, out, st.data(), failed_at, ranks_sigma,
                     ranks_factor, nullptr, 1);
  if (rc < 0) return rc;
  if (status) memcpy(status, st.data(), sizeof(int32_t) * m);
  // keep the base ranks as replay records (record s = slot s)
  HIPCHK(launch_slot_records(c->d_slots, m, c->rep_rank0, c->rep_rlog, c->rep_rpl, c->ntile, c->nlog, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  // d/dlog sigma2 in closed form: tiles and budgets both scale with sigma2, so the ranks are unchanged, L
  // scales by sigma and log L = -(quad / sigma2 + logdet + N log sigma2 + N log 2 pi) / 2 in sigma2
  // d/dlog range, d/dlog eta: central differences at the base ranks; the eta perturbations leave the
  // off-diagonal tiles unchanged, so they reuse the base's compressed factors
  std::vector<double> th4;
  std::vector<int> src, reuse4, warm4;
  for (int s = 0; s < m; ++s) {
    grad[3 * s + 0] = st[s] == TLR_OK ? 0.5 * out[3 * s + 2] - 0.5 * (double)c->n : NAN;
    grad[3 * s + 1] = grad[3 * s + 2] = NAN;
    if (st[s] != TLR_OK) continue;
    for (int par = 1; par <= 2; ++par)
      for (int sg = -1; sg <= 1; sg += 2) {
        const double* t = thetas + 3 * s;
        th4.insert(th4.end(), {t[0], t[1], t[2]});
        th4[th4.size() - 3 + par] *= exp(sg * h);
        src.push_back(s);
        reuse4.push_back(par == 2 ? s : -1);   // eta: the base's compressed factors as they are
        warm4.push_back(par == 1 ? s : -1);    // range: sketch warm-started from the base's factors
      }
  }
  const int n4 = (int)src.size();
  std::vector<double> o4(3 * (n4 > 0 ? n4 : 1));
  std::vector<int32_t> s4(n4 > 0 ? n4 : 1), f4(n4 > 0 ? n4 : 1);
  for (int off = 0; off < n4; off += c->max_batch) {
    const int mm = n4 - off < c->max_batch ? n4 - off : c->max_batch;
    rc = run_batch(c, mm, th4.data() + 3 * off, rule, value, precision, src.data() + off, o4.data() + 3 * off,
                   s4.data() + off, f4.data() + off, nullptr, nullptr, reuse4.data() + off, 0, warm4.data() + off);
    if (rc < 0) return rc;
  }
  for (int e = 0; e < n4; e += 4) {
    const int s = src[e];
    for (int par = 1; par <= 2; ++par) {
      const int ip = e + 2 * (par - 1) + 1, im = e + 2 * (par - 1);
      if (s4[ip] == TLR_OK && s4[im] == TLR_OK) grad[3 * s + par] = (o4[3 * ip] - o4[3 * im]) / (2.0 * h);
    }
  }
  return (m == 1 && st[0] != TLR_OK) ? TLR_NOT_PD : TLR_OK;
}

int tlr_loglik_grad(tlr_ctx* ctx, const double theta[3], int32_t rule, double value, int32_t precision, double h,
                    double out3[3], double grad3[3], int32_t* ranks_sigma, int32_t* ranks_factor, int32_t* failed_at) {
  int32_t st = 0, fa = -1;
  const int rc = tlr_loglik_grad_batch(ctx, 1, theta, rule, value, precision, h, out3, grad3, &st, &fa, ranks_sigma,
                                       ranks_factor);
  if (failed_at) *failed_at = fa;
  return rc;
}

// The factor at theta in slot 0: the resident one when slot 0 holds exactly this evaluation (the factor handle
// of the last tlr_loglik), else a new factorization.  out3 / status receive its result.
static int factor_at(tlr_ctx* c, const double theta[3], int32_t rule, double value, int32_t precision, double out3[3],
                     int32_t* status) {
  if (c->sstate[0].matches(theta, rule, value, precision)) {
    memcpy(out3, c->sstate[0].out3, 3 * sizeof(double));
    *status = c->sstate[0].status;
    return TLR_OK;
  }
  int32_t fa = -1;
  return run_batch(c, 1, theta, rule, value, precision, nullptr, out3, status, &fa, nullptr, nullptr);
}

int tlr_ctx_stats(const tlr_ctx* ctx, int64_t* factorizations) {
  if (!ctx) return set_err(TLR_ERR_ARG, "null context");
  if (factorizations) *factorizations = ctx->factorizations;
  return TLR_OK;
}

int tlr_predict(tlr_ctx* c, const double theta[3], int32_t rule, double value, int32_t precision, const double* Sp,
                int64_t n_pred, double* mu, double* sd, double out3[3]) {
  if (!c || !theta || !Sp || n_pred < 1 || n_pred > 65535 || !mu || !sd) return set_err(TLR_ERR_ARG, "invalid argument");
  double o3[3];
  int32_t st = 0;
  const int rc = factor_at(c, theta, rule, value, precision, o3, &st);
  if (rc < 0) return rc;
  if (out3) memcpy(out3, o3, sizeof(o3));
  if (st != TLR_OK) return TLR_NOT_PD;
  const int64_t n = c->n;
  const int m = (int)n_pred;
  // workspace: Xc (n x m) | alpha (n) | px, py (m) | mu, ss (m)
  const int64_t need = n * m + n + 4 * (int64_t)m;
  if (need > c->pred_cap) {
    if (c->pred_buf) HIPCHK(hipFree(c->pred_buf));
    c->pred_buf = nullptr;
    HIPCHK(hipMalloc(&c->pred_buf, need * sizeof(double)));
    c->pred_cap = need;
  }
  double* Xc = c->pred_buf;
  double* alpha = Xc + n * m;
  double* px = alpha + n;
  double* py = px + m;
  double* dmu = py + m;
  double* dss = dmu + m;
  std::vector<double> hx(m), hy(m);
  for (int j = 0; j < m; ++j) { hx[j] = Sp[2 * j]; hy[j] = Sp[2 * j + 1]; }
  hipStream_t s = c->stream;
  HIPCHK(hipMemcpyAsync(px, hx.data(), m * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(py, hy.data(), m * sizeof(double), hipMemcpyHostToDevice, s));
  SlotArgs sa = c->args();
  // alpha = L^{-T} z, z = L^{-1} y left in the slot by the factorization
  HIPCHK(hipMemcpyAsync(alpha, c->slots[0].p.rhs, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  HIPCHK(launch_solve(sa, 0, alpha, n, 1, 1, s));
  MaternParams mp;
  mp.sigma2 = theta[0];
  mp.inv_range_sqrt3 = sqrt(3.0) / theta[1];
  mp.nugget = 0.0;
  HIPCHK(launch_gen_cross(c->sx, c->sy, n, px, py, m, mp, Xc, n, s));
  HIPCHK(launch_col_dot(Xc, n, n, m, alpha, dmu, s));          // mu = Kp alpha
  HIPCHK(launch_solve(sa, 0, Xc, n, m, 0, s));                   // W = L^{-1} Kp^T
  HIPCHK(launch_col_dot(Xc, n, n, m, nullptr, dss, s));          // column sums of W^2
  HIPCHK(hipStreamSynchronize(s));
  std::vector<double> ss(m);
  HIPCHK(hipMemcpy(mu, dmu, m * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(ss.data(), dss, m * sizeof(double), hipMemcpyDeviceToHost));
  const double v0 = theta[0] * (1.0 + theta[2]);
  for (int j = 0; j < m; ++j) {
    const double v = v0 - ss[j];
    sd[j] = sqrt(v > 1e-12 ? v : 1e-12);
  }
  return TLR_OK;
}

int tlr_solve(tlr_ctx* c, const double theta[3], int32_t rule, double value, int32_t precision, double* B,
              int64_t ldb, int64_t m, int32_t mode, double out3[3]) {
  if (!c || !theta || !B || m < 1 || ldb < c->n || mode < 0 || mode > 2) return set_err(TLR_ERR_ARG, "invalid argument");
  double o3[3];
  int32_t st = 0;
  const int rc = factor_at(c, theta, rule, value, precision, o3, &st);
  if (rc < 0) return rc;
  if (out3) memcpy(out3, o3, sizeof(o3));
  if (st != TLR_OK) return TLR_NOT_PD;
  const int64_t n = c->n;
  const int64_t chunk = m < 4096 ? m : 4096;           // right-hand sides per pass through the device
  if (n * chunk > c->pred_cap) {
    if (c->pred_buf) HIPCHK(hipFree(c->pred_buf));
    c->pred_buf = nullptr;
    HIPCHK(hipMalloc(&c->pred_buf, n * chunk * sizeof(double)));
    c->pred_cap = n * chunk;
  }
  SlotArgs sa = c->args();
  hipStream_t s = c->stream;
  for (int64_t c0 = 0; c0 < m; c0 += chunk) {
    const int mc = (int)(m - c0 < chunk ? m - c0 : chunk);
    HIPCHK(hipMemcpy2DAsync(c->pred_buf, n * sizeof(double), B + c0 * ldb, ldb * sizeof(double), n * sizeof(double),
                            mc, hipMemcpyHostToDevice, s));
    if (mode != 1) HIPCHK(launch_solve(sa, 0, c->pred_buf, n, mc, 0, s));
    if (mode != 0) HIPCHK(launch_solve(sa, 0, c->pred_buf, n, mc, 1, s));
    HIPCHK(hipMemcpy2DAsync(B + c0 * ldb, ldb * sizeof(double), c->pred_buf, n * sizeof(double), n * sizeof(double),
                            mc, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return TLR_OK;
}

int tlr_loglik(tlr_ctx* ctx, const double theta[3], int32_t rule, double value, int32_t precision, double out3[3],
               int32_t* ranks_sigma, int32_t* ranks_factor, int32_t* failed_at) {
  int32_t st = 0, fa = -1;
  const int rc = tlr_loglik_batch(ctx, 1, theta, rule, value, precision, out3, &st, &fa, ranks_sigma, ranks_factor);
  if (failed_at) *failed_at = fa;
  return rc;
}

}  // extern "C"

// ------------------------------------------------------------------ dense analysis of the rule's tiles
// The error terms of the assembled approximation Sigma^TLR (codes/36_likelihood_error_terms.py:75-136) need
// the dense N x N matrices Sigma and Delta = Sigma^TLR - Sigma.  Sigma, Delta (from the rule's compressed
// tiles), tile norms and symmetric parts come from this library's kernels; Cholesky factorizations, the
// inverse, triangular solves, products and the symmetric eigensolver are rocSOLVER / rocBLAS calls on the
// context's stream.

namespace {

struct DevArr {   // device allocation released on every return
  double* p = nullptr;
  ~DevArr() { if (p) (void)hipFree(p); }
};

struct Blas {
  rocblas_handle h = nullptr;
  ~Blas() { if (h) (void)rocblas_destroy_handle(h); }
};

// the rule's tiles at theta into D (Delta on the off-diagonal tiles; D zeroed first), ||Delta_ij||_F per tile
// and the ranks (tri_index order): the initial compression of an evaluation at theta in slot 0
int assembled_delta(tlr_ctx* c, const double theta[3], int32_t rule, double value, double* D,
                    std::vector<double>& err, std::vector<int>& rk) {
  const int64_t nt = c->ntile;
  hipStream_t st = c->stream;
  c->sstate[0].valid = false;   // slot 0's resident factor is given up
  const double bmax = (double)(c->n < c->b ? c->n : c->b);
  c->h_thr[0] = rule == TLR_RULE_FIXED ? value : 0.0;
  c->h_thr[1] = rule == TLR_RULE_FIXED ? 0.0 : value * bmax * theta[0] * theta[2] / (double)c->n;
  c->h_mp[0].sigma2 = theta[0];
  c->h_mp[0].inv_range_sqrt3 = sqrt(3.0) / theta[1];
  c->h_mp[0].nugget = theta[0] * theta[2];
  const SlotPtrs& p = c->slots[0].p;
  HIPCHK(hipMemcpyAsync(c->d_thr, c->h_thr, sizeof(double) * 2, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c->d_mp, c->h_mp, sizeof(MaternParams), hipMemcpyHostToDevice, st));
  HIPCHK(launch_fill_int(p.status, ST_COUNT, 0, st));
  HIPCHK(launch_fill_int(p.status + ST_FAILED, 1, 0x7fffffff, st));
  HIPCHK(launch_fill_int(c->d_retry, 1, 0, st));
  HIPCHK(hipMemsetAsync(D, 0, sizeof(double) * c->n * c->n, st));
  // the error terms measure the truncation itself: the tiles come from a 64-column first round (a wider
  // sketch leaves a smaller subspace defect beside the truncation than the likelihood's default)
  int rc = compress_slot(c, 0, p, rule == TLR_RULE_FIXED, nullptr, 0, -1, 0, 64);
  if (rc < 0) {
    (void)hipStreamSynchronize(st);
    return rc;
  }
  double* err2 = c->bufR[0];   // the compression's cores are dead by now: nt doubles of tile errors
  HIPCHK(launch_assemble_delta(c->sx, c->sy, c->n, c->b, c->d_mp, p.U, p.V, (int64_t)c->b * RANK_CAP, p.rank, D,
                               c->n, err2, nt, st));
  err.assign(nt > 0 ? nt : 1, 0.0);
  rk.assign(nt > 0 ? nt : 1, 0);
  int sts[ST_COUNT];
  if (nt > 0) {
    HIPCHK(hipMemcpyAsync(err.data(), err2, nt * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(rk.data(), p.rank, nt * sizeof(int), hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipMemcpyAsync(sts, p.status, sizeof(sts), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (sts[ST_GUARD]) return set_err(TLR_ERR_HIP, "a compression guard tripped (non-finite tile)");
  if (sts[ST_RANKCAP]) return set_err(TLR_ERR_RANKCAP, "a tile needs more columns than the rank capacity");
  for (int64_t t = 0; t < nt; ++t) err[t] = sqrt(err[t]);
  return TLR_OK;
}

#define RBCHK(x)                                                                   \
  do {                                                                             \
    rocblas_status rs_ = (x);                                                      \
    if (rs_ != rocblas_status_success)                                             \
      return set_err(TLR_ERR_HIP, std::string("rocBLAS/rocSOLVER: ") + rocblas_status_to_string(rs_)); \
  } while (0)

// sum over an array of m doubles (m may exceed the 32-bit counts of the BLAS calls) of x*y or x*x
int long_dot(rocblas_handle h, int64_t m, const double* x, const double* y, double* out) {
  const int64_t chunk = (int64_t)1 << 30;
  double acc = 0.0;
  for (int64_t o = 0; o < m; o += chunk) {
    const rocblas_int len = (rocblas_int)(m - o < chunk ? m - o : chunk);
    double r = 0.0;
    RBCHK(rocblas_ddot(h, len, x + o, 1, y + o, 1, &r));
    acc += r;
  }
  *out = acc;
  return TLR_OK;
}

// log L = -(y^T alpha + 2 sum log diag L + N log 2 pi) / 2 from a Cholesky factor L (lower, in A) of Sigma;
// alpha in the n-vector a; the log-determinant and the quadratic form y^T alpha also when asked for
int dense_loglik(tlr_ctx* c, rocblas_handle h, const double* A, double* a, double* ll, double* logdet = nullptr,
                 double* quad = nullptr) {
  const int64_t n = c->n;
  hipStream_t st = c->stream;
  HIPCHK(hipMemcpyAsync(a, c->y, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  RBCHK(rocsolver_dpotrs(h, rocblas_fill_lower, (rocblas_int)n, 1, const_cast<double*>(A), (rocblas_int)n, a,
                         (rocblas_int)n));
  std::vector<double> dg(n);
  HIPCHK(hipMemcpy2DAsync(dg.data(), sizeof(double), A, (n + 1) * sizeof(double), sizeof(double), n,
                          hipMemcpyDeviceToHost, st));
  double ya = 0.0;
  RBCHK(rocblas_ddot(h, (rocblas_int)n, c->y, 1, a, 1, &ya));
  HIPCHK(hipStreamSynchronize(st));
  double ld = 0.0;
  for (int64_t i = 0; i < n; ++i) ld += log(dg[i]);
  *ll = -0.5 * (ya + 2.0 * ld + (double)n * 1.8378770664093453);
  if (logdet) *logdet = 2.0 * ld;
  if (quad) *quad = ya;
  return TLR_OK;
}

}  // namespace

extern "C" int tlr_dense_loglik(tlr_ctx* c, int32_t m, const double* thetas, double* out) {
  if (!c || m < 0 || (m > 0 && (!thetas || !out))) return set_err(TLR_ERR_ARG, "invalid argument");
  for (int s = 0; s < m; ++s) {
    const double* th = thetas + 3 * s;
    if (!(th[0] > 0.0) || !(th[1] > 0.0) || !(th[2] >= 0.0)) return set_err(TLR_ERR_ARG, "theta must be positive");
  }
  if (c->arena) return set_err(TLR_ERR_ARG, "not available on an allocation-free context");
  const int64_t n = c->n, n2 = n * n;
  if (n > 46340) return set_err(TLR_ERR_ARG, "the dense log-likelihood is limited to N <= 46340");
  hipStream_t st = c->stream;
  HIPCHK(hipStreamSynchronize(st));
  DevArr K, v;   // Sigma and its factor; alpha (n) and the factorization's info word
  HIPCHK(hipMalloc(&K.p, n2 * sizeof(double)));
  HIPCHK(hipMalloc(&v.p, n * sizeof(double) + 64));
  rocblas_int* dinfo = reinterpret_cast<rocblas_int*>(v.p + n);
  Blas bl;
  RBCHK(rocblas_create_handle(&bl.h));
  RBCHK(rocblas_set_stream(bl.h, st));
  for (int s = 0; s < m; ++s) {
    const double* th = thetas + 3 * s;
    MaternParams mp;
    mp.sigma2 = th[0];
    mp.inv_range_sqrt3 = sqrt(3.0) / th[1];
    mp.nugget = th[0] * th[2];
    HIPCHK(launch_dense_cov(c->sx, c->sy, n, mp, K.p, n, st));
    RBCHK(rocsolver_dpotrf(bl.h, rocblas_fill_lower, (rocblas_int)n, K.p, (rocblas_int)n, dinfo));
    rocblas_int info = 0;
    HIPCHK(hipMemcpyAsync(&info, dinfo, sizeof(info), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double* o = out + 3 * (int64_t)s;
    if (info != 0) {   // not positive definite
      o[0] = o[1] = o[2] = NAN;
      continue;
    }
    const int rc = dense_loglik(c, bl.h, K.p, v.p, &o[0], &o[1], &o[2]);
    if (rc < 0) return rc;
  }
  return TLR_OK;
}

extern "C" {

int tlr_bound_terms(tlr_ctx* c, const double theta[3], int32_t nrules, const int32_t* rules, const double* values,
                    double dense[2], double* kinv_tiles, double* cov_tiles, double* alpha_blocks, double* terms,
                    double* tile_err, int32_t* ranks) {
  if (!c || !theta || nrules < 0 || (nrules > 0 && (!rules || !values || !terms)) || !dense)
    return set_err(TLR_ERR_ARG, "invalid argument");
  if (!(theta[0] > 0.0) || !(theta[1] > 0.0) || !(theta[2] >= 0.0)) return set_err(TLR_ERR_ARG, "theta must be positive");
  for (int q = 0; q < nrules; ++q) {
    if (rules[q] != TLR_RULE_FIXED && rules[q] != TLR_RULE_ADAPTIVE)
      return set_err(TLR_ERR_ARG, "rule must be fixed or adaptive");
    if (!(values[q] > 0.0)) return set_err(TLR_ERR_ARG, "rule value must be positive");
  }
  if (c->arena) return set_err(TLR_ERR_ARG, "not available on an allocation-free context");
  const int64_t n = c->n, nt = c->ntile, n2 = n * n;
  const int nB = c->nB;
  if (n > 46340) return set_err(TLR_ERR_ARG, "the dense analysis is limited to N <= 46340");
  hipStream_t st = c->stream;
  HIPCHK(hipStreamSynchronize(st));
  DevArr K, L, Ki, D, W, v;   // v: alpha, z, w, u, alpha_t (n each) and the eigenvalues (2 n)
  HIPCHK(hipMalloc(&K.p, n2 * sizeof(double)));
  HIPCHK(hipMalloc(&L.p, n2 * sizeof(double)));
  HIPCHK(hipMalloc(&Ki.p, n2 * sizeof(double)));
  HIPCHK(hipMalloc(&D.p, n2 * sizeof(double)));
  HIPCHK(hipMalloc(&W.p, n2 * sizeof(double)));
  HIPCHK(hipMalloc(&v.p, 7 * n * sizeof(double) + 64));
  double *alpha = v.p, *z = v.p + n, *w = v.p + 2 * n, *u = v.p + 3 * n, *at = v.p + 4 * n, *eig = v.p + 5 * n,
         *offd = v.p + 6 * n;
  rocblas_int* dinfo = reinterpret_cast<rocblas_int*>(v.p + 7 * n);
  Blas bl;
  RBCHK(rocblas_create_handle(&bl.h));
  RBCHK(rocblas_set_stream(bl.h, st));
  rocblas_handle h = bl.h;
  const rocblas_int ni = (rocblas_int)n;
  rocblas_int info = 0;
  // Sigma, its Cholesky factor, alpha, z, log L
  MaternParams mp;
  mp.sigma2 = theta[0];
  mp.inv_range_sqrt3 = sqrt(3.0) / theta[1];
  mp.nugget = theta[0] * theta[2];
  HIPCHK(launch_dense_cov(c->sx, c->sy, n, mp, K.p, n, st));
  HIPCHK(hipMemcpyAsync(L.p, K.p, n2 * sizeof(double), hipMemcpyDeviceToDevice, st));
  RBCHK(rocsolver_dpotrf(h, rocblas_fill_lower, ni, L.p, ni, dinfo));
  HIPCHK(hipMemcpyAsync(&info, dinfo, sizeof(info), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (info != 0) return set_err(TLR_ERR_ARG, "Sigma(theta) is not positive definite");
  int rc = dense_loglik(c, h, L.p, alpha, &dense[0]);
  if (rc < 0) return rc;
  HIPCHK(hipMemcpyAsync(z, c->y, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  RBCHK(rocblas_dtrsv(h, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_non_unit, ni, L.p, ni, z, 1));
  RBCHK(rocblas_ddot(h, ni, z, 1, z, 1, &dense[1]));
  // Sigma^-1 (the lower triangle from potri, then mirrored)
  HIPCHK(hipMemcpyAsync(Ki.p, L.p, n2 * sizeof(double), hipMemcpyDeviceToDevice, st));
  RBCHK(rocsolver_dpotri(h, rocblas_fill_lower, ni, Ki.p, ni, dinfo));
  HIPCHK(launch_symmetrize(Ki.p, n, n, 0, st));
  std::vector<double> tmp(nt > 0 ? nt : 1);
  if (kinv_tiles && nt > 0) {
    HIPCHK(launch_tile_fro(Ki.p, n, n, c->b, c->bufR[0], nt, st));
    HIPCHK(hipMemcpyAsync(kinv_tiles, c->bufR[0], nt * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  if (cov_tiles && nt > 0) {
    HIPCHK(launch_tile_fro(K.p, n, n, c->b, c->bufR[0], nt, st));
    HIPCHK(hipMemcpyAsync(cov_tiles, c->bufR[0], nt * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  if (alpha_blocks) {
    std::vector<double> ah(n);
    HIPCHK(hipMemcpyAsync(ah.data(), alpha, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int i = 0; i < nB; ++i) {
      double s = 0.0;
      for (int64_t p = (int64_t)i * c->b; p < n && p < (int64_t)(i + 1) * c->b; ++p) s += ah[p] * ah[p];
      alpha_blocks[i] = sqrt(s);
    }
  }
  std::vector<double> err;
  std::vector<int> rk;
  std::vector<double> ev(n);
  const double one = 1.0, zero = 0.0;
  for (int q = 0; q < nrules; ++q) {
    double* tq = terms + 6 * (int64_t)q;
    rc = assembled_delta(c, theta, rules[q], values[q], D.p, err, rk);
    if (rc < 0) return rc;
    if (tile_err)
      for (int64_t t = 0; t < nt; ++t) tile_err[(int64_t)q * nt + t] = err[t];
    if (ranks) {
      int32_t* d = ranks + (int64_t)q * nB * nB;
      memset(d, 0, sizeof(int32_t) * nB * nB);
      for (int i = 1; i < nB; ++i)
        for (int j = 0; j < i; ++j) d[i * nB + j] = rk[tri_index(i, j)];
    }
    // log L of Sigma^TLR = Sigma + Delta (NaN when it is not positive definite)
    RBCHK(rocblas_dgeam(h, rocblas_operation_none, rocblas_operation_none, ni, ni, &one, K.p, ni, &one, D.p, ni, W.p,
                        ni));
    RBCHK(rocsolver_dpotrf(h, rocblas_fill_lower, ni, W.p, ni, dinfo));
    HIPCHK(hipMemcpyAsync(&info, dinfo, sizeof(info), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int e = 0; e < 6; ++e) tq[e] = NAN;
    if (info != 0) continue;
    rc = dense_loglik(c, h, W.p, at, &tq[0]);
    if (rc < 0) return rc;
    // first order: tr(Sigma^-1 Delta) (both symmetric: the sum of their entrywise product), alpha^T Delta alpha
    rc = long_dot(h, n2, Ki.p, D.p, &tq[1]);
    if (rc < 0) return rc;
    RBCHK(rocblas_dgemv(h, rocblas_operation_none, ni, ni, &one, D.p, ni, alpha, 1, &zero, w, 1));
    RBCHK(rocblas_ddot(h, ni, alpha, 1, w, 1, &tq[2]));
    // ||E z||^2 = (Delta alpha)^T Sigma^-1 (Delta alpha)
    RBCHK(rocblas_dgemv(h, rocblas_operation_none, ni, ni, &one, Ki.p, ni, w, 1, &zero, u, 1));
    RBCHK(rocblas_ddot(h, ni, w, 1, u, 1, &tq[5]));
    // E = L^-1 Delta L^-T: two triangular solves with N right-hand sides
    HIPCHK(hipMemcpyAsync(W.p, D.p, n2 * sizeof(double), hipMemcpyDeviceToDevice, st));
    RBCHK(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_non_unit,
                        ni, ni, &one, L.p, ni, W.p, ni));
    RBCHK(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                        rocblas_diagonal_non_unit, ni, ni, &one, L.p, ni, W.p, ni));
    double ef2 = 0.0;
    rc = long_dot(h, n2, W.p, W.p, &ef2);
    if (rc < 0) return rc;
    tq[3] = sqrt(ef2);
    // ||E||_2: the largest eigenvalue magnitude of its symmetric part
    HIPCHK(launch_symmetrize(W.p, n, n, 1, st));
    RBCHK(rocsolver_dsyevd(h, rocblas_evect_none, rocblas_fill_lower, ni, W.p, ni, eig, offd, dinfo));
    HIPCHK(hipMemcpyAsync(ev.data(), eig, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&info, dinfo, sizeof(info), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (info != 0) return set_err(TLR_ERR_HIP, "symmetric eigensolver did not converge");
    double e2 = 0.0;
    for (int64_t i = 0; i < n; ++i) e2 = fmax(e2, fabs(ev[i]));
    tq[4] = e2;
  }
  HIPCHK(hipStreamSynchronize(st));
  return TLR_OK;
}

}  // extern "C"
