// svdb_lowp.h — low-precision first pass + FP64 rerank (internal).
//
// Shadow of the store (per slot):
//   Xb      bf16 row, zero-padded to dp = round_up(d, 32)
//   aux_nx  fp32 ‖x‖² rounded down      (Euclidean ordering key)
//   aux_xn  fp32 ‖x‖ rounded up         (dot ordering key)
//   aux_inv fp32 1/‖x‖                  (cosine ordering key)
//   aux_pen fp32 0, or +inf for rows the first pass must never pick
// Dead slots and "degenerate" rows (zero, tiny, huge or non-finite) carry
// values that make every ordering key +inf; degenerate rows go to a short
// list that every query's rerank includes, so they are never missed.
//
// First-pass error bound (per query, for every row x):
//   |s - q·x| <= c_q·‖x‖ + δ,  s = fp32-accumulated q̂·x̂ (bf16 q̂, x̂)
//   c_q = γ‖q̂‖(1+ρ) + ‖q̂-q‖ + ρ‖q̂‖(1+γ) + FP64-key slack,
//   ρ = max over stored rows of ‖x̂-x‖/‖x‖, γ = (d+32)·2^-24, δ = d·2^-124.
// Each metric's fp32 ordering key O is built so that the exact rank key of
// every row satisfies R >= O·scale_q + off_q; a query is certified when its
// k-th exact FP64 key is below that bound for the best row left out.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "svdb_kernels.h"

namespace svdb {

constexpr int kLowpWaves = 8;                      // wavefronts per block
constexpr int kLowpQF = 4;                         // 16-query fragments per wavefront
constexpr int kLowpQPB = kLowpWaves * kLowpQF * 16;  // queries per block (512)
constexpr int kLowpTile = 128;                     // rows per LDS tile
constexpr int kLowpLP = 8;                         // entries per per-lane list
constexpr int kLowpLanesPerQuery = 4;              // lists per query per slice
constexpr int kDegenCap = 64;                      // degenerate rows kept in the list
constexpr int kMaxShort = 4096;                    // widest shortlist (the widening retries)
constexpr int kMixed = 3;                          // scan "metric": per-wavefront metric from QParams
constexpr int kSlotAlign = 64;                     // metric ranges start on wavefront boundaries

struct Shadow {
  uint16_t* Xb = nullptr;   // [cap][dp]
  size_t dp = 0;            // padded row length (elements)
  float* nx = nullptr;
  float* xn = nullptr;
  float* inv = nullptr;
  float* pen = nullptr;
  uint32_t* rho_bits = nullptr;   // fp32 bits of ρ (running max)
  uint32_t* degen_count = nullptr;
  uint32_t* degen = nullptr;      // [kDegenCap]
};

// FP8 (OCP e4m3fn) shadow with per-row power-of-two scales:
//   X8      e4m3 row of x / sx, zero-padded to dp = round_up(d, 256)
//   sx      fp32 2^e with max|x| / sx in (224, 448]
//   nx, xn  as in Shadow
//   inv     fp32 sx / ‖x‖ (cosine ordering key, sx folded in)
//   pen     as in Shadow
// x̂ = sx·e4m3(x/sx) exactly, so every FP8 ordering key is an exact
// power-of-two rescaling of the BF16-form key and the same bound applies
// with this shadow's ρ.
struct Shadow8 {
  uint8_t* X8 = nullptr;    // [cap][dp]
  size_t dp = 0;
  float* sx = nullptr;
  float* nx = nullptr;
  float* xn = nullptr;
  float* inv = nullptr;
  float* pen = nullptr;
  uint32_t* rho_bits = nullptr;
  uint32_t* degen_count = nullptr;
  uint32_t* degen = nullptr;      // [kDegenCap]
};

// Per-query parameters written by the prep kernel.
struct QParams {
  float P;        // epilogue multiplier (L2: a_q; dot: -c'_q; cos: unused)
  int flags;      // 1 = degenerate query or padding slot: FP64 only, never scanned
  double scale;   // R >= O*scale + off
  double off;
  double qn;      // FP64 ‖q‖ (canonical chain)
  double nq2;     // FP64 ‖q‖²
  int metric;     // kL2 / kCos / kDot (the first pass's key for this slot)
};

struct LowpScanArgs {
  const uint16_t* Xb;
  size_t dp;
  const float* w0;       // per-row array for the key (nx / xn / inv)
  const float* w1;       // cosine: pen
  const float* w2;       // mixed-metric scan only: xn (dot)
  const float* w3;       //                         inv (cosine)
  const uint16_t* Qb;    // [nq_pad][dp]
  const QParams* qp;
  uint32_t nq;
  uint32_t rows_per_slice;  // multiple of kLowpTile
  uint32_t slice_stride;    // rows from one slice's start to the next's (>= rows_per_slice)
  uint32_t nrows_pad;       // rows readable in Xb / aux (multiple of kLowpTile)
  uint32_t NL;              // lists per query = slices * 4
  float* cand_key;          // [nq][NL][LP]
  uint32_t* cand_slot;
  // Optional per-query starting threshold: query q's lists only take keys
  // below tau0[q * tau0_stride].  It must be the ordering key of an actual
  // row that has at least L-1 rows keyed at or below it (e.g. the L-th key
  // of a shortlist over a sample of the rows), so the global top-L lies
  // below it; a list left short then reports the threshold as its bound.
  const float* tau0;
  uint32_t tau0_stride;
};

struct ShortlistArgs {
  const float* cand_key;
  const uint32_t* cand_slot;
  uint32_t nq, NL, L;
  uint32_t* short_slot;     // [nq][L]  (UINT32_MAX when fewer)
  float* short_key;         // [nq][L]
  float* beta;              // [nq] bound on the ordering key of rows left out
  const uint32_t* qidx;     // optional: row i of the outputs is query qidx[i]
};

struct RerankArgs {
  StoreView st;
  const double* Q;          // FP64 queries, stride ldq
  size_t ldq;
  uint32_t d;
  const QParams* qp;
  const uint32_t* short_slot;
  const float* short_key;
  const float* beta;
  const uint32_t* degen;
  const uint32_t* degen_count;
  const float* w0;          // ordering-key row arrays (for the error report)
  const float* w1;
  uint32_t nq, L, k;
  uint32_t ostride;
  uint64_t* ids;
  double* scores;
  uint32_t* fail_count;     // queries not certified
  uint32_t* fail_list;      // their indices
  double* max_rel_err;      // [nq] first-pass score error of the returned scores
  const uint32_t* qidx;     // optional: shortlist row i is query qidx[i]
  // Optional carried state, indexed by query q: a stage's top k (FP64
  // ordering keys lk[q*k ..], slots ls[q*k ..]) and kmax[q], the largest
  // first-pass key of its shortlist and the earlier ones (every candidate
  // keyed below it was scored there, or pruned as unable to enter the top
  // k).  A query left uncertified writes them; with carry != 0 a stage
  // starts its list from them and skips candidates keyed below kmax.  Each
  // query's entries are read and written by one wavefront.
  uint64_t* lk;
  uint32_t* ls;
  float* kmax;
  int carry;
};

// Queries per block of the scan for a padded row length dp.
int lowp_queries_per_block(size_t dp);

hipError_t launch_shadow_init(const Shadow& sh, uint32_t first, uint32_t n, hipStream_t s);
hipError_t launch_quantize_rows(const Shadow& sh, const double* X, size_t ldx, const double* nx2,
                                const double* xn, const uint32_t* slots, uint32_t first,
                                uint32_t n, uint32_t d, hipStream_t s);
hipError_t launch_kill_rows(const Shadow& sh, const uint32_t* slots, uint32_t n, hipStream_t s);
hipError_t launch_query_prep(const double* Q, size_t ldq, uint32_t nq, uint32_t nq_pad,
                             uint32_t d, size_t dp, int metric, const uint32_t* rho_bits,
                             uint16_t* Qb, QParams* qp, hipStream_t s);
// minonly: one entry per list, its smallest key (the sample pre-pass).
hipError_t launch_scan_bf16(const LowpScanArgs& a, int metric, uint32_t nslices, bool minonly,
                            hipStream_t s);
// thr[q] = the L-th smallest of the n keys key[q*n .. q*n+n) (+inf when
// fewer than L are finite).  One wavefront per query.
hipError_t launch_kth_key(const float* key, uint32_t nq, uint32_t n, uint32_t L, float* thr,
                          hipStream_t s);
// ---- wide rows (svdb_wide.hip): BF16 at padded widths 512 and 768 ----
bool bf16_wide_dp_supported(size_t dp);
hipError_t launch_scan_bf16_wide(const LowpScanArgs& a, int metric, uint32_t nslices,
                                 bool minonly, hipStream_t s);

// ---- FP8 first pass (svdb_wide.hip) ----
// FP8 rows are padded to a multiple of 256 bytes; supported padded lengths.
bool fp8_dp_supported(size_t dp);
int fp8_queries_per_block(size_t dp);
hipError_t launch_shadow_init8(const Shadow8& sh, uint32_t first, uint32_t n, hipStream_t s);
hipError_t launch_quantize_rows8(const Shadow8& sh, const double* X, size_t ldx, const double* nx2,
                                 const double* xn, const uint32_t* slots, uint32_t first,
                                 uint32_t n, uint32_t d, hipStream_t s);
hipError_t launch_kill_rows8(const Shadow8& sh, const uint32_t* slots, uint32_t n, hipStream_t s);
// Q8: e4m3 queries [nq_pad][dp]; QParams as for BF16 with P / scale carrying
// the query's power-of-two scale (see svdb_fp8.hip).
hipError_t launch_query_prep8(const double* Q, size_t ldq, uint32_t nq, uint32_t nq_pad,
                              uint32_t d, size_t dp, int metric, const uint32_t* rho_bits,
                              uint8_t* Q8, QParams* qp, hipStream_t s);
// a.Xb / a.Qb point at e4m3 bytes; a.w0 / a.w1: L2 (nx, sx), dot (xn, sx),
// cosine (inv, pen).
hipError_t launch_scan_fp8(const LowpScanArgs& a, int metric, uint32_t nslices, bool minonly,
                           hipStream_t s);

hipError_t launch_shortlist(const ShortlistArgs& a, hipStream_t s);
hipError_t launch_rerank(const RerankArgs& a, int metric, hipStream_t s);
// n ranges in as few launches as possible (kMaxLaunchRanges per launch; for
// the shortlist, consecutive ranges that take the same kernel)
constexpr int kMaxLaunchRanges = 4;
hipError_t launch_shortlist_ranges(const ShortlistArgs* a, int n, hipStream_t s);
hipError_t launch_rerank_ranges(const RerankArgs* a, const int* metric, int n, hipStream_t s);
// Rebuilds a shadow's degenerate list (count, list[kDegenCap]) from the
// live slots whose penalty is +inf.
hipError_t launch_rebuild_degen(const float* pen, const uint64_t* alive, uint32_t nslots,
                                uint32_t* count, uint32_t* list, hipStream_t s);
hipError_t launch_gather_rows(const double* src, size_t lds, const uint32_t* idx, uint32_t n,
                              uint32_t d, double* dst, size_t ldd, hipStream_t s);
hipError_t launch_scatter_results(const uint64_t* ids, const double* sc, uint32_t k,
                                  const uint32_t* idx, uint32_t n, uint64_t* out_ids,
                                  double* out_sc, uint32_t ostride, hipStream_t s);

}  // namespace svdb
