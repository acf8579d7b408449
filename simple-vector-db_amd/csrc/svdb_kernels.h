// svdb_kernels.h — launch interface of the gfx950 kernels (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace svdb {

enum Metric { kL2 = 0, kCos = 1, kDot = 2 };

// Queries per wavefront in the FP64 scan; queries are wave-uniform there.
constexpr int kScanTQ = 16;
// Rows per wavefront step in the FP64 scan (two per lane).
constexpr int kScanRows = 128;
// Wavefronts per block in the FP64 scan and merge kernels.
constexpr int kWavesPerBlock = 4;

struct StoreView {
  const double* X;          // FP64 master rows, row stride ldx (elements)
  size_t ldx;
  const double* xn;         // sqrt(Σ fma(x_i,x_i,·)) per row
  const uint64_t* alive;    // 1 bit per slot
  const uint32_t* rank_base;// alive slots before each 64-slot word
  uint32_t nrows;           // slots in use
  int has_deletes;          // 0: logical id == slot
  const double* nx2;        // Σ fma(x_i,x_i,·) per row when the scan compares
                            // every column, else null
};

struct ScanF64Args {
  StoreView st;
  const double* Q;          // queries, stride ldq
  size_t ldq;
  const double* qn;         // per-query norm (cosine)
  const int* k;             // per-query k
  uint32_t nq;
  uint32_t D;               // leading columns compared
  uint32_t rows_per_slice;
  uint32_t P;               // partial lists per query
  uint32_t kstride;         // entries per partial list (>= max k)
  uint64_t* part_key;       // [nq][P][kstride]
  uint32_t* part_slot;
  // Per query, the smallest k-th rank key any partial list has reached so
  // far (all ones before the scan; the matrix-core scan lowers it with
  // atomic min and skips every row whose key is above it), or null.
  uint64_t* gtk;
  const double* qq;         // per-query Σ fma(q_i,q_i,·) (Euclidean on the
                            // matrix cores), or null
};

struct MergeArgs {
  StoreView st;
  const int* k;
  uint32_t nq;
  uint32_t P;
  uint32_t kstride;
  const uint64_t* part_key;
  const uint32_t* part_slot;
  uint32_t ostride;
  uint64_t* ids;            // logical ids, (uint64)-1 when missing
  double* scores;           // FP64 scores, NaN when missing
};

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);
// p[0..n) = v
hipError_t launch_fill_i32(int* p, uint32_t n, int v, hipStream_t s);
hipError_t launch_row_norms(const double* X, size_t ldx, const uint32_t* slots, uint32_t first,
                            uint32_t n, uint32_t d, double* nx2, double* xn, hipStream_t s);
// qn[i] = sqrt(Σ fma(q,q,·)) over the first D columns; qq[i] the sum itself
// when qq is not null.
hipError_t launch_query_norms(const double* Q, size_t ldq, uint32_t nq, uint32_t D, double* qn,
                              double* qq,
                              hipStream_t s);
// Whether the FP64 scan of this metric, width and row layout runs on the
// matrix cores, and how many partial lists per query it then writes for
// nslices row slices (the value ScanF64Args.P and MergeArgs.P must hold).
// (kmax: the largest k of the batch, ScanF64Args.kstride; nq its queries)
bool scan_f64_uses_mfma(int metric, uint32_t D, const StoreView& st, int kmax, uint32_t nq);
uint32_t scan_f64_lists(int metric, uint32_t D, const StoreView& st, int kmax, uint32_t nq,
                        uint32_t nslices);
hipError_t launch_scan_f64(const ScanF64Args& a, int metric, int kmax, uint32_t nslices,
                           hipStream_t s);
hipError_t launch_merge(const MergeArgs& a, int metric, int kmax, hipStream_t s);

// The nearest live row to one query over the first D columns (canonical
// Euclidean key, lowest slot on ties) in one launch: each block reduces its
// rows, the last block to finish reduces the blocks' results, writes the
// row's logical id ((uint64)-1 when no row is live) to out_id — host
// memory the caller reads after the launch — and puts the ticket back to 0.
constexpr uint32_t kNearestMaxD = 256;
constexpr uint32_t kNearestMaxBlocks = 1024;
struct Nearest1Args {
  StoreView st;
  uint32_t D;
  uint64_t* part_key;   // [kNearestMaxBlocks]
  uint32_t* part_slot;  // [kNearestMaxBlocks]
  uint32_t* ticket;     // 0 on entry and on exit
  uint64_t* out_id;     // [2]: the id, then seq once the id is visible
  uint64_t seq;
  double q[kNearestMaxD];
};
hipError_t launch_nearest1(const Nearest1Args& a, hipStream_t s);
// A loop of independent MFMAs (kind 0 FP64 16x16x4, 1 BF16 16x16x32, 2 FP8
// 16x16x128) over `blocks` blocks of 4 wavefronts; out holds blocks · 256
// floats.  mfma_loop_flop gives the flop of one such launch.
hipError_t launch_mfma_loop(int kind, uint32_t blocks, int iters, float* out, hipStream_t s);
double mfma_loop_flop(int kind, uint32_t blocks, int iters);
hipError_t launch_pair_f64(const double* X, size_t ldx, const uint32_t* sa, const uint32_t* sb,
                           uint32_t np, uint32_t D, int metric, double* out, hipStream_t s);
hipError_t launch_pair_ref32(const double* M, size_t ld, const uint32_t* sa, const uint32_t* sb,
                             uint32_t np, uint32_t n, int metric, float* out, hipStream_t s);
// One pair of at most kRef32ArgsMax elements each, carried in the kernel
// arguments (a then b in ab), the reference's float metric written to
// host memory: out[0] the float's bits, then out[1] = seq.
constexpr uint32_t kRef32ArgsMax = 224;
struct PairRef32Args {
  uint32_t n;
  int metric;
  uint64_t* out;  // [2], host memory
  uint64_t seq;
  double ab[2 * kRef32ArgsMax];
};
hipError_t launch_pair_ref32_args(const PairRef32Args& a, hipStream_t s);

}  // namespace svdb
