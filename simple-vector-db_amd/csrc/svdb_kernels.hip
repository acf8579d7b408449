// svdb_kernels.hip — gfx950 kernels of the similarity / nearest-vector path.
//
// Built with -ffp-contract=off: every a*b+c below that is not an explicit
// fma() is a rounded multiply then a rounded add; the explicit fma() calls
// are the canonical FP64 steps (correctly rounded v_fma_f64).  sqrt and '/'
// on doubles lower to the correctly rounded sequences (v_rsq_f64 + Newton
// steps + fixup; v_div_scale/v_div_fmas/v_div_fixup).
//
// Canonical FP64 arithmetic (matches oracle/svdb_oracle.c bit for bit):
//   L2 key  s = Σ_i fma(t, t, s), t = x_i - q_i, i = 0..D-1 in order
//   dot key s = Σ_i fma(q_i, x_i, s)
//   cosine  dot / (sqrt(Σ fma(q,q)) * sqrt(Σ fma(x,x)))
// Ranking: ascending "rank key" (L2: s; dot/cos: -score) mapped to a uint64
// that orders like the doubles (-0 folded to +0, NaN above +inf), then the
// slot (physical row), which orders like the logical index.
#include "svdb_kernels.h"

#include <math.h>

#include <algorithm>

namespace svdb {

// ----------------------------------------------------------------- helpers

__device__ __forceinline__ uint64_t rank_bits(double v) {
  v = v + 0.0;  // -0 → +0 (not folded: strict FP)
  uint64_t u = (uint64_t)__double_as_longlong(v);
  if (v != v) u = 0x7FF8000000000000ull;
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__device__ __forceinline__ double rank_value(uint64_t u) {
  uint64_t b = (u >> 63) ? (u & 0x7FFFFFFFFFFFFFFFull) : ~u;
  return __longlong_as_double((long long)b);
}

__device__ __forceinline__ bool lex_lt(uint64_t ka, uint32_t sa, uint64_t kb, uint32_t sb) {
  return (ka < kb) | ((ka == kb) & (sa < sb));
}

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t uniform32(uint32_t v) {
  return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ uint64_t read_lane64(uint64_t v, int l) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ double read_lane_f64(double v, int l) {
  return __longlong_as_double((long long)read_lane64((uint64_t)__double_as_longlong(v), l));
}

// Cosine rank keys -(dot / qx) of a lane's two rows (qx = ‖q‖·‖x‖, the
// canonical operand order).  The correctly rounded division costs a dozen
// instructions per (query, row), so it is skipped for the whole wavefront
// when a cheap estimate shows that no lane's rows can enter the list whose
// k-th rank key is tk (the keys are then +inf, which nothing inserts).  The
// estimate dot·rcp(qx) is within 2^-20 relative of the quotient (v_rcp_f64
// is good to about 2^-26, the product rounds once), so an estimated key
// above the k-th key by more than 2^-19 of its magnitude ranks after it;
// non-finite, zero or extreme operands and a list not yet full take the
// exact path.
__device__ __forceinline__ void cos_keys(double dot0, double qx0, double dot1, double qx1,
                                         uint64_t tk, double& k0, double& k1) {
  bool exact = true;
  if (tk != ~0ull) {
    const double thr = rank_value(tk);
    const double e0 = -(dot0 * __builtin_amdgcn_rcp(qx0));
    const double e1 = -(dot1 * __builtin_amdgcn_rcp(qx1));
    // (qx far from the extremes: its reciprocal stays a normal number)
    const bool skip0 = qx0 > 0x1p-1000 && qx0 < 0x1p1000 && isfinite(e0) &&
                       e0 - thr > 0x1p-19 * fabs(e0);
    const bool skip1 = qx1 > 0x1p-1000 && qx1 < 0x1p1000 && isfinite(e1) &&
                       e1 - thr > 0x1p-19 * fabs(e1);
    exact = !isfinite(thr) || __ballot(!(skip0 && skip1)) != 0;
  }
  if (exact) {
    k0 = -(dot0 / qx0);
    k1 = -(dot1 / qx1);
  } else {
    k0 = k1 = INFINITY;
  }
}

// A sorted list of (key, slot) spread over a wavefront: entry e lives in
// lane e % 64, register e / 64.  Unused entries hold the sentinel
// (UINT64_MAX, UINT32_MAX), which is above every real entry (slots are
// < UINT32_MAX).
template <int KPL>
struct WList {
  uint64_t k[KPL];
  uint32_t s[KPL];
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int p = 0; p < KPL; p++) {
      k[p] = ~0ull;
      s[p] = ~0u;
    }
  }
  // Entry kk-1 (the current k-th best), wave-uniform.
  __device__ __forceinline__ void kth(int kk, uint64_t& tk, uint32_t& ts) const {
    int e = kk - 1, p = e >> 6, ln = e & 63;
    uint64_t a = k[0];
    uint32_t b = s[0];
#pragma unroll
    for (int q = 1; q < KPL; q++)
      if (p == q) {
        a = k[q];
        b = s[q];
      }
    tk = uniform64(__shfl(a, ln));
    ts = uniform32(__shfl(b, ln));
  }
  // Insert (K, S) (wave-uniform) keeping the first kk entries the kk best.
  __device__ __forceinline__ void insert(uint64_t K, uint32_t S, int kk, int lane) {
    int pos = 0;
#pragma unroll
    for (int p = 0; p < KPL; p++) pos += __popcll(__ballot(lex_lt(k[p], s[p], K, S)));
    if (pos >= kk) return;
    uint64_t pk[KPL];
    uint32_t ps[KPL];
#pragma unroll
    for (int p = 0; p < KPL; p++) {
      pk[p] = __shfl_up(k[p], 1);
      ps[p] = __shfl_up(s[p], 1);
    }
#pragma unroll
    for (int p = 1; p < KPL; p++) {
      uint64_t ck = __shfl(k[p - 1], 63);
      uint32_t cs = __shfl(s[p - 1], 63);
      if (lane == 0) {
        pk[p] = ck;
        ps[p] = cs;
      }
    }
#pragma unroll
    for (int p = 0; p < KPL; p++) {
      int e = p * 64 + lane;
      if (e > pos) {
        k[p] = pk[p];
        s[p] = ps[p];
      } else if (e == pos) {
        k[p] = K;
        s[p] = S;
      }
    }
  }
};

__device__ __forceinline__ bool slot_alive(const uint64_t* alive, uint32_t r) {
  return (alive[r >> 6] >> (r & 63)) & 1ull;
}

__device__ __forceinline__ uint64_t logical_id(const StoreView& st, uint32_t slot) {
  if (!st.has_deletes) return slot;
  uint32_t w = slot >> 6;
  uint64_t below = st.alive[w] & ((1ull << (slot & 63)) - 1ull);
  return (uint64_t)st.rank_base[w] + (uint64_t)__popcll(below);
}

// ------------------------------------------------------------ store writes

// One thread per row: the sequential fma chain of the squared norm of a row
// already copied into the master (padding columns are zero).  Rows are the
// slots first..first+n-1, or slots[0..n) when a slot list is given.
__global__ void k_row_norms(const double* __restrict__ X, size_t ldx,
                            const uint32_t* __restrict__ slots, uint32_t first, uint32_t n,
                            uint32_t d, double* __restrict__ nx2, double* __restrict__ xn) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  uint32_t slot = slots ? slots[r] : first + r;
  const double* x = X + (size_t)slot * ldx;
  double acc = 0.0;
  for (uint32_t i = 0; i < d; i++) acc = fma(x[i], x[i], acc);
  nx2[slot] = acc;
  xn[slot] = sqrt(acc);
}

hipError_t launch_row_norms(const double* X, size_t ldx, const uint32_t* slots, uint32_t first,
                            uint32_t n, uint32_t d, double* nx2, double* xn, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_row_norms, dim3((n + 255) / 256), dim3(256), 0, s, X, ldx, slots, first, n,
                     d, nx2, xn);
  return hipGetLastError();
}

__global__ void k_query_norms(const double* __restrict__ Q, size_t ldq, uint32_t nq, uint32_t D,
                              double* __restrict__ qn, double* __restrict__ qq) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const double* q = Q + (size_t)i * ldq;
  double acc = 0.0;
  for (uint32_t c = 0; c < D; c++) acc = fma(q[c], q[c], acc);
  qn[i] = sqrt(acc);
  if (qq) qq[i] = acc;
}

hipError_t launch_query_norms(const double* Q, size_t ldq, uint32_t nq, uint32_t D, double* qn,
                              double* qq, hipStream_t s) {
  if (nq == 0) return hipSuccess;
  hipLaunchKernelGGL(k_query_norms, dim3((nq + 255) / 256), dim3(256), 0, s, Q, ldq, nq, D, qn,
                     qq);
  return hipGetLastError();
}

// Plain device copy, 16 bytes per lane, grid-stride (also the known-traffic
// reference for hardware byte counters: it reads and writes `bytes` once).
__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void k_copy1(const unsigned char* __restrict__ src, unsigned char* __restrict__ dst,
                        size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  const bool aligned = ((uintptr_t)src % 16 == 0) && ((uintptr_t)dst % 16 == 0);
  const size_t n16 = aligned ? bytes / 16 : 0;
  if (n16) {
    const size_t blocks = std::min<size_t>((n16 + 255) / 256, 256 * 32);
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, s, (const uint4*)src,
                       (uint4*)dst, n16);
  }
  const size_t rest = bytes - n16 * 16;
  if (rest)
    hipLaunchKernelGGL(k_copy1, dim3((unsigned)((rest + 255) / 256)), dim3(256), 0, s,
                       (const unsigned char*)src + n16 * 16, (unsigned char*)dst + n16 * 16, rest);
  return hipGetLastError();
}

__global__ void k_fill_i32(int* __restrict__ p, uint32_t n, int v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_fill_i32(int* p, uint32_t n, int v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i32, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

// ---------------------------------------------------------- FP64 scan
//
// Grid (query groups of kScanTQ, row slices); 4 wavefronts per block share
// the block's queries and take interleaved 128-row steps of the slice.  Each
// lane owns two rows per step (row base+lane and base+64+lane) and keeps the
// kScanTQ×2 FP64 accumulators; query values are wave-uniform loads.  After
// a step every (query, row) key is tested against that query's current k-th
// best (wave-uniform); hits are inserted into the query's wave-distributed
// sorted list.  Each wavefront writes its lists as one partial list per
// query.

template <int METRIC>
__device__ __forceinline__ void accum8(double (&acc)[kScanTQ][2], const double (&x0)[8],
                                       const double (&x1)[8], const double* __restrict__ Qg,
                                       size_t ldq, uint32_t c) {
  // Queries in groups of G: a group's 8·G values are scalar loads, and the
  // next group's loads are issued (into the other of two register sets)
  // before this group's fmas, with scheduling barriers keeping it that way —
  // left to itself the compiler hoists all 16 queries' loads to the top
  // (256 scalar registers) and spills most of them.  A group's 2·G
  // (query, row) chains interleave; every chain advances in element order.
  constexpr int G = 2;
  constexpr int NG = kScanTQ / G;
  double qv[2][G][8];
#pragma unroll
  for (int jj = 0; jj < G; jj++)
#pragma unroll
    for (int i = 0; i < 8; i++) qv[0][jj][i] = Qg[(size_t)jj * ldq + c + i];
#pragma unroll
  for (int g = 0; g < NG; g++) {
    if (g + 1 < NG) {
#pragma unroll
      for (int jj = 0; jj < G; jj++)
#pragma unroll
        for (int i = 0; i < 8; i++)
          qv[(g + 1) & 1][jj][i] = Qg[(size_t)((g + 1) * G + jj) * ldq + c + i];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < 8; i++) {
#pragma unroll
      for (int jj = 0; jj < G; jj++) {
        const int j = g * G + jj;
        const double q = qv[g & 1][jj][i];
        if (METRIC == kL2) {
          double t0 = x0[i] - q, t1 = x1[i] - q;
          acc[j][0] = fma(t0, t0, acc[j][0]);
          acc[j][1] = fma(t1, t1, acc[j][1]);
        } else {
          acc[j][0] = fma(q, x0[i], acc[j][0]);
          acc[j][1] = fma(q, x1[i], acc[j][1]);
        }
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int METRIC, int KPL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KPL == 1 ? 3 : 2))) void
k_scan_f64(ScanF64Args a) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  // XCD-aware block order: block ids ≡ x (mod 8) run on XCD x, which gets
  // whole slices (slice ≡ x mod 8) with all their query groups one after
  // the other; the groups of a slice then stream its rows through that
  // XCD's L2 in near lockstep instead of each from HBM.
  const uint32_t ngroups = (a.nq + kScanTQ - 1) / kScanTQ;
  const uint32_t xcd = blockIdx.x % 8, local = blockIdx.x / 8;
  const uint32_t q0 = (local % ngroups) * kScanTQ;
  const uint32_t slice = (local / ngroups) * 8 + xcd;
  const uint32_t r_begin = slice * a.rows_per_slice;
  const uint32_t r_end = min(r_begin + a.rows_per_slice, a.st.nrows);
  // Queries past nq reuse the last query's row (their lists are not written).
  const uint32_t qlast = a.nq - 1;
  const double* __restrict__ Qg = a.Q + (size_t)min(q0, qlast) * a.ldq;
  const bool full_group = q0 + kScanTQ <= a.nq;

  WList<KPL> L[kScanTQ];
  // Per-query state, lane j holding query j's: the current k-th best
  // (key, slot), k and (cosine) ‖q‖.  Kept in vector registers and read
  // into scalars one query at a time in the epilogue: as arrays of
  // wave-uniform values they needed about 100 scalar registers and the
  // kernel spilled scalars throughout its main loop.
  uint64_t tk_l = ~0ull;
  uint32_t ts_l = ~0u;
  int kq_l = 0;
  double qn_l = 0.0;
  if (lane < kScanTQ) {
    const uint32_t qi = min(q0 + lane, qlast);
    kq_l = a.k[qi];
    qn_l = METRIC == kCos ? a.qn[qi] : 0.0;
  }
#pragma unroll
  for (int j = 0; j < kScanTQ; j++) L[j].init();

  for (uint32_t base = r_begin + wave * kScanRows; base < r_end;
       base += kWavesPerBlock * kScanRows) {
    const uint32_t r0 = base + lane, r1 = base + 64 + lane;
    const uint32_t c0 = min(r0, a.st.nrows - 1), c1 = min(r1, a.st.nrows - 1);
    const double* __restrict__ x0p = a.st.X + (size_t)c0 * a.st.ldx;
    const double* __restrict__ x1p = a.st.X + (size_t)c1 * a.st.ldx;
    double acc[kScanTQ][2];
#pragma unroll
    for (int j = 0; j < kScanTQ; j++) acc[j][0] = acc[j][1] = 0.0;

    uint32_t c = 0;
    if (full_group) {
      for (; c + 8 <= a.D; c += 8) {
        double x0[8], x1[8];
#pragma unroll
        for (int i = 0; i < 8; i += 2) {
          double2 v0 = *reinterpret_cast<const double2*>(x0p + c + i);
          double2 v1 = *reinterpret_cast<const double2*>(x1p + c + i);
          x0[i] = v0.x; x0[i + 1] = v0.y;
          x1[i] = v1.x; x1[i + 1] = v1.y;
        }
        accum8<METRIC>(acc, x0, x1, Qg, a.ldq, c);
      }
    }
    for (; c < a.D; c++) {
      double xa = x0p[c], xb = x1p[c];
#pragma unroll
      for (int j = 0; j < kScanTQ; j++) {
        double q = a.Q[(size_t)min(q0 + j, qlast) * a.ldq + c];
        if (METRIC == kL2) {
          double t0 = xa - q, t1 = xb - q;
          acc[j][0] = fma(t0, t0, acc[j][0]);
          acc[j][1] = fma(t1, t1, acc[j][1]);
        } else {
          acc[j][0] = fma(q, xa, acc[j][0]);
          acc[j][1] = fma(q, xb, acc[j][1]);
        }
      }
    }

    const bool v0 = r0 < r_end && slot_alive(a.st.alive, c0);
    const bool v1 = r1 < r_end && slot_alive(a.st.alive, c1);
    double xn0 = 0.0, xn1 = 0.0;
    if (METRIC == kCos) {
      xn0 = a.st.xn[c0];
      xn1 = a.st.xn[c1];
    }
    // lane j: query q0 + j's global bound (the smallest k-th key any list
    // of the query has published; rows with keys above it are out)
    uint64_t gk_l = ~0ull;
    if (a.gtk && lane < kScanTQ)
      gk_l = __hip_atomic_load(a.gtk + min(q0 + lane, qlast), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int j = 0; j < kScanTQ; j++) {
      uint64_t tkj = read_lane64(tk_l, j);
      uint32_t tsj = __builtin_amdgcn_readlane(ts_l, j);
      const int kqj = __builtin_amdgcn_readlane(kq_l, j);
      const uint64_t gkj = read_lane64(gk_l, j);
      double s0 = acc[j][0], s1 = acc[j][1];
      if (METRIC == kDot) {
        s0 = -s0;
        s1 = -s1;
      } else if (METRIC == kCos) {
        const double qnj = read_lane_f64(qn_l, j);
        cos_keys(s0, qnj * xn0, s1, qnj * xn1, min(tkj, gkj), s0, s1);
      }
      const uint64_t k0 = rank_bits(s0), k1 = rank_bits(s1);
      // slots past nq (a partial group) keep no list
      const bool jv = q0 + j < a.nq;
      uint64_t m0 = __ballot(jv && v0 && k0 <= gkj && lex_lt(k0, r0, tkj, tsj));
      uint64_t m1 = __ballot(jv && v1 && k1 <= gkj && lex_lt(k1, r1, tkj, tsj));
      // After each insertion the k-th entry may have moved: lanes that no
      // longer beat it drop out (their insertion would change nothing).
      const bool moved = (m0 | m1) != 0;
      while (m0) {
        int src = __ffsll((long long)m0) - 1;
        m0 &= m0 - 1;
        L[j].insert(uniform64(__shfl(k0, src)), base + src, kqj, lane);
        L[j].kth(kqj, tkj, tsj);
        m0 &= __ballot(lex_lt(k0, r0, tkj, tsj));
      }
      m1 &= __ballot(lex_lt(k1, r1, tkj, tsj));
      while (m1) {
        int src = __ffsll((long long)m1) - 1;
        m1 &= m1 - 1;
        L[j].insert(uniform64(__shfl(k1, src)), base + 64 + src, kqj, lane);
        L[j].kth(kqj, tkj, tsj);
        m1 &= __ballot(lex_lt(k1, r1, tkj, tsj));
      }
      if (moved && lane == j) {
        tk_l = tkj;
        ts_l = tsj;
        // a full list publishes its k-th key
        if (a.gtk && tkj != ~0ull && tkj < gk_l)
          __hip_atomic_fetch_min(a.gtk + q0 + j, tkj, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }

  const uint32_t list = slice * kWavesPerBlock + wave;
#pragma unroll
  for (int j = 0; j < kScanTQ; j++) {
    if (q0 + j >= a.nq) break;
    const int kqj = __builtin_amdgcn_readlane(kq_l, j);
    size_t o = ((size_t)(q0 + j) * a.P + list) * a.kstride;
#pragma unroll
    for (int p = 0; p < KPL; p++) {
      int e = p * 64 + lane;
      if (e < kqj) {
        a.part_key[o + e] = L[j].k[p];
        a.part_slot[o + e] = L[j].s[p];
      }
    }
  }
}

// ------------------------------------------------- FP64 scan on MFMA
//
// The FP64 scan on the matrix cores.  v_mfma_f64_16x16x4f64 adds the four
// products of a k-step onto each accumulator one after the other, each as
// one correctly rounded fused multiply-add (the same bits as sequential
// fma chains, subnormal, infinite and NaN operands included), so chained
// over the k-steps in column order it computes exactly the canonical dot
// key s = Σ_i fma(q_i, x_i, s); cosine divides it as the VALU scan does.
// Euclidean keys, Σ fma(t, t, s) with t = x_i − q_i, have no matrix form:
// the dot product gives a = (xx + qq) − 2·dot (xx, qq the canonical
// squared norms), within E = c·(‖x‖ + ‖q‖)² + e of the canonical key
// (l2_bound), and only a pair whose lower bound a − E could still enter
// its query's list computes the canonical key (exact_l2); every other pair
// is provably outside the list.  Candidates wait in a per-wavefront LDS
// queue and are keyed 64 at a time, one pair per lane (a lane at a time
// had left the wavefront serial in those chains: Euclidean k = 10 10.4 →
// 7.3 ms, k = 100 142 → 16.7 ms on 1M × 128 for 1,024 queries).
//
// Block: 4 wavefronts, 16 queries each (64 per block), sharing the rows of
// one slice, which stream through LDS in stages of 64 rows × 32 columns
// (16 KB, global_load_lds_dwordx4, double-buffered, so the next stage's
// loads run under this stage's MFMAs).  16-byte units are XOR-swizzled by
// row so the operand reads of 16 rows hit distinct banks.  Lane l,
// g = l / 16, j = l % 16; k-step s of a stage covers its columns
// 4s..4s+3:
//   A (queries)  lane l holds Q[qw + j][col0 + 4s + g]   (global loads)
//   B (rows)     lane l holds X[row0 + 16t + j][col0 + 4s + g]  (LDS)
//   accumulator  tile t, register r of lane l: query qw + g + 4r, row
//                row0 + 16t + j
// After a row step's last stage each (query, row) key is tested against
// its query's k-th best (the wave-distributed lists of k_scan_f64) and
// against the query's global bound gtk: the smallest k-th key any list of
// the query has published (atomic min).  That list holds k rows with keys
// at most gtk, so a row whose key is above gtk is not among the query's k
// nearest; a key equal to it still goes to the list, whose (key, slot)
// order settles the tie.  Lists then see far fewer insertions than a list
// bounded by its own slice alone.
// Grid: (groups of 64 queries) × slices, in k_scan_f64's XCD order; each
// block writes one partial list per query (P = number of slices).

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void_t;
constexpr int kF64StageRows = 64;
constexpr int kF64StageCols = 32;
constexpr int kF64StageBytes = kF64StageRows * kF64StageCols * 8;

// Byte offset in a stage of 16-byte unit u (columns 2u, 2u+1) of row r.
__device__ __forceinline__ int f64_unit(int r, int u) { return (r * 16 + (u ^ (r & 15))) * 16; }

// The stage of rows row0.. (clamped to the store) and columns col0..col0+31:
// 16 direct-to-LDS instructions of 64 × 16 bytes, four per wavefront.
__device__ __forceinline__ void f64_issue_stage(const StoreView& st, unsigned char* buf,
                                                uint32_t row0, uint32_t col0, int wave,
                                                int lane) {
#pragma unroll
  for (int i = 0; i < kF64StageBytes / (1024 * kWavesPerBlock); i++) {
    const int p0 = (wave + kWavesPerBlock * i) * 64;  // first unit slot of this instruction
    const int p = p0 + lane;
    const int r = p >> 4, u = (p & 15) ^ (r & 15);
    const uint32_t row = min(row0 + (uint32_t)r, st.nrows - 1);
    __builtin_amdgcn_global_load_lds(st.X + (size_t)row * st.ldx + col0 + 2 * u,
                                     (lds_void_t*)(buf + p0 * 16), 16, 0, 0);
  }
}

// The canonical Euclidean key of row x and query q over D columns
// (D a multiple of 8, x 16-byte aligned): Σ_i fma(t, t, s), t = x_i − q_i.
__device__ __forceinline__ double exact_l2(const double* __restrict__ x,
                                          const double* __restrict__ q, uint32_t D) {
  double s = 0.0;
  for (uint32_t c = 0; c < D; c += 8) {
    double xv[8], qv[8];
#pragma unroll
    for (int i = 0; i < 8; i += 2) {
      const double2 v = *reinterpret_cast<const double2*>(x + c + i);
      xv[i] = v.x;
      xv[i + 1] = v.y;
    }
#pragma unroll
    for (int i = 0; i < 8; i++) qv[i] = q[c + i];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const double t = xv[i] - qv[i];
      s = fma(t, t, s);
    }
  }
  return s;
}

// Bound on |a − s*| for a = fl(fl(xx + qq) − 2·dot) and the canonical key
// s*, with xx, qq, dot and s* the canonical fma chains over D columns and
// u = 2^-53.  Each chain is within γ_D = D·u/(1 − D·u) of its exact value
// relative to its sum of |terms| (xx: ‖x‖², qq: ‖q‖², dot: ‖q‖‖x‖), the two
// roundings of a add at most 2u(1 + γ_D)·S with S = (‖x‖ + ‖q‖)², and the
// rounded differences t_i change ‖x − q‖² ≤ S by at most 2u + u², so
// |a − s*| ≤ (2γ_D + 4u + O(u²))·S, and the final a − E rounds by at most
// 2u·S more.  E = (4D + 128)·u·sn² with sn = fl(sqrt(xx)) + fl(sqrt(qq))
// (≥ (‖x‖ + ‖q‖)(1 − γ_D − 3u)) is more than twice that for any D the
// store allows, and e = (8D + 64)·2^-1074 covers the absolute error of
// subnormal results (at most one 2^-1074 per rounding, 5D + 4 of them).
// A non-finite operand gives a non-finite bound: the pair takes the
// exact path.
__device__ __forceinline__ double l2_lower(double xx, double qq, double dot, double xs,
                                          double qs, double crel, double eabs) {
  const double ap = (xx + qq) - 2.0 * dot;
  const double sn = xs + qs;
  const double E = crel * (sn * sn) + eabs;
  return ap - E;
}

template <int METRIC, int KPL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void
k_scan_f64_mfma(ScanF64Args a) {
  constexpr int NT = kF64StageRows / 16;
  constexpr int KS = kF64StageCols / 4;
  __shared__ __attribute__((aligned(16))) unsigned char buf0[kF64StageBytes];
  __shared__ __attribute__((aligned(16))) unsigned char buf1[kF64StageBytes];
  // Each wavefront's queue of candidate pairs (slot << 8 | query of the
  // wavefront; dot and cosine also the key): Euclidean queues a whole row
  // step (up to 1,024) on top of a part batch, dot and cosine (k <= 64)
  // drain it whenever it holds a batch, so 127 entries at most
  constexpr int QCAP = METRIC == kL2 ? 64 + 16 * 16 * NT : (KPL == 1 ? 128 : 1);
  __shared__ uint64_t cq[kWavesPerBlock][QCAP];
  __shared__ uint64_t cqk[kWavesPerBlock][METRIC == kL2 ? 1 : QCAP];
  // Per-row values of a row step (‖x‖, Σx², the alive words), staged
  // through LDS with the stage loads one row step ahead (two row steps in
  // flight, by parity), so the epilogue reads them from LDS instead of
  // waiting on global memory
  __shared__ __attribute__((aligned(16))) double rowv[2][2][KPL == 1 ? kF64StageRows : 1];
  __shared__ __attribute__((aligned(16))) uint32_t alivew[2][4];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane >> 4, jr = lane & 15;
  constexpr uint32_t QB = kScanTQ * kWavesPerBlock;
  const uint32_t ngroups = (a.nq + QB - 1) / QB;
  const uint32_t xcd = blockIdx.x % 8, local = blockIdx.x / 8;
  const uint32_t qw = (local % ngroups) * QB + wave * kScanTQ;  // this wave's first query
  const uint32_t slice = (local / ngroups) * 8 + xcd;
  const uint32_t r_begin = slice * a.rows_per_slice;
  const uint32_t r_end = min(r_begin + a.rows_per_slice, a.st.nrows);
  const uint32_t nsteps = r_end > r_begin ? (r_end - r_begin + kF64StageRows - 1) / kF64StageRows : 0;
  const uint32_t nchunks = a.D / kF64StageCols;
  const uint32_t nstages = nsteps * nchunks;
  const uint32_t qlast = a.nq - 1;
  const bool active = qw < a.nq;  // wave-uniform: a wave past nq only helps load
  const double* __restrict__ qp = a.Q + (size_t)min(qw + jr, qlast) * a.ldq;

  // Lane j holds query qw + j's current k-th best, k and norm.
  uint64_t tk_l = ~0ull;
  uint32_t ts_l = ~0u;
  int kq_l = 0;
  double qn_l = 0.0, qq_l = 0.0;
  if (lane < kScanTQ) {
    const uint32_t qi = min(qw + lane, qlast);
    kq_l = a.k[qi];
    qn_l = METRIC != kDot ? a.qn[qi] : 0.0;
    if (METRIC == kL2) qq_l = a.qq[qi];
  }
  // the k-th best of this lane's accumulator queries g + 4r (filter)
  uint64_t thk[4];
  uint32_t ths[4];
  double qn4[4], qq4[4];
  bool qv4[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    thk[r] = ~0ull;
    ths[r] = ~0u;
    qn4[r] = __shfl(qn_l, g + 4 * r);
    qq4[r] = __shfl(qq_l, g + 4 * r);
    qv4[r] = qw + g + 4 * r < a.nq;
  }
  // Euclidean: the constants of l2_lower
  const double crel = (4.0 * a.D + 128.0) * 0x1p-53;
  const double eabs = (8.0 * a.D + 64.0) * 0x1p-1074;
  WList<KPL> L[kScanTQ];
#pragma unroll
  for (int j = 0; j < kScanTQ; j++) L[j].init();
  // lane j: query qw + j's global bound, read once per row step
  uint64_t gk_l = ~0ull, gk_next = ~0ull;

  f64x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) acc[t] = (f64x4){0.0, 0.0, 0.0, 0.0};
  double qa[KS], qn[KS];
  auto load_q = [&](uint32_t col0) {
#pragma unroll
    for (int s = 0; s < KS; s++) qn[s] = qp[col0 + 4 * s + g];
  };

  // n queued pairs from cq[wave][start], one per lane: the key (Euclidean:
  // the canonical fma chain; dot and cosine: queued with the pair), then
  // the lists — for each query, the lanes whose pair beats its k-th best
  // insert one at a time, the k-th best re-read after each.
  uint32_t qcount = 0;
  auto run_batch = [&](uint32_t start, uint32_t n) __attribute__((always_inline)) {
    const bool v = (uint32_t)lane < n;
    const uint64_t e = v ? cq[wave][start + lane] : 0ull;
    const uint32_t srow = (uint32_t)(e >> 8);
    const int jq = (int)(e & 255u);
    uint64_t kk = ~0ull;
    if constexpr (METRIC == kL2) {
      if (v)
        kk = rank_bits(exact_l2(a.st.X + (size_t)srow * a.st.ldx,
                                a.Q + (size_t)min(qw + (uint32_t)jq, qlast) * a.ldq, a.D));
    } else {
      if (v) kk = cqk[wave][start + lane];
    }
#pragma unroll
    for (int j = 0; j < kScanTQ; j++) {
      if (qw + j >= a.nq) break;
      uint64_t tkj = read_lane64(tk_l, j);
      uint32_t tsj = __builtin_amdgcn_readlane(ts_l, j);
      const int kqj = __builtin_amdgcn_readlane(kq_l, j);
      const uint64_t gkj = read_lane64(gk_l, j);
      uint64_t m = __ballot(v && jq == j && kk <= gkj && lex_lt(kk, srow, tkj, tsj));
      if (!m) continue;
      while (m) {
        const int src = __ffsll((long long)m) - 1;
        m &= m - 1;
        L[j].insert(uniform64(__shfl(kk, src)), uniform32(__shfl(srow, src)), kqj, lane);
        L[j].kth(kqj, tkj, tsj);
        m &= __ballot(v && jq == j && lex_lt(kk, srow, tkj, tsj));
      }
      if (lane == j) {
        tk_l = tkj;
        ts_l = tsj;
        if (a.gtk && tkj != ~0ull && tkj < gk_l)
          __hip_atomic_fetch_min(a.gtk + qw + j, tkj, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      if (g == (j & 3)) {
        thk[j >> 2] = tkj;
        ths[j >> 2] = tsj;
      }
    }
  };

  // The per-row values of the row step at base into rowv / alivew[par]:
  // wave 0 the norms, wave 1 the squared norms (4-byte direct-to-LDS loads,
  // lane l: dword l&1 of row 32i + l/2, rows clamped to the store), wave 2
  // the (at most two) alive words the 64 rows span.
  // (one list register per lane only: at two the staging's registers cost
  // spills, and the row step's loads stay direct)
  constexpr bool ROWV = KPL == 1;
  auto issue_rowvals = [&](int par, uint32_t base) __attribute__((always_inline)) {
    if constexpr (!ROWV) return;
    const uint32_t last = a.st.nrows - 1;
    if ((wave == 0 && METRIC != kDot) || (wave == 1 && METRIC == kL2)) {
      const double* src = wave == 0 ? a.st.xn : a.st.nx2;
#pragma unroll
      for (int i = 0; i < 2; i++) {
        const uint32_t row = min(base + 32 * i + (lane >> 1), last);
        __builtin_amdgcn_global_load_lds(reinterpret_cast<const uint32_t*>(src + row) + (lane & 1),
                                         (lds_void_t*)&rowv[par][wave][32 * i], 4, 0, 0);
      }
    } else if (wave == 2 && lane < 4) {
      const uint32_t w = min((base >> 6) + (lane >> 1), last >> 6);
      __builtin_amdgcn_global_load_lds(reinterpret_cast<const uint32_t*>(a.st.alive + w) + (lane & 1),
                                       (lds_void_t*)&alivew[par][0], 4, 0, 0);
    }
  };

  // Per-row-step epilogue: keys of the step's 64 rows, list insertions.
  auto epilogue = [&](uint32_t base, int par) __attribute__((always_inline)) {
    uint32_t rowi[NT];
    bool rv[NT];
    double xn[NT], xx[NT];
    const uint32_t w0 = base >> 6;
#pragma unroll
    for (int t = 0; t < NT; t++) {
      rowi[t] = base + 16 * t + jr;
      if constexpr (ROWV) {
        const uint32_t wi = (rowi[t] >> 6) - w0;  // 0 or 1
        const uint64_t word = ((uint64_t)alivew[par][2 * wi + 1] << 32) | alivew[par][2 * wi];
        rv[t] = rowi[t] < r_end && ((word >> (rowi[t] & 63)) & 1ull);
        xn[t] = METRIC != kDot ? rowv[par][0][16 * t + jr] : 0.0;
        xx[t] = METRIC == kL2 ? rowv[par][1][16 * t + jr] : 0.0;
      } else {
        const uint32_t slot = min(rowi[t], a.st.nrows - 1);
        rv[t] = rowi[t] < r_end && slot_alive(a.st.alive, slot);
        xn[t] = METRIC != kDot ? a.st.xn[slot] : 0.0;
        xx[t] = METRIC == kL2 ? a.st.nx2[slot] : 0.0;
      }
    }
    gk_l = gk_next;
    uint64_t gk4[4];
#pragma unroll
    for (int r = 0; r < 4; r++) gk4[r] = __shfl(gk_l, g + 4 * r);
    if constexpr (METRIC == kL2) {
      // Candidates — pairs whose lower bound could still enter the list —
      // join the wavefront's queue; full batches of 64 get their canonical
      // keys one pair per lane, all lanes at once, and go to the lists.
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const double lb = l2_lower(xx[t], qq4[r], acc[t][r], xn[t], qn4[r], crel, eabs);
          const uint64_t kb = rank_bits(lb);
          const bool c = qv4[r] && rv[t] &&
                         (!(lb > -INFINITY && lb < INFINITY) || (kb <= thk[r] && kb <= gk4[r]));
          const uint64_t m = __ballot(c);
          if (m) {
            if (c)
              cq[wave][qcount + __popcll(m & ((1ull << lane) - 1ull))] =
                  ((uint64_t)rowi[t] << 8) | (uint64_t)(g + 4 * r);
            qcount += __popcll(m);
          }
        }
      while (qcount >= 64) {
        qcount -= 64;
        run_batch(qcount, 64);
      }
      return;
    }
    // dot and cosine: the keys are exact already.  With one list register
    // per lane (k <= 64), pairs that beat their query's k-th best (and its
    // global bound) join the queue, drained a batch of 64 at a time (1M ×
    // 128, 1,024 queries, k = 10: dot 7.8 → 6.6 ms, cosine 8.1 → 7.2); with
    // two, the registers the queue path needs cost spills and the pairs go
    // to the lists directly (k = 100: 12.8 ms against 13.8 queued).
    if constexpr (KPL == 1) {
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const double sc = METRIC == kCos ? -(acc[t][r] / (qn4[r] * xn[t])) : -acc[t][r];
          const uint64_t kt = rank_bits(sc);
          const bool c = qv4[r] && rv[t] && kt <= gk4[r] && lex_lt(kt, rowi[t], thk[r], ths[r]);
          const uint64_t m = __ballot(c);
          if (m) {
            if (c) {
              const uint32_t pos = qcount + __popcll(m & ((1ull << lane) - 1ull));
              cq[wave][pos] = ((uint64_t)rowi[t] << 8) | (uint64_t)(g + 4 * r);
              cqk[wave][pos] = kt;
            }
            qcount += __popcll(m);
            if (qcount >= 64) {
              qcount -= 64;
              run_batch(qcount, 64);
            }
          }
        }
    } else {
      uint64_t key[NT][4];
      bool hit = false;
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const double sc = METRIC == kCos ? -(acc[t][r] / (qn4[r] * xn[t])) : -acc[t][r];
          key[t][r] = rank_bits(sc);
          hit |= qv4[r] && rv[t] && key[t][r] <= gk4[r] &&
                 lex_lt(key[t][r], rowi[t], thk[r], ths[r]);
        }
      if (__builtin_expect(__ballot(hit) == 0, 1)) return;
#pragma unroll
      for (int j = 0; j < kScanTQ; j++) {
        const int r = j >> 2;
        const bool mine = g == (j & 3);
        if (qw + j >= a.nq) break;
        uint64_t tkj = read_lane64(tk_l, j);
        uint32_t tsj = __builtin_amdgcn_readlane(ts_l, j);
        const int kqj = __builtin_amdgcn_readlane(kq_l, j);
        const uint64_t gkj = read_lane64(gk_l, j);
        bool moved = false;
#pragma unroll
        for (int t = 0; t < NT; t++) {
          uint64_t m = __ballot(mine && rv[t] && key[t][r] <= gkj &&
                                lex_lt(key[t][r], rowi[t], tkj, tsj));
          while (m) {
            const int src = __ffsll((long long)m) - 1;
            m &= m - 1;
            L[j].insert(uniform64(__shfl(key[t][r], src)), base + 16 * t + (src & 15), kqj,
                        lane);
            L[j].kth(kqj, tkj, tsj);
            m &= __ballot(mine && lex_lt(key[t][r], rowi[t], tkj, tsj));
            moved = true;
          }
        }
        if (moved) {
          if (lane == j) {
            tk_l = tkj;
            ts_l = tsj;
            // a full list publishes its k-th key
            if (a.gtk && tkj != ~0ull && tkj < gk_l)
              __hip_atomic_fetch_min(a.gtk + qw + j, tkj, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
          }
          if (mine) {
            thk[r] = tkj;
            ths[r] = tsj;
          }
        }
      }
    }
  };

  if (nstages > 0) {
    f64_issue_stage(a.st, buf0, r_begin, 0, wave, lane);
    issue_rowvals(0, r_begin);
    load_q(0);
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  // One stage: issue the next stage's loads into `nxt`, MFMAs over `cur`,
  // the row step's epilogue after its last stage, then wait for the loads
  // and for every wave to be done with `cur`.  Stages alternate between the
  // two buffers as distinct objects known at compile time, so reads of one
  // need not wait for the direct-to-LDS loads still filling the other.
  auto stage = [&](const unsigned char* cur, unsigned char* nxt, uint32_t st)
                   __attribute__((always_inline)) {
    const uint32_t step = st / nchunks, chunk = st % nchunks;
    // the global bounds for this row step's epilogue, read at its first
    // stage so the load's latency runs under the step's MFMAs
    if (chunk == 0 && a.gtk && lane < kScanTQ)
      gk_next = __hip_atomic_load(a.gtk + min(qw + lane, qlast), __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int s = 0; s < KS; s++) qa[s] = qn[s];
    if (st + 1 < nstages) {
      const uint32_t nstep = (st + 1) / nchunks, nchunk = (st + 1) % nchunks;
      f64_issue_stage(a.st, nxt, r_begin + nstep * kF64StageRows, nchunk * kF64StageCols, wave,
                      lane);
      if (nchunk == 0) issue_rowvals(nstep & 1, r_begin + nstep * kF64StageRows);
      load_q(nchunk * kF64StageCols);
    }
    if (active) {
#pragma unroll
      for (int s = 0; s < KS; s++) {
        // B operand of tile t: row 16t + j, column 4s + g = unit 2s + g/2, half g&1
        double b[NT];
#pragma unroll
        for (int t = 0; t < NT; t++)
          b[t] = *reinterpret_cast<const double*>(cur + f64_unit(16 * t + jr, 2 * s + (g >> 1)) +
                                                  (g & 1) * 8);
#pragma unroll
        for (int t = 0; t < NT; t++)
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(qa[s], b[t], acc[t], 0, 0, 0);
      }
      if (chunk + 1 == nchunks) {
        epilogue(r_begin + step * kF64StageRows, step & 1);
#pragma unroll
        for (int t = 0; t < NT; t++) acc[t] = (f64x4){0.0, 0.0, 0.0, 0.0};
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  };
  for (uint32_t st0 = 0; st0 < nstages; st0 += 2) {
    stage(buf0, buf1, st0);
    if (st0 + 1 < nstages) stage(buf1, buf0, st0 + 1);
  }
  if (active && qcount) run_batch(0, qcount);

  if (!active) return;
  const uint32_t list = slice;
#pragma unroll
  for (int j = 0; j < kScanTQ; j++) {
    if (qw + j >= a.nq) break;
    const int kqj = __builtin_amdgcn_readlane(kq_l, j);
    size_t o = ((size_t)(qw + j) * a.P + list) * a.kstride;
#pragma unroll
    for (int p = 0; p < KPL; p++) {
      int e = p * 64 + lane;
      if (e < kqj) {
        a.part_key[o + e] = L[j].k[p];
        a.part_slot[o + e] = L[j].s[p];
      }
    }
  }
}

// Euclidean needs the squared row norms over the compared columns (all of
// them); a column prefix (kd-tree dimension below the row width) stays on
// the VALU scan, as do widths that are not whole stages.
bool scan_f64_uses_mfma(int metric, uint32_t D, const StoreView& st, int kmax, uint32_t nq) {
  (void)nq;
  (void)kmax;
  return D > 0 && D % kF64StageCols == 0 && st.ldx % 2 == 0 && ((uintptr_t)st.X & 15) == 0 &&
         (metric != kL2 || st.nx2 != nullptr);
}

// The matrix-core scan takes one list per (query, slice), with slices for
// about 512 blocks up to k = 32 (two per CU, one round: fewer, longer lists
// fill fewer times — 1M × 128, 1,024 queries, k = 10: 32 slices, dot 8.25
// → 7.67 ms against 128) and about 1,024 above (k = 100: 12.7 ms at 64
// slices, 13.5 at 128, 17.3 at 16 — the block's insertions dominating), at
// least 64 rows a slice (a one-query call over 10,000 rows: 8 slices of
// 1,250 rows took 193 µs per call, 64 slices 88 µs), and no more lists than
// the scratch holds (nslices · kWavesPerBlock per query).
uint32_t scan_f64_lists(int metric, uint32_t D, const StoreView& st, int kmax, uint32_t nq,
                        uint32_t nslices) {
  if (!scan_f64_uses_mfma(metric, D, st, kmax, nq)) return nslices * kWavesPerBlock;
  const uint32_t groups = (nq + kScanTQ * kWavesPerBlock - 1) / (kScanTQ * kWavesPerBlock);
  const uint32_t blocks = kmax <= 32 ? 512 : 1024;
  uint32_t P = (blocks + groups - 1) / groups;
  P = std::min<uint32_t>(P, (st.nrows + 63) / 64);
  P = std::min<uint32_t>(P, nslices * kWavesPerBlock);
  return std::max<uint32_t>(8, (P + 7) / 8 * 8);
}

template <int METRIC, int KPL>
static hipError_t scan_f64_t(const ScanF64Args& a, uint32_t nslices, hipStream_t s) {
  // nslices is a multiple of 8 (one set of slices per XCD); 1-D grid
  if (nslices % 8) return hipErrorInvalidValue;
  if (a.P != scan_f64_lists(METRIC, a.D, a.st, (int)a.kstride, a.nq, nslices))
    return hipErrorInvalidValue;
  // On the matrix cores when the width is whole stages and rows take
  // 16-byte loads (Euclidean: all columns, with the squared norms at
  // hand): one block per (64 queries, slice), one partial list per query
  // and slice.
  if (scan_f64_uses_mfma(METRIC, a.D, a.st, (int)a.kstride, a.nq)) {
    if (METRIC == kL2 && !a.qq) return hipErrorInvalidValue;
    constexpr uint32_t QB = kScanTQ * kWavesPerBlock;
    ScanF64Args m = a;
    m.rows_per_slice = (a.st.nrows + a.P - 1) / a.P;
    dim3 grid(((a.nq + QB - 1) / QB) * a.P);
    hipLaunchKernelGGL((k_scan_f64_mfma<METRIC, KPL>), grid, dim3(256), 0, s, m);
    return hipGetLastError();
  }
  dim3 grid(((a.nq + kScanTQ - 1) / kScanTQ) * nslices);
  hipLaunchKernelGGL((k_scan_f64<METRIC, KPL>), grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_scan_f64(const ScanF64Args& a, int metric, int kmax, uint32_t nslices,
                           hipStream_t s) {
  if (a.nq == 0) return hipSuccess;
  const bool two = kmax > 64;
  switch (metric) {
    case kL2: return two ? scan_f64_t<kL2, 2>(a, nslices, s) : scan_f64_t<kL2, 1>(a, nslices, s);
    case kCos: return two ? scan_f64_t<kCos, 2>(a, nslices, s) : scan_f64_t<kCos, 1>(a, nslices, s);
    case kDot: return two ? scan_f64_t<kDot, 2>(a, nslices, s) : scan_f64_t<kDot, 1>(a, nslices, s);
  }
  return hipErrorInvalidValue;
}

// ------------------------------------------------ one query, nearest row
//
// kdtree_nearest called once per request (the reference's handlers do) on a
// store small enough to scan in microseconds: the launch, not the scan, is
// the cost, so the whole answer is one kernel — no query upload (the query
// travels in the kernel arguments), no merge launch, no result copy (the
// last block writes the id to host memory).  The blocks' results meet
// through a ticket: every block publishes its (key, slot) and takes a
// ticket (release/acquire at agent scope); the block holding the last
// ticket reads them all.

__device__ __forceinline__ void min_pair(uint64_t& bk, uint32_t& bs, uint64_t k, uint32_t s) {
  if (lex_lt(k, s, bk, bs)) {
    bk = k;
    bs = s;
  }
}

__global__ __launch_bounds__(256) void k_nearest1(const Nearest1Args a) {
  __shared__ uint64_t wk[4];
  __shared__ uint32_t ws[4];
  __shared__ int last;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t bk = ~0ull;
  uint32_t bs = ~0u;
  for (uint32_t r = blockIdx.x * 256 + threadIdx.x; r < a.st.nrows; r += gridDim.x * 256) {
    if (!slot_alive(a.st.alive, r)) continue;
    const double* __restrict__ x = a.st.X + (size_t)r * a.st.ldx;
    double s = 0.0;
    for (uint32_t i = 0; i < a.D; i++) {
      const double t = x[i] - a.q[i];
      s = fma(t, t, s);
    }
    min_pair(bk, bs, rank_bits(s), r);
  }
  auto block_min = [&]() {
    for (int off = 32; off >= 1; off >>= 1) {
      const uint64_t ok = (uint64_t)__shfl_xor((long long)bk, off);
      const uint32_t os = (uint32_t)__shfl_xor((int)bs, off);
      min_pair(bk, bs, ok, os);
    }
    if (lane == 0) {
      wk[wave] = bk;
      ws[wave] = bs;
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int w = 1; w < 4; w++) min_pair(bk, bs, wk[w], ws[w]);
  };
  block_min();
  if (threadIdx.x == 0) {
    __hip_atomic_store(a.part_key + blockIdx.x, bk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(a.part_slot + blockIdx.x, bs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t t =
        __hip_atomic_fetch_add(a.ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = t == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  bk = ~0ull;
  bs = ~0u;
  for (uint32_t b = threadIdx.x; b < gridDim.x; b += 256)
    min_pair(bk, bs, __hip_atomic_load(a.part_key + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
             __hip_atomic_load(a.part_slot + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  __syncthreads();  // wk / ws are reused
  block_min();
  if (threadIdx.x == 0) {
    __hip_atomic_store(a.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    a.out_id[0] = bs == ~0u ? ~0ull : logical_id(a.st, bs);
    __threadfence_system();
    __hip_atomic_store(a.out_id + 1, a.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

hipError_t launch_nearest1(const Nearest1Args& a, hipStream_t s) {
  if (a.D == 0 || a.D > kNearestMaxD || a.st.nrows == 0) return hipErrorInvalidValue;
  const uint32_t blocks =
      std::min<uint32_t>(kNearestMaxBlocks, (a.st.nrows + 255) / 256);
  hipLaunchKernelGGL(k_nearest1, dim3(blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

// ---------------------------------------------------------- list merge
//
// One wavefront per query merges the query's P sorted partial lists into
// its final top-k, then writes logical ids and FP64 scores.  A partial list
// whose first entry does not beat the current k-th best is skipped whole.

template <int METRIC, int KPL>
__global__ __launch_bounds__(256) void k_merge(MergeArgs a) {
  const int lane = threadIdx.x & 63;
  const uint32_t q = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (q >= a.nq) return;
  const int kk = a.k[q];
  WList<KPL> L;
  L.init();
  uint64_t tk = ~0ull;
  uint32_t ts = ~0u;
  for (uint32_t p = 0; p < a.P; p++) {
    const size_t o = ((size_t)q * a.P + p) * a.kstride;
    const uint64_t e0k = uniform64(a.part_key[o]);
    const uint32_t e0s = uniform32(a.part_slot[o]);
    if (!lex_lt(e0k, e0s, tk, ts)) continue;
#pragma unroll
    for (int pp = 0; pp < KPL; pp++) {
      const int e = pp * 64 + lane;
      uint64_t ck = ~0ull;
      uint32_t cs = ~0u;
      if (e < kk) {
        ck = a.part_key[o + e];
        cs = a.part_slot[o + e];
      }
      uint64_t m = __ballot(lex_lt(ck, cs, tk, ts));
      while (m) {
        int src = __ffsll((long long)m) - 1;
        m &= m - 1;
        L.insert(uniform64(__shfl(ck, src)), uniform32(__shfl(cs, src)), kk, lane);
        L.kth(kk, tk, ts);
        m &= __ballot(lex_lt(ck, cs, tk, ts));
      }
    }
    L.kth(kk, tk, ts);
  }
#pragma unroll
  for (int pp = 0; pp < KPL; pp++) {
    const int e = pp * 64 + lane;
    if (e < kk) {
      uint64_t id = ~0ull;
      double sc = __longlong_as_double(0x7FF8000000000000ll);
      if (L.s[pp] != ~0u) {
        double v = rank_value(L.k[pp]);
        // NaN keys come back as the one canonical quiet NaN (its sign is
        // not meaningful); a zero dot/cos rank key is a +0 score.
        if (v == v) sc = METRIC == kL2 ? sqrt(v) : (v == 0.0 ? 0.0 : -v);
        id = logical_id(a.st, L.s[pp]);
      }
      a.ids[(size_t)q * a.ostride + e] = id;
      a.scores[(size_t)q * a.ostride + e] = sc;
    }
  }
}

template <int METRIC, int KPL>
static hipError_t merge_t(const MergeArgs& a, hipStream_t s) {
  dim3 grid((a.nq + kWavesPerBlock - 1) / kWavesPerBlock);
  hipLaunchKernelGGL((k_merge<METRIC, KPL>), grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_merge(const MergeArgs& a, int metric, int kmax, hipStream_t s) {
  if (a.nq == 0) return hipSuccess;
  const bool two = kmax > 64;
  switch (metric) {
    case kL2: return two ? merge_t<kL2, 2>(a, s) : merge_t<kL2, 1>(a, s);
    case kCos: return two ? merge_t<kCos, 2>(a, s) : merge_t<kCos, 1>(a, s);
    case kDot: return two ? merge_t<kDot, 2>(a, s) : merge_t<kDot, 1>(a, s);
  }
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------- pair metrics
//
// FP64: one wavefront per pair, coalesced loads; lane l accumulates the
// elements i ≡ l (mod 64) in increasing order with fma, then six xor
// butterfly steps (offsets 32..1).  Every lane ends with the same sum.

template <int METRIC>
__global__ __launch_bounds__(256) void k_pair_f64(const double* __restrict__ X, size_t ldx,
                                                  const uint32_t* __restrict__ sa,
                                                  const uint32_t* __restrict__ sb, uint32_t np,
                                                  uint32_t D, double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const uint32_t p = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (p >= np) return;
  const double* __restrict__ A = X + (size_t)sa[p] * ldx;
  const double* __restrict__ B = X + (size_t)sb[p] * ldx;
  double d = 0.0, na = 0.0, nb = 0.0;
  for (uint32_t i = lane; i < D; i += 64) {
    double u = A[i], v = B[i];
    if (METRIC == kL2) {
      double t = u - v;
      d = fma(t, t, d);
    } else {
      d = fma(u, v, d);
      if (METRIC == kCos) {
        na = fma(u, u, na);
        nb = fma(v, v, nb);
      }
    }
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    d = d + __shfl_xor(d, off);
    if (METRIC == kCos) {
      na = na + __shfl_xor(na, off);
      nb = nb + __shfl_xor(nb, off);
    }
  }
  if (lane == 0) {
    double r = d;
    if (METRIC == kL2) r = sqrt(d);
    else if (METRIC == kCos) r = d / (sqrt(na) * sqrt(nb));
    out[p] = r;
  }
}

hipError_t launch_pair_f64(const double* X, size_t ldx, const uint32_t* sa, const uint32_t* sb,
                           uint32_t np, uint32_t D, int metric, double* out, hipStream_t s) {
  if (np == 0) return hipSuccess;
  dim3 grid((np + kWavesPerBlock - 1) / kWavesPerBlock), block(256);
  switch (metric) {
    case kL2: hipLaunchKernelGGL(k_pair_f64<kL2>, grid, block, 0, s, X, ldx, sa, sb, np, D, out); break;
    case kCos: hipLaunchKernelGGL(k_pair_f64<kCos>, grid, block, 0, s, X, ldx, sa, sb, np, D, out); break;
    case kDot: hipLaunchKernelGGL(k_pair_f64<kDot>, grid, block, 0, s, X, ldx, sa, sb, np, D, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// The reference's own float arithmetic (vector_database.c:301-352), one
// thread per pair, so the legacy float-returning metrics are bit-identical
// to the reference: FP64 products added into float accumulators
// (acc = (float)((double)acc + a*b)); Euclidean squares the float-rounded
// difference in float; sqrt and the cosine division in double.
// The reference's float metric of one pair (vector_database.c:301-352):
// float accumulators, each step widened to double and rounded back, as its
// C compiles.
__device__ float ref32_metric(const double* __restrict__ A, const double* __restrict__ B,
                              uint32_t n, int metric) {
  if (metric == kCos) {
    float dp = 0.0f, na = 0.0f, nb = 0.0f;
    for (uint32_t i = 0; i < n; i++) {
      dp = (float)((double)dp + A[i] * B[i]);
      na = (float)((double)na + A[i] * A[i]);
      nb = (float)((double)nb + B[i] * B[i]);
    }
    return (float)((double)dp / (sqrt((double)na) * sqrt((double)nb)));
  }
  if (metric == kL2) {
    float sum = 0.0f;
    for (uint32_t i = 0; i < n; i++) {
      float diff = (float)(A[i] - B[i]);
      float sq = diff * diff;
      sum = sum + sq;
    }
    return (float)sqrt((double)sum);
  }
  float acc = 0.0f;
  for (uint32_t i = 0; i < n; i++) acc = (float)((double)acc + A[i] * B[i]);
  return acc;
}

__global__ void k_pair_ref32(const double* __restrict__ M, size_t ld,
                             const uint32_t* __restrict__ sa, const uint32_t* __restrict__ sb,
                             uint32_t np, uint32_t n, int metric, float* __restrict__ out) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  out[p] = ref32_metric(M + (size_t)sa[p] * ld, M + (size_t)sb[p] * ld, n, metric);
}

// One pair from the kernel arguments (the drop-in per-call functions): the
// result, then seq once it is visible, into host memory the caller polls.
__global__ void k_pair_ref32_args(const PairRef32Args a) {
  if (threadIdx.x != 0) return;
  const float r = ref32_metric(a.ab, a.ab + a.n, a.n, a.metric);
  a.out[0] = (uint64_t)__float_as_uint(r);
  __threadfence_system();
  __hip_atomic_store(a.out + 1, a.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

hipError_t launch_pair_ref32_args(const PairRef32Args& a, hipStream_t s) {
  if (a.n > kRef32ArgsMax) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_pair_ref32_args, dim3(1), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_pair_ref32(const double* M, size_t ld, const uint32_t* sa, const uint32_t* sb,
                             uint32_t np, uint32_t n, int metric, float* out, hipStream_t s) {
  if (np == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pair_ref32, dim3((np + 255) / 256), dim3(256), 0, s, M, ld, sa, sb, np, n,
                     metric, out);
  return hipGetLastError();
}

}  // namespace svdb
