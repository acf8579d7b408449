// svdb_lowp.hip — BF16 MFMA first pass, shortlist, FP64 rerank with
// certification (gfx950).  See svdb_lowp.h for the shadow layout and the
// error bound.  Built with -ffp-contract=off like the FP64 kernels.
#include "svdb_lowp.h"

#include <math.h>

namespace svdb {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------- helpers

__device__ __forceinline__ uint16_t bf16_of(double x) {
  float f = __double2float_rn(x);
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u) return (u & 0x007FFFFFu) ? (uint16_t)0x7FC0 : (uint16_t)(u >> 16);
  u += 0x7FFFu + ((u >> 16) & 1u);
  uint16_t b = (uint16_t)(u >> 16);
  if ((b & 0x7F80u) == 0) b &= 0x8000u;  // no bf16 subnormals: flush to zero
  return b;
}

__device__ __forceinline__ double bf16_val(uint16_t b) {
  return (double)__uint_as_float((uint32_t)b << 16);
}

__device__ __forceinline__ uint64_t rank_bits64(double v) {
  v = v + 0.0;
  uint64_t u = (uint64_t)__double_as_longlong(v);
  if (v != v) u = 0x7FF8000000000000ull;
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__device__ __forceinline__ double rank_val64(uint64_t u) {
  uint64_t b = (u >> 63) ? (u & 0x7FFFFFFFFFFFFFFFull) : ~u;
  return __longlong_as_double((long long)b);
}

__device__ __forceinline__ uint32_t rank_bits32(float v) {
  uint32_t u = __float_as_uint(v + 0.0f);
  if (v != v) u = 0x7FC00000u;
  return (u >> 31) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ bool lt2(uint64_t ka, uint32_t sa, uint64_t kb, uint32_t sb) {
  return (ka < kb) | ((ka == kb) & (sa < sb));
}

__device__ __forceinline__ uint64_t uni64(uint64_t v) {
  uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t uni32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Wave-distributed sorted list (entry e in lane e%64, register e/64).
template <int KPL>
struct WL {
  uint64_t k[KPL];
  uint32_t s[KPL];
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int p = 0; p < KPL; p++) {
      k[p] = ~0ull;
      s[p] = ~0u;
    }
  }
  __device__ __forceinline__ void at(int e, uint64_t& tk, uint32_t& ts) const {
    int p = e >> 6, ln = e & 63;
    uint64_t a = k[0];
    uint32_t b = s[0];
#pragma unroll
    for (int q = 1; q < KPL; q++)
      if (p == q) {
        a = k[q];
        b = s[q];
      }
    tk = uni64(__shfl(a, ln));
    ts = uni32(__shfl(b, ln));
  }
  __device__ __forceinline__ void insert(uint64_t K, uint32_t S, int kk, int lane) {
    int pos = 0;
#pragma unroll
    for (int p = 0; p < KPL; p++) pos += __popcll(__ballot(lt2(k[p], s[p], K, S)));
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

// ------------------------------------------------------------- shadow rows

__global__ void k_shadow_init(Shadow sh, uint32_t first, uint32_t n) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  uint32_t slot = first + r;
  sh.nx[slot] = INFINITY;
  sh.xn[slot] = -INFINITY;
  sh.inv[slot] = 0.0f;
  sh.pen[slot] = INFINITY;
}

hipError_t launch_shadow_init(const Shadow& sh, uint32_t first, uint32_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_shadow_init, dim3((n + 255) / 256), dim3(256), 0, s, sh, first, n);
  return hipGetLastError();
}

__device__ void kill_shadow_row(const Shadow& sh, uint32_t slot) {
  uint16_t* xb = sh.Xb + (size_t)slot * sh.dp;
  for (size_t i = 0; i < sh.dp; i++) xb[i] = 0;
  sh.nx[slot] = INFINITY;
  sh.xn[slot] = -INFINITY;
  sh.inv[slot] = 0.0f;
  sh.pen[slot] = INFINITY;
}

// One thread per row (after k_row_norms has written nx2 / xn).
__global__ void k_quantize_rows(Shadow sh, const double* __restrict__ X, size_t ldx,
                                const double* __restrict__ nx2, const double* __restrict__ xnorm,
                                const uint32_t* __restrict__ slots, uint32_t first, uint32_t n,
                                uint32_t d) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint32_t slot = slots ? slots[r] : first + r;
  const double* x = X + (size_t)slot * ldx;
  uint16_t* xb = sh.Xb + (size_t)slot * sh.dp;
  double esq = 0.0, maxabs = 0.0;
  bool finite = true;
  for (uint32_t i = 0; i < d; i++) {
    double v = x[i];
    finite = finite && isfinite(v);
    maxabs = fmax(maxabs, fabs(v));
    uint16_t b = bf16_of(v);
    xb[i] = b;
    double e = bf16_val(b) - v;
    esq = fma(e, e, esq);
  }
  for (size_t i = d; i < sh.dp; i++) xb[i] = 0;
  const double norm = xnorm[slot];
  const double lim = 1152921504606846976.0;  // 2^60
  const bool degenerate = !finite || !(norm >= 1.0 / lim) || !(norm <= lim) || !(maxabs <= lim);
  if (degenerate) {
    kill_shadow_row(sh, slot);
    uint32_t cnt = atomicAdd(sh.degen_count, 0u);
    bool present = false;
    for (uint32_t j = 0; j < cnt && j < (uint32_t)kDegenCap; j++) present |= sh.degen[j] == slot;
    if (!present) {
      uint32_t idx = atomicAdd(sh.degen_count, 1u);
      if (idx < (uint32_t)kDegenCap) sh.degen[idx] = slot;
    }
    return;
  }
  sh.nx[slot] = __double2float_rd(nx2[slot]);
  sh.xn[slot] = __double2float_ru(norm);
  sh.inv[slot] = __double2float_rn(1.0 / norm);
  sh.pen[slot] = 0.0f;
  const double up = 1.0 + 0x1p-40;
  const double ratio = (sqrt(esq) * up) / norm * up;
  atomicMax(sh.rho_bits, __float_as_uint(__double2float_ru(ratio)));
}

hipError_t launch_quantize_rows(const Shadow& sh, const double* X, size_t ldx, const double* nx2,
                                const double* xn, const uint32_t* slots, uint32_t first,
                                uint32_t n, uint32_t d, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_quantize_rows, dim3((n + 127) / 128), dim3(128), 0, s, sh, X, ldx, nx2, xn,
                     slots, first, n, d);
  return hipGetLastError();
}

__global__ void k_kill_rows(Shadow sh, const uint32_t* __restrict__ slots, uint32_t n) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) kill_shadow_row(sh, slots[r]);
}

hipError_t launch_kill_rows(const Shadow& sh, const uint32_t* slots, uint32_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_kill_rows, dim3((n + 63) / 64), dim3(64), 0, s, sh, slots, n);
  return hipGetLastError();
}

// ------------------------------------------------------------- query prep
//
// One thread per query: FP64 ‖q‖ (the canonical chain), bf16 q̂ scaled by
// -2 (Euclidean) or -1 (dot, cosine) — exact power-of-two scalings — and
// the per-query bound constants.

__global__ void k_query_prep(const double* __restrict__ Q, size_t ldq, uint32_t nq,
                             uint32_t nq_pad, uint32_t d, size_t dp, int metric,
                             const uint32_t* __restrict__ rho_bits, uint16_t* __restrict__ Qb,
                             QParams* __restrict__ qp) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq_pad) return;
  uint16_t* qb = Qb + (size_t)i * dp;
  QParams p{};
  p.metric = metric;
  if (i >= nq) {
    for (size_t c = 0; c < dp; c++) qb[c] = 0;
    p.flags = 1;
    qp[i] = p;
    return;
  }
  const double* q = Q + (size_t)i * ldq;
  const double sc = metric == kL2 ? -2.0 : -1.0;
  double nq2 = 0.0, esq = 0.0, hsq = 0.0, maxabs = 0.0;
  bool finite = true;
  for (uint32_t c = 0; c < d; c++) {
    double v = q[c];
    nq2 = fma(v, v, nq2);
    finite = finite && isfinite(v);
    maxabs = fmax(maxabs, fabs(v));
    uint16_t b = bf16_of(v);
    double h = bf16_val(b);
    double e = h - v;
    esq = fma(e, e, esq);
    hsq = fma(h, h, hsq);
    qb[c] = bf16_of(h * sc);
  }
  for (size_t c = d; c < dp; c++) qb[c] = 0;
  const double qn = sqrt(nq2);
  const double lim = 1152921504606846976.0;  // 2^60
  const double up = 1.0 + 0x1p-40;
  const double rho = (double)__uint_as_float(*rho_bits);
  const double gam = (double)(d + 32) * 0x1p-24;
  const double qh = sqrt(hsq) * up, eq = sqrt(esq) * up;
  const double qh1 = qh * (1.0 + rho) * (1.0 + gam) * up;  // >= |s| / ‖x‖
  const double delta = (double)d * 0x1p-124;
  const double c = (gam * qh * (1.0 + rho) + eq + rho * qh * (1.0 + gam) +
                    (double)(d + 8) * 0x1p-52 * qn) * up;
  p.qn = qn;
  p.nq2 = nq2;
  p.flags = (!finite || !(qn >= 1.0 / lim) || !(qn <= lim) || !(maxabs <= lim)) ? 1 : 0;
  if (metric == kL2) {
    const double a = 1.0 - c / qn - 0x1p-24 * (1.0 + qh1 / qn) - 0x1p-30;
    p.P = __double2float_rd(a);
    p.scale = 1.0;
    p.off = nq2 - qn * (c + 0x1p-24 * qh1) - 2.0 * delta;
    if (!(a > 0.0)) p.flags = 1;
  } else if (metric == kDot) {
    const double cp = (c + 0x1p-23 * (qh1 + c) * 1.01) * up;
    p.P = __double2float_rd(-cp);
    p.scale = 1.0;
    p.off = -delta;
  } else {
    p.P = 0.0f;
    p.scale = 1.0 / qn;
    p.off = -(c + 0x1p-23 * qh1 * 1.01 + delta * lim) / qn - (double)(d + 16) * 0x1p-52;
  }
  qp[i] = p;
}

hipError_t launch_query_prep(const double* Q, size_t ldq, uint32_t nq, uint32_t nq_pad,
                             uint32_t d, size_t dp, int metric, const uint32_t* rho_bits,
                             uint16_t* Qb, QParams* qp, hipStream_t s) {
  if (nq_pad == 0) return hipSuccess;
  hipLaunchKernelGGL(k_query_prep, dim3((nq_pad + 127) / 128), dim3(128), 0, s, Q, ldq, nq, nq_pad,
                     d, dp, metric, rho_bits, Qb, qp);
  return hipGetLastError();
}

// ------------------------------------------------------------- shortlist
//
// One wavefront per query: merge the query's NL candidate lists (8 entries
// each) into the L best by (ordering key, slot), and the bound
// beta = min(8th entry of every list, L-th merged key): no row outside the
// shortlist has a smaller ordering key.

template <int KPL>
__device__ __forceinline__ void shortlist_body(const ShortlistArgs& a, uint32_t blk) {
  const int lane = threadIdx.x & 63;
  const uint32_t i = blk * 4 + (threadIdx.x >> 6);
  if (i >= a.nq) return;
  const uint32_t q = a.qidx ? a.qidx[i] : i;  // candidate lists of query q -> output row i
  const uint32_t n = a.NL * kLowpLP;
  const size_t base = (size_t)q * n;
  WL<KPL> L;
  L.init();
  const int kk = (int)a.L;
  uint64_t tk = ~0ull;
  uint32_t ts = ~0u;
  float lane_min = INFINITY;
  // Entries are read kSlU batches of 64 at a time, so the loads of a batch
  // group are in flight together (most entries fail the ballot and cost
  // only their load).
  constexpr int kSlU = 8;
  for (uint32_t c0 = 0; c0 < n; c0 += 64 * kSlU) {
    float key[kSlU];
    uint32_t slot[kSlU];
#pragma unroll
    for (int u = 0; u < kSlU; u++) {
      const uint32_t c = c0 + u * 64 + lane;
      key[u] = INFINITY;
      slot[u] = ~0u;
      if (c < n) {
        key[u] = a.cand_key[base + c];
        slot[u] = a.cand_slot[base + c];
      }
    }
#pragma unroll
    for (int u = 0; u < kSlU; u++) {
      const uint32_t c = c0 + u * 64 + lane;
      if (c < n && (c % kLowpLP) == kLowpLP - 1) lane_min = fminf(lane_min, key[u]);
      const uint64_t K = slot[u] == ~0u ? ~0ull : (uint64_t)rank_bits32(key[u]);
      uint64_t m = __ballot(slot[u] != ~0u && lt2(K, slot[u], tk, ts));
      while (m) {
        int src = __ffsll((long long)m) - 1;
        m &= m - 1;
        L.insert(uni64(__shfl(K, src)), uni32(__shfl(slot[u], src)), kk, lane);
        // lanes that no longer beat the (moved) kk-th entry drop out
        L.at(kk - 1, tk, ts);
        m &= __ballot(lt2(K, slot[u], tk, ts));
      }
      L.at(kk - 1, tk, ts);
    }
  }
  // min over lanes of the lists' 8th entries
  for (int off = 32; off >= 1; off >>= 1) lane_min = fminf(lane_min, __shfl_xor(lane_min, off));
  float oL = INFINITY;
  if (ts != ~0u) {
    const uint32_t u = (uint32_t)tk;
    oL = __uint_as_float((u >> 31) ? (u & 0x7FFFFFFFu) : ~u);
  }
#pragma unroll
  for (int p = 0; p < KPL; p++) {
    const int e = p * 64 + lane;
    if (e < kk) {
      const uint32_t u = (uint32_t)L.k[p];
      a.short_slot[(size_t)i * a.L + e] = L.s[p];
      a.short_key[(size_t)i * a.L + e] =
          L.s[p] == ~0u ? INFINITY : __uint_as_float((u >> 31) ? (u & 0x7FFFFFFFu) : ~u);
    }
  }
  if (lane == 0) a.beta[i] = fminf(lane_min, oL);
}

// Wide shortlists (L > 256): one 256-thread block per query.  The query's
// n <= 256*kSelPT candidate keys sit in registers as order-preserving
// uint32 (empty entries: all ones); four 8-bit histogram passes in LDS find
// the L-th smallest key K*; every entry below K* and enough entries equal to
// it are written out, unordered (the rerank does not need an order).  A row
// left out has a key >= K* or >= its list's 8th entry, so
// beta = min(8th entry of every list, K*) as in the sorted form.
constexpr int kSelPT = 64;

struct SelShared {
  uint32_t hist[256];
  uint32_t prefix, mask, rank, nvalid, lt, eq;
  float min[4];
};

__device__ __forceinline__ void shortlist_sel_body(const ShortlistArgs& a, uint32_t blk,
                                                   SelShared& sm) {
  uint32_t* hist = sm.hist;
  uint32_t &s_prefix = sm.prefix, &s_mask = sm.mask, &s_rank = sm.rank, &s_nvalid = sm.nvalid,
           &s_lt = sm.lt, &s_eq = sm.eq;
  float* s_min = sm.min;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t i = blk;
  const uint32_t q = a.qidx ? a.qidx[i] : i;
  const uint32_t n = a.NL * kLowpLP;
  const size_t base = (size_t)q * n;
  uint32_t u[kSelPT];
  float lmin = INFINITY;
  uint32_t nv = 0;
#pragma unroll
  for (int t = 0; t < kSelPT; t++) {
    const uint32_t c = t * 256 + tid;
    u[t] = 0xFFFFFFFFu;
    if (c < n) {
      const float key = a.cand_key[base + c];
      if ((c % kLowpLP) == kLowpLP - 1) lmin = fminf(lmin, key);
      if (a.cand_slot[base + c] != ~0u) {
        u[t] = rank_bits32(key);
        nv++;
      }
    }
  }
  for (int off = 32; off >= 1; off >>= 1) {
    lmin = fminf(lmin, __shfl_xor(lmin, off));
    nv += __shfl_xor(nv, off);
  }
  if (tid == 0) {
    s_prefix = 0;
    s_mask = 0;
    s_rank = a.L;
    s_nvalid = 0;
    s_lt = 0;
    s_eq = 0;
  }
  __syncthreads();
  if (lane == 0) {
    s_min[wave] = lmin;
    atomicAdd(&s_nvalid, nv);
  }
  __syncthreads();
  const uint32_t nvalid = s_nvalid;
  lmin = fminf(fminf(s_min[0], s_min[1]), fminf(s_min[2], s_min[3]));
  uint32_t kstar = 0xFFFFFFFFu;  // fewer than L valid entries: take them all
  if (nvalid >= a.L) {
    for (int pass = 0; pass < 4; pass++) {
      const int shift = 24 - 8 * pass;
      hist[tid] = 0;
      __syncthreads();
      const uint32_t prefix = s_prefix, mask = s_mask;
#pragma unroll
      for (int t = 0; t < kSelPT; t++)
        if (u[t] != 0xFFFFFFFFu && (u[t] & mask) == prefix)
          atomicAdd(&hist[(u[t] >> shift) & 255u], 1u);
      __syncthreads();
      if (wave == 0) {
        // lane l holds bins 4l..4l+3; an inclusive scan over lanes finds
        // the bin where the running count reaches the target rank
        const uint32_t h0 = hist[4 * lane], h1 = hist[4 * lane + 1], h2 = hist[4 * lane + 2],
                       h3 = hist[4 * lane + 3];
        const uint32_t sum = h0 + h1 + h2 + h3;
        uint32_t inc = sum;
        for (int off = 1; off < 64; off <<= 1) {
          const uint32_t v = __shfl_up(inc, off);
          if (lane >= off) inc += v;
        }
        const uint32_t exc = inc - sum, rank = s_rank;
        if (exc < rank && rank <= inc) {
          uint32_t cum = exc, bin = 4 * lane;
          const uint32_t hs[4] = {h0, h1, h2, h3};
          for (int b = 0; b < 4; b++) {
            if (cum + hs[b] >= rank) {
              bin = 4 * lane + b;
              break;
            }
            cum += hs[b];
          }
          s_rank = rank - cum;
          s_prefix = prefix | (bin << shift);
          s_mask = mask | (255u << shift);
        }
      }
      __syncthreads();
    }
    kstar = s_prefix;
  }
  // entries below K*: all of them; equal to K*: the first s_rank that come
  const uint32_t need_eq = nvalid >= a.L ? s_rank : 0u;
  const uint32_t n_lt = nvalid >= a.L ? a.L - need_eq : 0u;
#pragma unroll
  for (int t = 0; t < kSelPT; t++) {
    if (u[t] == 0xFFFFFFFFu) continue;
    uint32_t pos = ~0u;
    if (nvalid < a.L || u[t] < kstar) {
      pos = atomicAdd(&s_lt, 1u);
    } else if (u[t] == kstar) {
      const uint32_t e = atomicAdd(&s_eq, 1u);
      if (e < need_eq) pos = n_lt + e;
    }
    if (pos != ~0u) {
      const uint32_t c = t * 256 + tid;
      a.short_slot[(size_t)i * a.L + pos] = a.cand_slot[base + c];
      a.short_key[(size_t)i * a.L + pos] = a.cand_key[base + c];
    }
  }
  __syncthreads();
  // unused tail (fewer than L valid entries)
  for (uint32_t e = (nvalid < a.L ? nvalid : a.L) + tid; e < a.L; e += 256) {
    a.short_slot[(size_t)i * a.L + e] = ~0u;
    a.short_key[(size_t)i * a.L + e] = INFINITY;
  }
  if (tid == 0) {
    const float oL =
        nvalid >= a.L ? __uint_as_float((kstar >> 31) ? (kstar & 0x7FFFFFFFu) : ~kstar) : INFINITY;
    a.beta[i] = fminf(lmin, oL);
  }
}

// Several ranges' shortlists (or reranks) in one launch: range r takes
// blocks first[r] .. first[r+1]-1 (a launch per range costs its latency
// once per range, and small ranges leave most of the GPU idle).
template <typename A>
struct RangeBatch {
  A r[kMaxLaunchRanges];
  int metric[kMaxLaunchRanges];
  uint32_t first[kMaxLaunchRanges + 1];
  int n;
};

template <typename A>
__device__ __forceinline__ int range_of(const RangeBatch<A>& b, uint32_t blk) {
  int i = 0;
  while (i + 1 < b.n && blk >= b.first[i + 1]) i++;
  return i;
}

template <int KPL>
__global__ __launch_bounds__(256) void k_shortlist(RangeBatch<ShortlistArgs> b) {
  const int i = range_of(b, blockIdx.x);
  shortlist_body<KPL>(b.r[i], blockIdx.x - b.first[i]);
}

__global__ __launch_bounds__(256) void k_shortlist_sel(RangeBatch<ShortlistArgs> b) {
  __shared__ SelShared sm;
  const int i = range_of(b, blockIdx.x);
  shortlist_sel_body(b.r[i], blockIdx.x - b.first[i], sm);
}

// Up to 64 wide: one wavefront per query keeps a sorted list (few
// insertions past the first 64).  Wider: the block-per-query select, whose
// cost does not grow with L (a 256-wide sorted list in one wavefront takes
// hundreds of serial insertions per query); the sorted 256-wide form
// remains for candidate sets too large for its registers.
static int shortlist_variant(const ShortlistArgs& a) {
  const bool sel_fits = (size_t)a.NL * kLowpLP <= 256u * kSelPT;
  if (a.L == 0 || a.L > (uint32_t)kMaxShort) return -1;
  if (a.L <= 64) return 1;
  if (sel_fits) return 0;
  if (a.L <= 256) return 4;
  return -1;
}

hipError_t launch_shortlist_ranges(const ShortlistArgs* a, int n, hipStream_t s) {
  for (int r0 = 0; r0 < n;) {
    // one launch per run of ranges that take the same kernel
    RangeBatch<ShortlistArgs> b{};
    const int v = shortlist_variant(a[r0]);
    if (v < 0) return hipErrorInvalidValue;
    uint32_t blocks = 0;
    b.n = 0;
    int r = r0;
    for (; r < n && b.n < kMaxLaunchRanges && shortlist_variant(a[r]) == v; r++) {
      b.r[b.n] = a[r];
      b.first[b.n] = blocks;
      blocks += v == 0 ? a[r].nq : (a[r].nq + 3) / 4;
      b.n++;
    }
    b.first[b.n] = blocks;
    r0 = r;
    if (blocks == 0) continue;
    if (v == 1) hipLaunchKernelGGL(k_shortlist<1>, dim3(blocks), dim3(256), 0, s, b);
    else if (v == 4) hipLaunchKernelGGL(k_shortlist<4>, dim3(blocks), dim3(256), 0, s, b);
    else hipLaunchKernelGGL(k_shortlist_sel, dim3(blocks), dim3(256), 0, s, b);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_shortlist(const ShortlistArgs& a, hipStream_t s) {
  return launch_shortlist_ranges(&a, 1, s);
}

// ------------------------------------------------ L-th smallest of n keys
//
// One wavefront per query; the keys (n <= 64 * kKthPerLane) sit in
// registers as order-preserving uint32; a 32-step binary search on that
// order counts, with ballots, how many lie at or below the midpoint.  The
// result is one of the keys: the L-th smallest.

constexpr int kKthPerLane = 16;

__global__ __launch_bounds__(256) void k_kth_key(const float* __restrict__ key, uint32_t nq,
                                                 uint32_t n, uint32_t L, float* __restrict__ thr) {
  const int lane = threadIdx.x & 63;
  const uint32_t q = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (q >= nq) return;
  uint32_t u[kKthPerLane];
#pragma unroll
  for (int i = 0; i < kKthPerLane; i++) {
    const uint32_t c = i * 64 + lane;
    u[i] = c < n ? rank_bits32(key[(size_t)q * n + c]) : 0xFFFFFFFFu;
  }
  const uint32_t inf_bits = rank_bits32(INFINITY);
  uint32_t lo = 0, hi = 0xFFFFFFFFu;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    uint32_t cnt = 0;
#pragma unroll
    for (int i = 0; i < kKthPerLane; i++) cnt += __popcll(__ballot(u[i] <= mid));
    if (cnt >= L) hi = mid;
    else lo = mid + 1;
  }
  if (lane == 0)
    thr[q] = (lo >= inf_bits || n < L) ? INFINITY
                                       : __uint_as_float((lo >> 31) ? (lo & 0x7FFFFFFFu) : ~lo);
}

hipError_t launch_kth_key(const float* key, uint32_t nq, uint32_t n, uint32_t L, float* thr,
                          hipStream_t s) {
  if (nq == 0) return hipSuccess;
  if (n > 64u * kKthPerLane || L == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_kth_key, dim3((nq + 3) / 4), dim3(256), 0, s, key, nq, n, L, thr);
  return hipGetLastError();
}

// ------------------------------------------------------------- FP64 rerank
//
// One wavefront per query.  Lanes take the candidates 64 at a time (the
// shortlist, then the degenerate rows) and compute their canonical FP64 rank
// key with the sequential fma chain; the k best by (key, slot) are
// selected; the query is certified when its k-th key is below the bound
// beta*scale + off (minus the FP64 rounding slack of that expression and,
// for Euclidean, of the keys themselves).  Certified queries get their ids
// and scores written; the others are listed for the FP64 scan.

// Candidate rows reach the lanes through LDS: per round of 64 candidates
// the wave copies kRC columns of all 64 rows at a time (lane l fetches
// column l%16 of candidates 4u + l/16, u = 0..15: each load instruction
// reads 4 contiguous 128-byte row pieces), then each lane continues its own
// candidate's sequential fma chain over those columns.  The FP64 keys are
// the canonical chains, in index order, as before.
constexpr int kRC = 16;

template <int METRIC>
__device__ __forceinline__ void rerank_body(const RerankArgs& a, uint32_t blk, double* stg,
                                            uint32_t* queue) {
  const int lane = threadIdx.x & 63;
  const uint32_t i = blk * 4 + (threadIdx.x >> 6);
  if (i >= a.nq) return;
  // shortlist row i belongs to query q (its parameters, FP64 row, outputs)
  const uint32_t q = a.qidx ? a.qidx[i] : i;
  const QParams qp = a.qp[q];
  const double* __restrict__ qv = a.Q + (size_t)q * a.ldq;
  const uint32_t ndeg = min(*a.degen_count, (uint32_t)kDegenCap);
  const int kk = (int)a.k;

  // FP64 lower bound on the ordering value of a row whose first-pass key is
  // at least `bo` (the rounding slack of the key's conversion included)
  auto lower = [&](float bo) __attribute__((always_inline)) -> double {
    if (bo == INFINITY) return INFINITY;
    const double bs = (double)bo * qp.scale;
    double bR = bs + qp.off;
    bR -= 0x1p-40 * (fabs(bs) + fabs(qp.off));
    if (METRIC == kL2 && bR > 0.0) bR *= 1.0 - (double)(a.d + 4) * 0x1p-52;
    return bR;
  };
  // a flagged query's keys carry no bound: no pruning
  const bool prune = !qp.flags;

  WL<2> L;
  L.init();
  uint64_t tk = ~0ull;
  uint32_t ts = ~0u;
  // the earlier stage's list, and the key below which its shortlists
  // already covered every candidate
  const bool carry = a.carry && a.lk;
  float kprev = -INFINITY;
  if (carry) {
#pragma unroll
    for (int p = 0; p < 2; p++) {
      const int e = p * 64 + lane;
      if (e < kk) {
        L.k[p] = a.lk[(size_t)q * kk + e];
        L.s[p] = a.ls[(size_t)q * kk + e];
      }
    }
    L.at(kk - 1, tk, ts);
    kprev = a.kmax[q];
  }
  float kmax_lane = -INFINITY;
  const uint32_t ncand = a.L + ndeg;

  // Scores up to 64 candidates, lane l's `slot` when `valid`, and enters
  // them into the list.
  auto score = [&](uint32_t slot, bool valid) __attribute__((always_inline)) {
    // rows this lane fetches: candidates 4u + lane/16
    const double* src[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int cand = u * 4 + (lane >> 4);
      const uint32_t sl = __shfl(slot, cand);
      const int ok = __shfl(valid ? 1 : 0, cand);
      src[u] = ok ? a.st.X + (size_t)sl * a.st.ldx : nullptr;
    }
    double acc = 0.0;
    const double* my = stg + lane * (kRC + 1);
    // The next chunk's loads go out once this chunk's values sit in LDS
    // (into the same registers), so they are in flight during its FMAs.
    double vals[16];
    auto fetch = [&](uint32_t c0) __attribute__((always_inline)) {
      const uint32_t e = c0 + (lane & 15);
#pragma unroll
      for (int u = 0; u < 16; u++) vals[u] = (src[u] && e < a.d) ? src[u][e] : 0.0;
    };
    fetch(0);
    for (uint32_t c0 = 0; c0 < a.d; c0 += kRC) {
#pragma unroll
      for (int u = 0; u < 16; u++) stg[(u * 4 + (lane >> 4)) * (kRC + 1) + (lane & 15)] = vals[u];
      __builtin_amdgcn_wave_barrier();
      if (c0 + kRC < a.d) fetch(c0 + kRC);
      __builtin_amdgcn_sched_barrier(0);
      const uint32_t ne = min((uint32_t)kRC, a.d - c0);
#pragma unroll
      for (int t = 0; t < kRC; t++) {
        if ((uint32_t)t < ne) {
          const double x = my[t];
          if (METRIC == kL2) {
            const double tt = x - qv[c0 + t];
            acc = fma(tt, tt, acc);
          } else {
            acc = fma(qv[c0 + t], x, acc);
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
    uint64_t K = ~0ull;
    if (valid) {
      double v = acc;
      if (METRIC == kDot) v = -acc;
      else if (METRIC == kCos) v = -(acc / (qp.qn * a.st.xn[slot]));
      K = rank_bits64(v);
    }
    uint64_t m = __ballot(valid && lt2(K, slot, tk, ts));
    while (m) {
      int src = __ffsll((long long)m) - 1;
      m &= m - 1;
      const uint32_t S = uni32(__shfl(slot, src));
      // a row the carried list already holds (keyed at kprev, or a
      // degenerate row) must not enter twice
      if (carry && __ballot(L.s[0] == S || L.s[1] == S) != 0) continue;
      L.insert(uni64(__shfl(K, src)), S, kk, lane);
      // lanes that no longer beat the (moved) k-th entry drop out
      L.at(kk - 1, tk, ts);
      m &= __ballot(lt2(K, slot, tk, ts));
    }
    L.at(kk - 1, tk, ts);
  };

  // Candidates that still need their FP64 row are packed, in order, into
  // this wavefront's queue and scored 64 at a time, so a group fetches 64
  // rows however many of the shortlist's candidates were skipped.
  uint32_t nqd = 0;  // queued (wave-uniform; at most 127)
  uint32_t j0 = 0;
  for (;;) {
    while (nqd < 64 && j0 < ncand) {
      const uint32_t j = j0 + lane;
      uint32_t slot = ~0u;
      if (j < a.L) slot = a.short_slot[(size_t)i * a.L + j];
      else if (j < ncand) slot = a.degen[j - a.L];
      bool valid = slot != ~0u && slot < a.st.nrows &&
                   ((a.st.alive[slot >> 6] >> (slot & 63)) & 1ull);
      const float key = j < a.L ? a.short_key[(size_t)i * a.L + j] : -INFINITY;
      kmax_lane = fmaxf(kmax_lane, key);
      // scored in an earlier stage (its best k are in the carried list)
      if (carry && j < a.L && key < kprev) valid = false;
      // Prune threshold: the list's k-th value once full (+inf: fetch every
      // candidate).  A row whose lower bound exceeds it cannot be in the
      // top k.  (A NaN bound compares false and keeps its row.)
      const double thr = (prune && ts != ~0u) ? rank_val64(tk) : INFINITY;
      if (valid && j < a.L && thr < INFINITY && lower(key) > thr) valid = false;
      // a degenerate row that is also shortlisted must not enter twice
      if (valid && j >= a.L) {
        for (uint32_t t = 0; t < a.L; t++)
          if (a.short_slot[(size_t)i * a.L + t] == slot) valid = false;
      }
      const uint64_t vm = __ballot(valid);
      if (valid) queue[nqd + __popcll(vm & ((1ull << lane) - 1ull))] = slot;
      nqd += __popcll(vm);
      j0 += 64;
    }
    __builtin_amdgcn_wave_barrier();
    if (nqd == 0) break;
    const uint32_t n = min(nqd, 64u);
    const uint32_t slot = (uint32_t)lane < n ? queue[lane] : ~0u;
    const uint32_t rest = (uint32_t)lane + 64 < nqd ? queue[lane + 64] : 0u;
    __builtin_amdgcn_wave_barrier();
    if ((uint32_t)lane + 64 < nqd) queue[lane] = rest;
    nqd -= n;
    __builtin_amdgcn_wave_barrier();
    score(slot, (uint32_t)lane < n);
  }

  // certification
  // more degenerate rows than the list holds: nothing can be certified
  bool cert = false;
  const double tau = ts == ~0u ? INFINITY : rank_val64(tk);
  if (!qp.flags && *a.degen_count <= (uint32_t)kDegenCap) cert = tau < lower(a.beta[i]);
  if (!cert) {
    if (a.lk) {
#pragma unroll
      for (int p = 0; p < 2; p++) {
        const int e = p * 64 + lane;
        if (e < kk) {
          a.lk[(size_t)q * kk + e] = L.k[p];
          a.ls[(size_t)q * kk + e] = L.s[p];
        }
      }
      for (int off = 32; off >= 1; off >>= 1) kmax_lane = fmaxf(kmax_lane, __shfl_xor(kmax_lane, off));
      if (lane == 0) a.kmax[q] = fmaxf(kmax_lane, kprev);
    }
    if (lane == 0) {
      uint32_t idx = atomicAdd(a.fail_count, 1u);
      a.fail_list[idx] = q;
    }
    return;
  }
  // First-pass keys of this lane's results, for the error report: the
  // shortlist passes through the wavefront's LDS area in chunks of kOC
  // (slot, key) pairs and every lane matches its two result slots there.
  constexpr uint32_t kOC = 1024;  // 8 KB of the 64 * (kRC + 1) doubles
  static_assert(kOC * 8 <= 64 * (kRC + 1) * sizeof(double), "chunk exceeds the stage");
  float O2[2] = {INFINITY, INFINITY};
  {
    uint32_t* oslot = reinterpret_cast<uint32_t*>(stg);
    float* okey = reinterpret_cast<float*>(oslot + kOC);
    for (uint32_t t0 = 0; t0 < a.L; t0 += kOC) {
      const uint32_t nt = min(kOC, a.L - t0);
      for (uint32_t t = lane; t < nt; t += 64) {
        oslot[t] = a.short_slot[(size_t)i * a.L + t0 + t];
        okey[t] = a.short_key[(size_t)i * a.L + t0 + t];
      }
      __builtin_amdgcn_wave_barrier();
      for (uint32_t t = 0; t < nt; t++) {
        const uint32_t sl = oslot[t];
        if (sl == L.s[0]) O2[0] = okey[t];
        if (sl == L.s[1]) O2[1] = okey[t];
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  double err = 0.0;
#pragma unroll
  for (int p = 0; p < 2; p++) {
    const int e = p * 64 + lane;
    if (e < kk) {
      uint64_t id = ~0ull;
      double sc = __longlong_as_double(0x7FF8000000000000ll);
      const uint32_t slot = L.s[p];
      if (slot != ~0u) {
        const double v = rank_val64(L.k[p]);
        if (v == v) sc = METRIC == kL2 ? sqrt(v) : (v == 0.0 ? 0.0 : -v);
        id = a.st.has_deletes
                 ? (uint64_t)a.st.rank_base[slot >> 6] +
                       (uint64_t)__popcll(a.st.alive[slot >> 6] & ((1ull << (slot & 63)) - 1ull))
                 : slot;
        // first-pass estimate of this score, for the error report
        const float O = O2[p];
        if (O != INFINITY && sc == sc && sc != 0.0) {
          const double xn = a.st.xn[slot];
          double est;
          if (METRIC == kL2) {
            // (FP8 keys are the BF16-form key / sq; scale = sq there, 1 for BF16)
            const double sp = ((double)O - (double)qp.P * (double)a.w0[slot]) * qp.scale;
            est = sqrt(fmax(qp.nq2 + xn * xn + sp, 0.0));
          } else if (METRIC == kDot) {
            est = -((double)O - (double)qp.P * (double)a.w0[slot]) * qp.scale;
          } else {
            est = -(double)O / qp.qn;
          }
          err = fmax(err, fabs(est - sc) / fabs(sc));
        }
      }
      a.ids[(size_t)q * a.ostride + e] = id;
      a.scores[(size_t)q * a.ostride + e] = sc;
    }
  }
  for (int off = 32; off >= 1; off >>= 1) err = fmax(err, __shfl_xor(err, off));
  if (lane == 0) a.max_rel_err[q] = err;
}

// one wavefront per shortlist row; each range its own metric
__global__ __launch_bounds__(256) void k_rerank(RangeBatch<RerankArgs> b) {
  __shared__ double stage[4][64 * (kRC + 1)];
  __shared__ uint32_t queue[4][128];
  const int i = range_of(b, blockIdx.x);
  const uint32_t blk = blockIdx.x - b.first[i];
  double* stg = stage[threadIdx.x >> 6];
  uint32_t* qu = queue[threadIdx.x >> 6];
  switch (b.metric[i]) {
    case kL2: rerank_body<kL2>(b.r[i], blk, stg, qu); break;
    case kCos: rerank_body<kCos>(b.r[i], blk, stg, qu); break;
    default: rerank_body<kDot>(b.r[i], blk, stg, qu); break;
  }
}

// ranges of one metric: that metric's code only, held to 128 VGPRs for 4
// wavefronts per SIMD (the row gather is latency-bound; a few spilled
// dwords cost less than the lost occupancy: measured 12-20 % faster than
// the same code at 3 per SIMD)
template <int METRIC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_rerank1(
    RangeBatch<RerankArgs> b) {
  __shared__ double stage[4][64 * (kRC + 1)];
  __shared__ uint32_t queue[4][128];
  const int i = range_of(b, blockIdx.x);
  rerank_body<METRIC>(b.r[i], blockIdx.x - b.first[i], stage[threadIdx.x >> 6],
                      queue[threadIdx.x >> 6]);
}

hipError_t launch_rerank_ranges(const RerankArgs* a, const int* metric, int n, hipStream_t s) {
  for (int r0 = 0; r0 < n; r0 += kMaxLaunchRanges) {
    RangeBatch<RerankArgs> b{};
    uint32_t blocks = 0;
    b.n = 0;
    bool one = true;
    for (int r = r0; r < n && b.n < kMaxLaunchRanges; r++) {
      if (metric[r] != kL2 && metric[r] != kCos && metric[r] != kDot) return hipErrorInvalidValue;
      b.r[b.n] = a[r];
      b.metric[b.n] = metric[r];
      b.first[b.n] = blocks;
      blocks += (a[r].nq + 3) / 4;
      one &= metric[r] == metric[r0];
      b.n++;
    }
    b.first[b.n] = blocks;
    if (blocks == 0) continue;
    if (!one) hipLaunchKernelGGL(k_rerank, dim3(blocks), dim3(256), 0, s, b);
    else if (metric[r0] == kL2) hipLaunchKernelGGL(k_rerank1<kL2>, dim3(blocks), dim3(256), 0, s, b);
    else if (metric[r0] == kCos) hipLaunchKernelGGL(k_rerank1<kCos>, dim3(blocks), dim3(256), 0, s, b);
    else hipLaunchKernelGGL(k_rerank1<kDot>, dim3(blocks), dim3(256), 0, s, b);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_rerank(const RerankArgs& a, int metric, hipStream_t s) {
  return launch_rerank_ranges(&a, &metric, 1, s);
}

// ------------------------------------------------------------- fallback plumbing

// Rebuilds a shadow's degenerate-row list from the store: a live slot whose
// penalty is +inf holds a degenerate row (dead slots carry +inf too, and
// are skipped).  Slots deleted or rewritten since their row entered the
// list drop out, so the list holds only rows that are still degenerate.
// `count` must be zero on entry (k_zero_u32); the list's order is
// immaterial (the rerank ranks its rows by key and slot).
__global__ void k_zero_u32(uint32_t* __restrict__ p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__global__ void k_collect_degen(const float* __restrict__ pen, const uint64_t* __restrict__ alive,
                                uint32_t nslots, uint32_t* __restrict__ count,
                                uint32_t* __restrict__ list) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslots) return;
  if (!((alive[s >> 6] >> (s & 63)) & 1ull) || !(pen[s] == INFINITY)) return;
  const uint32_t idx = atomicAdd(count, 1u);
  if (idx < (uint32_t)kDegenCap) list[idx] = s;
}

hipError_t launch_rebuild_degen(const float* pen, const uint64_t* alive, uint32_t nslots,
                                uint32_t* count, uint32_t* list, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, count, 1u);
  if (nslots)
    hipLaunchKernelGGL(k_collect_degen, dim3((nslots + 255) / 256), dim3(256), 0, s, pen, alive,
                       nslots, count, list);
  return hipGetLastError();
}

__global__ void k_gather_rows(const double* __restrict__ src, size_t lds,
                              const uint32_t* __restrict__ idx, uint32_t n, uint32_t d,
                              double* __restrict__ dst, size_t ldd) {
  const uint32_t r = blockIdx.x;
  if (r >= n) return;
  const double* s = src + (size_t)idx[r] * lds;
  for (uint32_t c = threadIdx.x; c < d; c += blockDim.x) dst[(size_t)r * ldd + c] = s[c];
}

hipError_t launch_gather_rows(const double* src, size_t lds, const uint32_t* idx, uint32_t n,
                              uint32_t d, double* dst, size_t ldd, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_rows, dim3(n), dim3(64), 0, s, src, lds, idx, n, d, dst, ldd);
  return hipGetLastError();
}

__global__ void k_scatter_results(const uint64_t* __restrict__ ids, const double* __restrict__ sc,
                                  uint32_t k, const uint32_t* __restrict__ idx, uint32_t n,
                                  uint64_t* __restrict__ out_ids, double* __restrict__ out_sc,
                                  uint32_t ostride) {
  const uint32_t r = blockIdx.x;
  if (r >= n) return;
  const size_t o = (size_t)idx[r] * ostride;
  for (uint32_t j = threadIdx.x; j < k; j += blockDim.x) {
    out_ids[o + j] = ids[(size_t)r * k + j];
    out_sc[o + j] = sc[(size_t)r * k + j];
  }
}

hipError_t launch_scatter_results(const uint64_t* ids, const double* sc, uint32_t k,
                                  const uint32_t* idx, uint32_t n, uint64_t* out_ids,
                                  double* out_sc, uint32_t ostride, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter_results, dim3(n), dim3(64), 0, s, ids, sc, k, idx, n, out_ids,
                     out_sc, ostride);
  return hipGetLastError();
}

}  // namespace svdb
