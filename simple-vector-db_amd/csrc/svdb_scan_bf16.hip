// svdb_scan_bf16.hip — the BF16 MFMA first pass over the store's shadow
// (gfx950).  Built with -ffp-contract=off; every fma below is explicit.
//
// Block: 8 wavefronts, each holding QF fragments of 16 queries in registers
// as B operands of v_mfma_f32_16x16x32_bf16, so a block serves 128*QF
// queries.  Grid: query blocks × row slices, ordered so each XCD owns whole
// slices (the blocks of one slice share that XCD's L2).  A slice is
// streamed in 128-row tiles straight into LDS (global_load_lds_dwordx4, no
// registers), double-buffered, 16-byte chunks XOR-swizzled by row so the
// 16-row A-operand reads hit distinct banks.  Per 16-row subtile a lane
// gets 4 scores per fragment — rows 4*(lane/16)+r, query lane%16 — and its
// fp32 ordering key is one FMA of the score and a per-row value.  Keys below
// the lane's current 8th best enter the lane's private sorted list (an
// insertion into 8 registers); each list is one candidate list of its
// query and its 8th entry bounds every row the lane left out.  Subtiles are
// software-pipelined: subtile s+1's MFMAs are issued before subtile s's
// epilogue, whose per-row values were read into registers beforehand.
#include "svdb_lowp.h"

#include <math.h>

namespace svdb {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void;

// One 16-byte chunk per lane from global memory into LDS at
// lds_base + 16*lane (the LDS pointer must be wave-uniform).
__device__ __forceinline__ void load_lds16(const void* g, void* lds_base) {
  __builtin_amdgcn_global_load_lds(g, (lds_void*)lds_base, 16, 0, 0);
}
__device__ __forceinline__ void load_lds4(const void* g, void* lds_base) {
  __builtin_amdgcn_global_load_lds(g, (lds_void*)lds_base, 4, 0, 0);
}

template <int KS, int NAUX = 2>
struct TileGeom {
  static constexpr int CPR = 4 * KS;                 // 16-byte chunks per row
  static constexpr int TILE_B = kLowpTile * CPR * 16;
  static constexpr int AUX_B = kLowpTile * 4 * NAUX; // NAUX fp32 arrays
  static constexpr int BUF_B = TILE_B + AUX_B;
  static constexpr int CPW = kLowpTile * CPR / (64 * kLowpWaves);  // chunk-instrs per wave
  // LDS chunk of (row, column chunk c): swizzled within the row.
  __device__ static __forceinline__ int chunk(int row, int c) {
    return row * CPR + (c ^ (row & (CPR - 1)));
  }
};

// Issue the direct-to-LDS loads of one tile (rows row0..row0+127) into buf.
// NAUX per-row arrays follow the tile: w0, w1 (cosine), and in the mixed
// scan w2, w3; waves 2j, 2j+1 load array j.
template <int KS, int NAUX>
__device__ __forceinline__ void issue_tile(const LowpScanArgs& a, unsigned char* buf,
                                           uint32_t row0, int wave, int lane) {
  using G = TileGeom<KS, NAUX>;
#pragma unroll
  for (int j = 0; j < G::CPW; j++) {
    const int p0 = (wave + kLowpWaves * j) * 64;  // first LDS chunk of this instruction
    const int p = p0 + lane;
    const int row = p / G::CPR, pos = p % G::CPR;
    const int c = pos ^ (row & (G::CPR - 1));
    load_lds16(a.Xb + (size_t)(row0 + row) * a.dp + c * 8, buf + p0 * 16);
  }
  float* aux = reinterpret_cast<float*>(buf + G::TILE_B);
  const int arr = wave >> 1, part = wave & 1;
  if (arr < NAUX) {
    const float* w = arr == 0 ? a.w0 : arr == 1 ? a.w1 : arr == 2 ? a.w2 : a.w3;
    load_lds4(w + row0 + part * 64 + lane, aux + arr * kLowpTile + part * 64);
  }
}

// MINONLY: each lane keeps only the smallest key of its rows (one list
// entry, cand_key[q * NL + list]); the sample pre-pass uses it to bound the
// L-th key cheaply.
// METRIC kMixed: each wavefront's queries share one metric (QParams.metric
// of its first slot; the engine aligns metric ranges to wavefronts), read at
// run time; the tile then carries all four per-row arrays.
template <int METRIC, int KS, int QF, bool MINONLY>
__global__ __launch_bounds__(512) void k_scan_bf16(LowpScanArgs a) {
  constexpr bool MIXED = METRIC == kMixed;
  constexpr int NAUX = MIXED ? 4 : (METRIC == kCos ? 2 : 1);
  using G = TileGeom<KS, NAUX < 2 ? 2 : NAUX>;
  __shared__ __attribute__((aligned(16))) unsigned char buf0[G::BUF_B];
  __shared__ __attribute__((aligned(16))) unsigned char buf1[G::BUF_B];
  constexpr int LP = kLowpLP;
  constexpr bool COS = METRIC == kCos;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int ql = lane & 15, rg = lane >> 4;
  const uint32_t qpb = kLowpWaves * QF * 16;
  // XCD-aware block order: consecutive block ids go to the 8 XCDs in turn,
  // so ids ≡ x (mod 8) run on XCD x.  Each XCD gets whole slices
  // (slice ≡ x mod 8) and all query blocks of a slice follow one another
  // there, streaming the same rows through that XCD's L2.
  const uint32_t nqb = (a.nq + qpb - 1) / qpb;
  const uint32_t xcd = blockIdx.x % 8, local = blockIdx.x / 8;
  const uint32_t qbase = (local % nqb) * qpb + wave * QF * 16;
  const uint32_t slice = (local / nqb) * 8 + xcd;
  const uint32_t r0 = slice * a.slice_stride;
  const uint32_t r1 = min(r0 + a.rows_per_slice, a.nrows_pad);
  const uint32_t ntiles = r1 > r0 ? (r1 - r0) / kLowpTile : 0;
  const bool active = qbase < a.nq;  // wave-uniform
  // mixed scan: this wavefront's metric, and where its per-row array sits
  // in the tile's aux block (0 nx, 2 xn, 3 inv; 1 is pen)
  const int wm = MIXED ? a.qp[min(qbase, a.nq - 1)].metric : METRIC;
  const bool wcos = MIXED ? wm == kCos : COS;
  const int warr = MIXED ? (wm == kL2 ? 0 : wm == kDot ? 2 : 3) : 0;

  bf16x8 B[QF][KS];
  float P[QF], tau[QF];
  float lk[QF][LP];
  uint32_t ls[QF][LP];
#pragma unroll
  for (int f = 0; f < QF; f++) {
    const uint32_t q = qbase + f * 16 + ql;  // < nq_pad (Qb is padded)
    const uint16_t* qrow = a.Qb + (size_t)q * a.dp + rg * 8;
#pragma unroll
    for (int ks = 0; ks < KS; ks++) B[f][ks] = *reinterpret_cast<const bf16x8*>(qrow + ks * 32);
    const QParams& qq = a.qp[min(q, a.nq - 1)];
    P[f] = qq.P;
    // Padding lanes (q >= nq, or a flagged slot: a degenerate query or the
    // padding between metric ranges) never insert; real ones start from the
    // optional threshold (a NaN threshold counts as none).
    float t0 = -INFINITY;
    if (q < a.nq && !qq.flags) {
      t0 = INFINITY;
      if (a.tau0) {
        const float t = a.tau0[(size_t)q * a.tau0_stride];
        if (t < INFINITY) t0 = t;
      }
    }
    tau[f] = MINONLY ? INFINITY : t0;
#pragma unroll
    for (int j = 0; j < LP; j++) {
      lk[f][j] = t0;
      ls[f][j] = ~0u;
    }
  }

  // Pending subtile (MFMAs issued, epilogue not yet run).  It starts with
  // values whose keys are all +inf (L2: w = +inf with P > 0; dot: w = -inf
  // with P < 0; cosine: pen = +inf), so the epilogue runs unconditionally
  // and shares its basic block with the next subtile's MFMAs.
  f32x4 pacc[QF];
#pragma unroll
  for (int f = 0; f < QF; f++) pacc[f] = f32x4{0.f, 0.f, 0.f, 0.f};
  const float w_init = wcos ? 0.f : (wm == kDot ? -INFINITY : INFINITY);
  float4 pw = make_float4(w_init, w_init, w_init, w_init);
  const float p_init = wcos ? INFINITY : 0.f;
  float4 ppen = make_float4(p_init, p_init, p_init, p_init);
  uint32_t prow = 0;

  auto key = [&](const f32x4 (&acc)[QF], int f, int r, const float (&wv)[4],
                 const float (&pv)[4]) {
    return MIXED ? fmaf(wcos ? acc[f][r] : P[f], wv[r], wcos ? pv[r] : acc[f][r])
                 : (COS ? fmaf(acc[f][r], wv[r], pv[r]) : fmaf(P[f], wv[r], acc[f][r]));
  };
  // One test per subtile for all fragments: in the common case no key
  // beats its list's 8th entry and the insertion code is skipped with a
  // single branch; the rare insertion path recomputes the keys.
  auto need_of = [&](const f32x4 (&acc)[QF], const float (&wv)[4], const float (&pv)[4]) {
    bool need = false;
#pragma unroll
    for (int f = 0; f < QF; f++) {
      const float mn = fminf(fminf(key(acc, f, 0, wv, pv), key(acc, f, 1, wv, pv)),
                             fminf(key(acc, f, 2, wv, pv), key(acc, f, 3, wv, pv)));
      need |= mn < tau[f];
    }
    return need;
  };
  auto insert_all = [&](const f32x4 (&acc)[QF], const float (&wv)[4], const float (&pv)[4],
                        uint32_t srow) {
    if constexpr (MINONLY) {
#pragma unroll
      for (int f = 0; f < QF; f++) {
        const float mn = fminf(fminf(key(acc, f, 0, wv, pv), key(acc, f, 1, wv, pv)),
                               fminf(key(acc, f, 2, wv, pv), key(acc, f, 3, wv, pv)));
        tau[f] = fminf(tau[f], mn);
      }
      return;
    }
#pragma unroll
    for (int f = 0; f < QF; f++) {
      float k4[4];
#pragma unroll
      for (int r = 0; r < 4; r++) k4[r] = key(acc, f, r, wv, pv);
      const float mn = fminf(fminf(k4[0], k4[1]), fminf(k4[2], k4[3]));
      if (mn < tau[f]) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
          if (k4[r] < lk[f][LP - 1]) {
            lk[f][LP - 1] = k4[r];
            ls[f][LP - 1] = srow + r;
#pragma unroll
            for (int j = LP - 1; j > 0; j--) {
              const bool sw = lk[f][j] < lk[f][j - 1];
              const float tk = sw ? lk[f][j - 1] : lk[f][j];
              const uint32_t tsl = sw ? ls[f][j - 1] : ls[f][j];
              lk[f][j - 1] = sw ? lk[f][j] : lk[f][j - 1];
              ls[f][j - 1] = sw ? ls[f][j] : ls[f][j - 1];
              lk[f][j] = tk;
              ls[f][j] = tsl;
            }
          }
        }
        tau[f] = lk[f][LP - 1];
      }
    }
  };
  auto epilogue = [&](const f32x4 (&acc)[QF], const float4& w, const float4& pen, uint32_t srow) {
    const float wv[4] = {w.x, w.y, w.z, w.w};
    const float pv[4] = {pen.x, pen.y, pen.z, pen.w};
    // (the mixed scan has no registers to spare for the combined test)
    if constexpr (!MINONLY && !MIXED)
      if (!need_of(acc, wv, pv)) return;
    insert_all(acc, wv, pv, srow);
  };

  // Two LDS buffers as distinct objects, picked at compile time (tiles are
  // taken in pairs and the pair loop is unrolled): reads of one buffer then
  // need not wait for the direct-to-LDS loads still filling the other.
  if (ntiles > 0) issue_tile<KS, NAUX>(a, buf0, r0, wave, lane);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (uint32_t t0 = 0; t0 < ntiles; t0 += 2) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const uint32_t t = t0 + h;
      if (t >= ntiles) break;
      const unsigned char* cur = h ? buf1 : buf0;
      if (t + 1 < ntiles)
        issue_tile<KS, NAUX>(a, h ? buf0 : buf1, r0 + (t + 1) * kLowpTile, wave, lane);
      const uint32_t trow = r0 + t * kLowpTile;
      const float* aux = reinterpret_cast<const float*>(cur + G::TILE_B);
      // A wave whose queries all lie past nq (a small batch) only helps
      // fill the tiles.  L2 and dot at widths <= 128: the A fragments of the next subtile are
      // read into the registers of this subtile's as soon as the MFMAs that
      // use them have issued, so their LDS latency runs under the epilogue
      // test rather than before the next MFMAs (measured 1.7 % on dot at
      // 128 wide, same box; Euclidean within its run-to-run spread);
      // cosine and the mixed scan have no registers for it and read each
      // subtile's fragments at its start.
      constexpr bool PFA = KS <= 4 && (METRIC == kL2 || METRIC == kDot);
      if constexpr (PFA) {
        bf16x8 A[KS];
        if (active) {
#pragma unroll
          for (int ks = 0; ks < KS; ks++)
            A[ks] = *reinterpret_cast<const bf16x8*>(cur + G::chunk(ql, ks * 4 + rg) * 16);
        }
#pragma unroll
        for (int sub = 0; sub < (active ? kLowpTile / 16 : 0); sub++) {
          const int rl = sub * 16 + rg * 4;
          // this subtile's per-row values, read after its first MFMAs have
          // issued: they are used one subtile later, and read here, before
          // the MFMAs, the first MFMA's LDS wait covered their read too
          // (same-box A/B, Euclidean at 16,384 queries: -2 %)
          float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
          float4 pen = make_float4(0.f, 0.f, 0.f, 0.f);
          f32x4 acc[QF];
#pragma unroll
          for (int f = 0; f < QF; f++) acc[f] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int ks = 0; ks < KS; ks++) {
#pragma unroll
            for (int f = 0; f < QF; f++)
              acc[f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[ks], B[f][ks], acc[f], 0, 0, 0);
            // fragment ks is dead now: its registers take the next subtile's
            if (sub + 1 < kLowpTile / 16 || ks == 0) {
              __builtin_amdgcn_sched_barrier(0);
              if (sub + 1 < kLowpTile / 16)
                A[ks] = *reinterpret_cast<const bf16x8*>(
                    cur + G::chunk((sub + 1) * 16 + ql, ks * 4 + rg) * 16);
              if (ks == 0) {
                w = *reinterpret_cast<const float4*>(aux + warr * kLowpTile + rl);
                if (wcos) pen = *reinterpret_cast<const float4*>(aux + kLowpTile + rl);
              }
              __builtin_amdgcn_sched_barrier(0);
            }
          }
          // the previous subtile's epilogue: its common-case test is VALU work
          // in the same basic block as the MFMAs above, free to be scheduled
          // among them
          {
            const float wv[4] = {pw.x, pw.y, pw.z, pw.w};
            const float pv[4] = {ppen.x, ppen.y, ppen.z, ppen.w};
            bool need = true;
            if constexpr (!MINONLY && !MIXED) need = need_of(pacc, wv, pv);
            // a wave-uniform, unlikely branch: the insertion code is laid out
            // away from the loop and the common path falls through to the
            // next subtile with no exec-mask changes (same-box A/B: -0.7 %
            // Euclidean, -1.1 % cosine)
            if (__builtin_expect(__ballot(need) != 0, 0)) insert_all(pacc, wv, pv, prow);
          }
#pragma unroll
          for (int f = 0; f < QF; f++) pacc[f] = acc[f];
          pw = w;
          ppen = pen;
          prow = trow + rl;
        }
      } else {
#pragma unroll
        for (int sub = 0; sub < (active ? kLowpTile / 16 : 0); sub++) {
          const int arow = sub * 16 + ql;
          bf16x8 A[KS];
#pragma unroll
          for (int ks = 0; ks < KS; ks++)
            A[ks] = *reinterpret_cast<const bf16x8*>(cur + G::chunk(arow, ks * 4 + rg) * 16);
          const int rl = sub * 16 + rg * 4;
          // per-row values after the first MFMAs, as above, at widths <= 128
          // (at 256 the later read costs registers the kernel does not have)
          constexpr bool LATE = KS <= 4;
          float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
          float4 pen = make_float4(0.f, 0.f, 0.f, 0.f);
          if constexpr (!LATE) {
            w = *reinterpret_cast<const float4*>(aux + warr * kLowpTile + rl);
            if (wcos) pen = *reinterpret_cast<const float4*>(aux + kLowpTile + rl);
          }
          f32x4 acc[QF];
#pragma unroll
          for (int f = 0; f < QF; f++) acc[f] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int ks = 0; ks < KS; ks++) {
#pragma unroll
            for (int f = 0; f < QF; f++)
              acc[f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[ks], B[f][ks], acc[f], 0, 0, 0);
            if (LATE && ks == 0) {
              __builtin_amdgcn_sched_barrier(0);
              w = *reinterpret_cast<const float4*>(aux + warr * kLowpTile + rl);
              if (wcos) pen = *reinterpret_cast<const float4*>(aux + kLowpTile + rl);
              __builtin_amdgcn_sched_barrier(0);
            }
          }
          // the previous subtile's epilogue: its common-case test is VALU work
          // in the same basic block as the MFMAs above, free to be scheduled
          // among them
          {
            const float wv[4] = {pw.x, pw.y, pw.z, pw.w};
            const float pv[4] = {ppen.x, ppen.y, ppen.z, ppen.w};
            bool need = true;
            if constexpr (!MINONLY && !MIXED) need = need_of(pacc, wv, pv);
            if (__builtin_expect(__ballot(need) != 0, 0)) insert_all(pacc, wv, pv, prow);
          }
#pragma unroll
          for (int f = 0; f < QF; f++) pacc[f] = acc[f];
          pw = w;
          ppen = pen;
          prow = trow + rl;
        }
      }
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
    }
  }
  epilogue(pacc, pw, ppen, prow);

  const uint32_t list = slice * kLowpLanesPerQuery + rg;
#pragma unroll
  for (int f = 0; f < QF; f++) {
    const uint32_t q = qbase + f * 16 + ql;
    if (q >= a.nq) continue;
    if constexpr (MINONLY) {
      a.cand_key[(size_t)q * a.NL + list] = tau[f];
      continue;
    }
    const size_t o = ((size_t)q * a.NL + list) * LP;
#pragma unroll
    for (int j = 0; j < LP; j++) {
      a.cand_key[o + j] = lk[f][j];
      a.cand_slot[o + j] = ls[f][j];
    }
  }
}

template <int METRIC, int KS, int QF>
static hipError_t scan_bf16_t(const LowpScanArgs& a, uint32_t nslices, bool minonly,
                              hipStream_t s) {
  // the mixed scan has no pre-pass form: the engine runs the pre-pass per
  // metric range with the single-metric kernels
  if constexpr (METRIC == kMixed) {
    if (minonly) return hipErrorInvalidValue;
  }
  auto fn = k_scan_bf16<METRIC, KS, QF, false>;
  if constexpr (METRIC != kMixed)
    if (minonly) fn = k_scan_bf16<METRIC, KS, QF, true>;
  constexpr size_t shm = 0;  // both tile buffers are static LDS
  const uint32_t qpb = kLowpWaves * QF * 16;
  // nslices is a multiple of 8 (one set of slices per XCD); 1-D grid.
  if (nslices % 8) return hipErrorInvalidValue;
  dim3 grid(((a.nq + qpb - 1) / qpb) * nslices);
  hipLaunchKernelGGL(fn, grid, dim3(512), shm, s, a);
  return hipGetLastError();
}

template <int METRIC>
static hipError_t scan_bf16_m(const LowpScanArgs& a, uint32_t nslices, bool minonly,
                              hipStream_t s) {
  switch (a.dp / 32) {
    case 1: return scan_bf16_t<METRIC, 1, 4>(a, nslices, minonly, s);
    case 2: return scan_bf16_t<METRIC, 2, 4>(a, nslices, minonly, s);
    case 4: return scan_bf16_t<METRIC, 4, 4>(a, nslices, minonly, s);
    case 8: return scan_bf16_t<METRIC, 8, 2>(a, nslices, minonly, s);
  }
  return hipErrorInvalidValue;
}

int lowp_queries_per_block(size_t dp) {
  if (dp > 768) return 4 * 16;           // k_scan_wide, 4 wavefronts of one fragment
  if (dp > 256) return kLowpWaves * 16;  // k_scan_wide, one fragment per wavefront
  return dp / 32 == 8 ? kLowpWaves * 2 * 16 : kLowpQPB;
}

hipError_t launch_scan_bf16(const LowpScanArgs& a, int metric, uint32_t nslices, bool minonly,
                            hipStream_t s) {
  if (a.nq == 0 || nslices == 0) return hipSuccess;
  if (a.dp > 256) return launch_scan_bf16_wide(a, metric, nslices, minonly, s);
  switch (metric) {
    case kL2: return scan_bf16_m<kL2>(a, nslices, minonly, s);
    case kCos: return scan_bf16_m<kCos>(a, nslices, minonly, s);
    case kDot: return scan_bf16_m<kDot>(a, nslices, minonly, s);
    case kMixed: return scan_bf16_m<kMixed>(a, nslices, minonly, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace svdb
