// svdb_wide.hip — FP8 (OCP e4m3fn) shadow, and the first passes over wide
// rows: FP8 (padded widths 256..1536) and BF16 (512, 768, 1536) (gfx950).  Built
// with -ffp-contract=off; every fma below is explicit.
//
// Rows are stored as e4m3 codes of x / sx with a per-row power-of-two scale
// sx (max |x| / sx in (224, 448]), so x̂ = sx · e4m3(x / sx) exactly and
// every product below is an exact power-of-two rescaling of the BF16-form
// key: the bound of svdb_lowp.h holds with this shadow's ρ (≈ 2^-5 instead
// of ≈ 2^-9).  Queries are quantized the same way, with their own scale sq.
//
// Scan (k_scan_wide): queries held as B operands, row tiles streamed
// direct-to-LDS into two static buffers, XCD owning whole slices, per-lane
// 8-entry sorted lists, a branch-free per-fragment test among the next
// subtile's MFMAs and one rare wave-uniform insertion branch, with
// v_mfma_scale_f32_16x16x128_f8f6f4 (unit E8M0 scales): lane l supplies
// row (or query) l%16, bytes k = 32·(l/16) .. +31 of each 128-byte K step;
// the result has rows 4·(l/16)+r for query l%16.  Ordering keys (2 VALU ops):
//   Euclidean, dot: O = fma(P/sq, w0, acc·sx)   (w0 = ‖x‖² rd / ‖x‖ ru)
//   cosine:         O = fma(acc·sq, sx/‖x‖, pen)
// Both equal the BF16-form key divided by sq (Euclidean, dot) or equal it
// (cosine) exactly; QParams.scale carries sq back.
#include "svdb_lowp.h"

#include <math.h>

namespace svdb {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(3))) void lds_void;

// ------------------------------------------------------------- e4m3

// Round to nearest even onto e4m3fn; |v| <= 448 and finite.
__device__ __forceinline__ uint32_t e4m3_of(double v) {
  const uint32_t s = signbit(v) ? 0x80u : 0u;
  const double a = fabs(v);
  if (a < 0x1p-6) return s | (uint32_t)rint(a * 0x1p9);  // subnormal (8 -> min normal)
  uint64_t b = (uint64_t)__double_as_longlong(a);
  b += 0xFFFFFFFFFFFFull + ((b >> 49) & 1ull);
  const int e = (int)((b >> 52) & 0x7FF) - 1023;
  return s | ((uint32_t)(e + 7) << 3) | (uint32_t)((b >> 49) & 7ull);
}

__device__ __forceinline__ double e4m3_val(uint32_t x) {
  const int e = (x >> 3) & 15, m = x & 7;
  const double v = e ? ldexp(1.0 + m * 0.125, e - 7) : m * 0x1p-9;
  return (x & 0x80u) ? -v : v;
}

// 2^e with maxabs / 2^e in (224, 448]; maxabs > 0 finite.
__device__ __forceinline__ double pow2_scale(double maxabs) {
  int e;
  const double m = frexp(maxabs / 448.0, &e);  // maxabs/448 = m·2^e, m in [0.5, 1)
  return ldexp(1.0, m == 0.5 ? e - 1 : e);
}

bool fp8_dp_supported(size_t dp) {
  return dp == 256 || dp == 512 || dp == 768 || dp == 1024 || dp == 1536;
}

int fp8_queries_per_block(size_t dp) { return kLowpWaves * (dp <= 768 ? 2 : 1) * 16; }

// ------------------------------------------------------------- shadow rows

__global__ void k_shadow_init8(Shadow8 sh, uint32_t first, uint32_t n) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint32_t slot = first + r;
  sh.sx[slot] = 0.0f;
  sh.nx[slot] = INFINITY;
  sh.xn[slot] = -INFINITY;
  sh.inv[slot] = 0.0f;
  sh.pen[slot] = INFINITY;
}

hipError_t launch_shadow_init8(const Shadow8& sh, uint32_t first, uint32_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_shadow_init8, dim3((n + 255) / 256), dim3(256), 0, s, sh, first, n);
  return hipGetLastError();
}

// One wavefront per row: lane l packs the codes of elements 4l..4l+3 of
// each 256-element chunk into one 32-bit store.
__global__ __launch_bounds__(256) void k_quantize_rows8(Shadow8 sh, const double* __restrict__ X,
                                                        size_t ldx, const double* __restrict__ nx2,
                                                        const double* __restrict__ xnorm,
                                                        const uint32_t* __restrict__ slots,
                                                        uint32_t first, uint32_t n, uint32_t d) {
  const int lane = threadIdx.x & 63;
  const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= n) return;
  const uint32_t slot = slots ? slots[r] : first + r;
  const double* x = X + (size_t)slot * ldx;
  uint32_t* row = reinterpret_cast<uint32_t*>(sh.X8 + (size_t)slot * sh.dp);
  double maxabs = 0.0;
  bool finite = true;
  for (uint32_t i = lane; i < d; i += 64) {
    const double v = x[i];
    finite = finite && isfinite(v);
    maxabs = fmax(maxabs, fabs(v));
  }
  for (int o = 32; o >= 1; o >>= 1) maxabs = fmax(maxabs, __shfl_xor(maxabs, o));
  finite = __all(finite);
  const double norm = xnorm[slot];
  const double lim = 1152921504606846976.0;  // 2^60
  const bool degenerate = !finite || !(norm >= 1.0 / lim) || !(norm <= lim) || !(maxabs <= lim);
  const double sx = degenerate ? 1.0 : pow2_scale(maxabs);
  double esq = 0.0;
  for (uint32_t c0 = 0; c0 < sh.dp; c0 += 256) {
    uint32_t word = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint32_t i = c0 + 4 * lane + j;
      uint32_t code = 0;
      if (i < d && !degenerate) {
        const double v = x[i];
        code = e4m3_of(v / sx);
        const double e = e4m3_val(code) * sx - v;
        esq = fma(e, e, esq);
      }
      word |= code << (8 * j);
    }
    row[(c0 >> 2) + lane] = word;
  }
  for (int o = 32; o >= 1; o >>= 1) esq += __shfl_xor(esq, o);
  if (lane != 0) return;
  if (degenerate) {
    sh.sx[slot] = 0.0f;
    sh.nx[slot] = INFINITY;
    sh.xn[slot] = -INFINITY;
    sh.inv[slot] = 0.0f;
    sh.pen[slot] = INFINITY;
    uint32_t cnt = atomicAdd(sh.degen_count, 0u);
    bool present = false;
    for (uint32_t j = 0; j < cnt && j < (uint32_t)kDegenCap; j++) present |= sh.degen[j] == slot;
    if (!present) {
      uint32_t idx = atomicAdd(sh.degen_count, 1u);
      if (idx < (uint32_t)kDegenCap) sh.degen[idx] = slot;
    }
    return;
  }
  sh.sx[slot] = (float)sx;
  sh.nx[slot] = __double2float_rd(nx2[slot]);
  sh.xn[slot] = __double2float_ru(norm);
  sh.inv[slot] = __double2float_rn(sx / norm);
  sh.pen[slot] = 0.0f;
  // the sum above is in no fixed order: (1 + 2^-40) covers its rounding for
  // d <= 2^11 as it does for the sequential chain
  const double up = 1.0 + 0x1p-40;
  const double ratio = (sqrt(esq) * up) / norm * up;
  atomicMax(sh.rho_bits, __float_as_uint(__double2float_ru(ratio)));
}

hipError_t launch_quantize_rows8(const Shadow8& sh, const double* X, size_t ldx, const double* nx2,
                                 const double* xn, const uint32_t* slots, uint32_t first,
                                 uint32_t n, uint32_t d, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_quantize_rows8, dim3((n + 3) / 4), dim3(256), 0, s, sh, X, ldx, nx2, xn,
                     slots, first, n, d);
  return hipGetLastError();
}

__global__ void k_kill_rows8(Shadow8 sh, const uint32_t* __restrict__ slots, uint32_t n) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint32_t slot = slots[r];
  uint32_t* row = reinterpret_cast<uint32_t*>(sh.X8 + (size_t)slot * sh.dp);
  for (size_t i = 0; i < sh.dp / 4; i++) row[i] = 0;
  sh.sx[slot] = 0.0f;
  sh.nx[slot] = INFINITY;
  sh.xn[slot] = -INFINITY;
  sh.inv[slot] = 0.0f;
  sh.pen[slot] = INFINITY;
}

hipError_t launch_kill_rows8(const Shadow8& sh, const uint32_t* slots, uint32_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_kill_rows8, dim3((n + 63) / 64), dim3(64), 0, s, sh, slots, n);
  return hipGetLastError();
}

// ------------------------------------------------------------- query prep
//
// One thread per query: FP64 ‖q‖ (the canonical chain), e4m3 codes of
// q·sc / sq (sc = -2 Euclidean, -1 dot and cosine; sq its power-of-two
// scale), ĥ = sq·e4m3(q·sc/sq)/sc, and the bound constants of the BF16 form
// with ‖ĥ‖, ‖ĥ - q‖ and this shadow's ρ; then the power-of-two rescaling of
// P and scale.

__global__ void k_query_prep8(const double* __restrict__ Q, size_t ldq, uint32_t nq,
                              uint32_t nq_pad, uint32_t d, size_t dp, int metric,
                              const uint32_t* __restrict__ rho_bits, uint8_t* __restrict__ Q8,
                              QParams* __restrict__ qp) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq_pad) return;
  uint8_t* q8 = Q8 + (size_t)i * dp;
  QParams p{};
  p.metric = metric;
  if (i >= nq) {
    for (size_t c = 0; c < dp; c++) q8[c] = 0;
    p.flags = 1;
    qp[i] = p;
    return;
  }
  const double* q = Q + (size_t)i * ldq;
  const double sc = metric == kL2 ? -2.0 : -1.0;
  double nq2 = 0.0, maxabs = 0.0;
  bool finite = true;
  for (uint32_t c = 0; c < d; c++) {
    const double v = q[c];
    nq2 = fma(v, v, nq2);
    finite = finite && isfinite(v);
    maxabs = fmax(maxabs, fabs(v));
  }
  const double qn = sqrt(nq2);
  const double lim = 1152921504606846976.0;  // 2^60
  p.flags = (!finite || !(qn >= 1.0 / lim) || !(qn <= lim) || !(maxabs <= lim)) ? 1 : 0;
  const double sq = p.flags ? 1.0 : pow2_scale(maxabs * fabs(sc));
  double esq = 0.0, hsq = 0.0;
  for (uint32_t c = 0; c < d; c++) {
    const double v = q[c];
    const uint32_t code = p.flags ? 0u : e4m3_of(v * sc / sq);
    q8[c] = (uint8_t)code;
    const double h = e4m3_val(code) * sq / sc;
    const double e = h - v;
    esq = fma(e, e, esq);
    hsq = fma(h, h, hsq);
  }
  for (size_t c = d; c < dp; c++) q8[c] = 0;
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
  const float inv_sq = (float)(1.0 / sq);  // a power of two
  if (metric == kL2) {
    const double a = 1.0 - c / qn - 0x1p-24 * (1.0 + qh1 / qn) - 0x1p-30;
    p.P = __double2float_rd(a) * inv_sq;
    p.scale = sq;
    p.off = nq2 - qn * (c + 0x1p-24 * qh1) - 2.0 * delta;
    if (!(a > 0.0)) p.flags = 1;
  } else if (metric == kDot) {
    const double cp = (c + 0x1p-23 * (qh1 + c) * 1.01) * up;
    p.P = __double2float_rd(-cp) * inv_sq;
    p.scale = sq;
    p.off = -delta;
  } else {
    p.P = (float)sq;
    p.scale = 1.0 / qn;
    p.off = -(c + 0x1p-23 * qh1 * 1.01 + delta * lim) / qn - (double)(d + 16) * 0x1p-52;
  }
  qp[i] = p;
}

hipError_t launch_query_prep8(const double* Q, size_t ldq, uint32_t nq, uint32_t nq_pad,
                              uint32_t d, size_t dp, int metric, const uint32_t* rho_bits,
                              uint8_t* Q8, QParams* qp, hipStream_t s) {
  if (nq_pad == 0) return hipSuccess;
  hipLaunchKernelGGL(k_query_prep8, dim3((nq_pad + 127) / 128), dim3(128), 0, s, Q, ldq, nq,
                     nq_pad, d, dp, metric, rho_bits, Q8, qp);
  return hipGetLastError();
}

// ------------------------------------------------------------- scan

__device__ __forceinline__ void load_lds16(const void* g, void* lds_base) {
  __builtin_amdgcn_global_load_lds(g, (lds_void*)lds_base, 16, 0, 0);
}
__device__ __forceinline__ void load_lds4(const void* g, void* lds_base) {
  __builtin_amdgcn_global_load_lds(g, (lds_void*)lds_base, 4, 0, 0);
}

// KS K-steps of 128 bytes per row, NW wavefronts per block.  Tiles of TR
// rows (TR divides 128, the store's slot granularity) of at most 48 KB, two
// per block.
template <int KS, int NW = kLowpWaves>
struct Geom8 {
  static constexpr int RB = KS * 128;  // row bytes
  static constexpr int TR = KS <= 2 ? 128 : (KS <= 6 ? 64 : (KS <= 12 ? 32 : 16));
  static constexpr int CPR = RB / 16;  // 16-byte chunks per row
  static constexpr int TILE_B = TR * RB;
  static constexpr int AUX_B = TR * 4 * 2;
  static constexpr int BUF_B = TILE_B + AUX_B;
  static constexpr int CPT = TR * CPR / (64 * NW);  // chunk loads per lane
  static_assert(TR * CPR % (64 * NW) == 0, "tile chunks must split over the block");
  static_assert(BUF_B * 2 <= 160 * 1024, "two tiles must fit LDS");
  // LDS chunk of (row, column chunk c): XOR-swizzled within aligned groups
  // of 8 chunks, so 8 consecutive rows of one column chunk hit 8 distinct
  // 16-byte bank groups.
  __device__ static __forceinline__ int chunk(int row, int c) { return row * CPR + (c ^ (row & 7)); }
};

template <int KS, int NW>
__device__ __forceinline__ void issue_tile8(const LowpScanArgs& a, unsigned char* buf,
                                            uint32_t row0, int wave, int lane) {
  using G = Geom8<KS, NW>;
  const uint8_t* X8 = reinterpret_cast<const uint8_t*>(a.Xb);
#pragma unroll
  for (int j = 0; j < G::CPT; j++) {
    const int p0 = (wave + NW * j) * 64;  // first LDS chunk of this instruction
    const int p = p0 + lane;
    const int row = p / G::CPR, pos = p % G::CPR;
    const int c = pos ^ (row & 7);
    load_lds16(X8 + (size_t)(row0 + row) * G::RB + c * 16, buf + p0 * 16);
  }
  // per-row values: waves 0-1 w0, waves 2-3 w1, 64 rows per wave
  float* aux = reinterpret_cast<float*>(buf + G::TILE_B);
  if (wave < 4) {
    const int arr = wave >> 1, part = wave & 1;
    if (part * 64 < G::TR && part * 64 + lane < G::TR) {
      const float* w = arr ? a.w1 : a.w0;
      load_lds4(w + row0 + part * 64 + lane, aux + arr * G::TR + part * 64);
    }
  }
}

// BF: BF16 rows (64 elements per 128-byte K step, two
// v_mfma_f32_16x16x32_bf16 per step: lane l gives elements 8·(l/16)..+7 of
// each 32-element half; one-FMA keys as in k_scan_bf16); else FP8.
// NW = 4 (one wavefront per SIMD) leaves each wavefront 512 registers, for
// the 192-register query fragment of a 1536-wide BF16 row.
template <int METRIC, int KS, int QF, bool MINONLY, bool BF, int NW>
__global__ __launch_bounds__(64 * NW) void k_scan_wide(LowpScanArgs a) {
  using G = Geom8<KS, NW>;
  __shared__ __attribute__((aligned(16))) unsigned char buf0[G::BUF_B];
  __shared__ __attribute__((aligned(16))) unsigned char buf1[G::BUF_B];
  constexpr int LP = kLowpLP;
  constexpr bool COS = METRIC == kCos;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int ql = lane & 15, rg = lane >> 4;
  const uint32_t qpb = NW * QF * 16;
  const uint32_t nqb = (a.nq + qpb - 1) / qpb;
  const uint32_t xcd = blockIdx.x % 8, local = blockIdx.x / 8;
  const uint32_t qbase = (local % nqb) * qpb + wave * QF * 16;
  const uint32_t slice = (local / nqb) * 8 + xcd;
  const uint32_t r0 = slice * a.slice_stride;
  const uint32_t r1 = min(r0 + a.rows_per_slice, a.nrows_pad);
  const uint32_t ntiles = r1 > r0 ? (r1 - r0) / G::TR : 0;
  const bool active = qbase < a.nq;  // wave-uniform

  const uint8_t* Q8 = reinterpret_cast<const uint8_t*>(a.Qb);
  // B fragments: FP8 bytes 128·ks + 32·rg .. +31; BF16 the two 16-byte
  // pieces at 128·ks + 16·rg and 128·ks + 64 + 16·rg
  i32x8 B[QF][KS];
  float P[QF], tau[QF];
  float lk[QF][LP];
  uint32_t ls[QF][LP];
#pragma unroll
  for (int f = 0; f < QF; f++) {
    const uint32_t q = qbase + f * 16 + ql;  // < nq_pad (Q8 is padded)
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      if constexpr (BF) {
        const uint8_t* qrow = Q8 + (size_t)q * G::RB + rg * 16 + ks * 128;
        const i32x4 lo = *reinterpret_cast<const i32x4*>(qrow);
        const i32x4 hi = *reinterpret_cast<const i32x4*>(qrow + 64);
        B[f][ks] = i32x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
      } else {
        B[f][ks] = *reinterpret_cast<const i32x8*>(Q8 + (size_t)q * G::RB + rg * 32 + ks * 128);
      }
    }
    const QParams& qq = a.qp[min(q, a.nq - 1)];
    P[f] = qq.P;
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

  auto key = [&](const f32x4 (&acc)[QF], int f, int r, const float4& w0, const float4& w1) {
    const float v0 = r == 0 ? w0.x : r == 1 ? w0.y : r == 2 ? w0.z : w0.w;
    const float v1 = r == 0 ? w1.x : r == 1 ? w1.y : r == 2 ? w1.z : w1.w;
    return BF ? (COS ? fmaf(acc[f][r], v0, v1) : fmaf(P[f], v0, acc[f][r]))
              : (COS ? fmaf(acc[f][r] * P[f], v0, v1) : fmaf(P[f], v0, acc[f][r] * v1));
  };
  // Fragment f's common-case test: does any of its four keys beat the
  // list's last entry?  Branch-free VALU work, issued among the next
  // subtile's MFMAs (the minimum-only pre-pass folds its keys here).
  auto need_of = [&](const f32x4 (&acc)[QF], int f, const float4& w0, const float4& w1) {
    const float mn = fminf(fminf(key(acc, f, 0, w0, w1), key(acc, f, 1, w0, w1)),
                           fminf(key(acc, f, 2, w0, w1), key(acc, f, 3, w0, w1)));
    if constexpr (MINONLY) {
      tau[f] = fminf(tau[f], mn);
      return false;
    }
    return mn < tau[f];
  };
  // The rare insertion path recomputes the keys.
  auto insert_all = [&](const f32x4 (&acc)[QF], const float4& w0, const float4& w1,
                        uint32_t srow) {
#pragma unroll
    for (int f = 0; f < QF; f++) {
      float k4[4];
#pragma unroll
      for (int r = 0; r < 4; r++) k4[r] = key(acc, f, r, w0, w1);
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
  auto epilogue = [&](const f32x4 (&acc)[QF], const float4& w0, const float4& w1, uint32_t srow) {
    bool need = false;
#pragma unroll
    for (int f = 0; f < QF; f++) need |= need_of(acc, f, w0, w1);
    if (!MINONLY && __ballot(need) != 0) insert_all(acc, w0, w1, srow);
  };

  // The previous subtile's accumulators and per-row values: its test runs
  // among the next subtile's MFMAs, so the key work (VALU) overlaps MFMAs in
  // flight instead of leaving the matrix core idle — with one wavefront per
  // SIMD nothing else covers it.  The initial values make every key +inf
  // (Euclidean: P > 0 with w0 = +inf; dot: P < 0 with w0 = -inf; cosine:
  // w1 = +inf), so the first test passes nothing.
  f32x4 pacc[QF];
#pragma unroll
  for (int f = 0; f < QF; f++) pacc[f] = f32x4{0.f, 0.f, 0.f, 0.f};
  const float w0_init = COS ? 0.f : (METRIC == kDot ? -INFINITY : INFINITY);
  float4 pw0 = make_float4(w0_init, w0_init, w0_init, w0_init);
  const float w1_init = COS ? INFINITY : 1.f;
  float4 pw1 = make_float4(w1_init, w1_init, w1_init, w1_init);
  uint32_t prow = 0;

  if (ntiles > 0) issue_tile8<KS, NW>(a, buf0, r0, wave, lane);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (uint32_t t0 = 0; t0 < ntiles; t0 += 2) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const uint32_t t = t0 + h;
      if (t >= ntiles) break;
      const unsigned char* cur = h ? buf1 : buf0;
      if (t + 1 < ntiles) issue_tile8<KS, NW>(a, h ? buf0 : buf1, r0 + (t + 1) * G::TR, wave, lane);
      const uint32_t trow = r0 + t * G::TR;
      const float* aux = reinterpret_cast<const float*>(cur + G::TILE_B);
      // A operand of K step ks: bytes 128·ks + 32·rg .. +31 of row arow
      // (two swizzled 16-byte chunks).  The scheduling barriers keep each
      // step's reads where they are written; hoisting them all would spill.
      auto loadA = [&](int arow, int ks) {
        const int c0 = BF ? ks * 8 + rg : ks * 8 + rg * 2;
        const int c1 = BF ? c0 + 4 : c0 + 1;
        const i32x4 lo = *reinterpret_cast<const i32x4*>(cur + G::chunk(arow, c0) * 16);
        const i32x4 hi = *reinterpret_cast<const i32x4*>(cur + G::chunk(arow, c1) * 16);
        return i32x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
      };
      // The A operands stream through two register buffers across the
      // tile's subtiles: after global step g's MFMAs have issued, its buffer
      // takes step g+2's operand, so each read has a whole step of MFMAs to
      // arrive, and a subtile's first steps do not wait on LDS.
      constexpr int NSUB = G::TR / 16;
      static_assert(KS % 2 == 0, "the operand buffers alternate by K step");
      i32x8 Ab[2];
      if (active) {
        Ab[0] = loadA(ql, 0);
        Ab[1] = loadA(ql, 1);
      }
#pragma unroll
      for (int sub = 0; sub < (active ? NSUB : 0); sub++) {
        f32x4 acc[QF];
#pragma unroll
        for (int f = 0; f < QF; f++) acc[f] = f32x4{0.f, 0.f, 0.f, 0.f};
        // The previous subtile's test, one fragment per K step (the rest at
        // the last step): VALU work in the same basic block as this
        // subtile's MFMAs, issued between them — with one wavefront per
        // SIMD nothing else would fill the matrix core meanwhile.  One
        // wave-uniform, unlikely branch after the last step takes the rare
        // insertions out of line.
        bool need = false;
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
          const i32x8 A = Ab[ks & 1];
#pragma unroll
          for (int f = 0; f < QF; f++) {
            if constexpr (BF) {
              const i32x8 b = B[f][ks];
              const bf16x8 a0 = __builtin_bit_cast(bf16x8, i32x4{A[0], A[1], A[2], A[3]});
              const bf16x8 a1 = __builtin_bit_cast(bf16x8, i32x4{A[4], A[5], A[6], A[7]});
              const bf16x8 b0 = __builtin_bit_cast(bf16x8, i32x4{b[0], b[1], b[2], b[3]});
              const bf16x8 b1 = __builtin_bit_cast(bf16x8, i32x4{b[4], b[5], b[6], b[7]});
              acc[f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[f], 0, 0, 0);
              acc[f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[f], 0, 0, 0);
            } else {
              acc[f] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(A, B[f][ks], acc[f], 0, 0,
                                                                        0, 127, 0, 127);
            }
          }
#pragma unroll
          for (int f = 0; f < QF; f++)
            if (f == ks || (ks == KS - 1 && f > ks)) need |= need_of(pacc, f, pw0, pw1);
          const int nx = sub * KS + ks + 2;
          if (nx < NSUB * KS) Ab[ks & 1] = loadA((nx / KS) * 16 + ql, nx % KS);
          // one MFMA, then up to four VALU instructions, and so on
#pragma unroll
          for (int i = 0; i < QF * (BF ? 2 : 1); i++) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);
          }
          __builtin_amdgcn_sched_barrier(0);
        }
        if (!MINONLY && __builtin_expect(__ballot(need) != 0, 0))
          insert_all(pacc, pw0, pw1, prow);
        const int rl = sub * 16 + rg * 4;
#pragma unroll
        for (int f = 0; f < QF; f++) pacc[f] = acc[f];
        pw0 = *reinterpret_cast<const float4*>(aux + rl);
        pw1 = *reinterpret_cast<const float4*>(aux + G::TR + rl);
        prow = trow + rl;
      }
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
    }
  }
  if (active) epilogue(pacc, pw0, pw1, prow);
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

template <int METRIC, int KS, int QF, bool BF = false, int NW = kLowpWaves>
static hipError_t scan_fp8_t(const LowpScanArgs& a, uint32_t nslices, bool minonly,
                             hipStream_t s) {
  auto fn = minonly ? k_scan_wide<METRIC, KS, QF, true, BF, NW>
                    : k_scan_wide<METRIC, KS, QF, false, BF, NW>;
  const uint32_t qpb = NW * QF * 16;
  // nslices is a multiple of 8 (one set of slices per XCD); rows per slice
  // and the slice stride are multiples of the 128-row slot granularity.
  if (nslices % 8 || a.rows_per_slice % Geom8<KS, NW>::TR || a.nrows_pad % 128)
    return hipErrorInvalidValue;
  dim3 grid(((a.nq + qpb - 1) / qpb) * nslices);
  hipLaunchKernelGGL(fn, grid, dim3(64 * NW), 0, s, a);
  return hipGetLastError();
}

template <int METRIC>
static hipError_t scan_fp8_m(const LowpScanArgs& a, uint32_t nslices, bool minonly,
                             hipStream_t s) {
  // Up to 768 wide: 4 wavefronts of 64 queries (one per SIMD, query
  // fragments partly in AGPRs).  Every wavefront reads each tile row's
  // operand from LDS, so twice the queries per wavefront halves LDS traffic
  // per MFMA, which bounds this scan (768 wide: 27.9 -> 17.8 ms per 5M rows
  // x 4096 queries against 8 wavefronts of 32).  Wider rows, and the BF16
  // form, measured faster with 8 wavefronts.
  switch (a.dp) {
    case 256: return scan_fp8_t<METRIC, 2, 4, false, 4>(a, nslices, minonly, s);
    case 512: return scan_fp8_t<METRIC, 4, 4, false, 4>(a, nslices, minonly, s);
    case 768: return scan_fp8_t<METRIC, 6, 4, false, 4>(a, nslices, minonly, s);
    case 1024: return scan_fp8_t<METRIC, 8, 1>(a, nslices, minonly, s);
    case 1536: return scan_fp8_t<METRIC, 12, 1>(a, nslices, minonly, s);
  }
  return hipErrorInvalidValue;
}

template <int METRIC>
static hipError_t scan_bf16w_m(const LowpScanArgs& a, uint32_t nslices, bool minonly,
                               hipStream_t s) {
  switch (a.dp) {
    case 512: return scan_fp8_t<METRIC, 8, 1, true>(a, nslices, minonly, s);
    case 768: return scan_fp8_t<METRIC, 12, 1, true>(a, nslices, minonly, s);
    case 1536: return scan_fp8_t<METRIC, 24, 1, true, 4>(a, nslices, minonly, s);
  }
  return hipErrorInvalidValue;
}

bool bf16_wide_dp_supported(size_t dp) { return dp == 512 || dp == 768 || dp == 1536; }

hipError_t launch_scan_bf16_wide(const LowpScanArgs& a, int metric, uint32_t nslices,
                                 bool minonly, hipStream_t s) {
  if (a.nq == 0 || nslices == 0) return hipSuccess;
  switch (metric) {
    case kL2: return scan_bf16w_m<kL2>(a, nslices, minonly, s);
    case kCos: return scan_bf16w_m<kCos>(a, nslices, minonly, s);
    case kDot: return scan_bf16w_m<kDot>(a, nslices, minonly, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_scan_fp8(const LowpScanArgs& a, int metric, uint32_t nslices, bool minonly,
                           hipStream_t s) {
  if (a.nq == 0 || nslices == 0) return hipSuccess;
  switch (metric) {
    case kL2: return scan_fp8_m<kL2>(a, nslices, minonly, s);
    case kCos: return scan_fp8_m<kCos>(a, nslices, minonly, s);
    case kDot: return scan_fp8_m<kDot>(a, nslices, minonly, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace svdb
