// svdb_engine.cpp — device store, search pipeline and pair metrics.
#include "svdb_engine.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>

namespace svdb {

// ------------------------------------------------------------------ errors

static thread_local std::string g_err;
// a static message set without allocating (the allocation-free entry point's
// errors); takes precedence over g_err until the next set_error
static thread_local const char* g_err_static = nullptr;

void set_error(int code, const std::string& msg) {
  (void)code;
  g_err_static = nullptr;
  g_err = msg;
}

void set_error_static(int code, const char* msg) {
  (void)code;
  g_err_static = msg;
}

const char* last_error() { return g_err_static ? g_err_static : g_err.c_str(); }

// Countdown of the armed fault: 0 = disarmed; n > 0 = the n-th fault_point
// from now throws.  Process-wide.
static std::atomic<int64_t> g_fault_countdown{0};

void fault_point() {
  int64_t c = g_fault_countdown.load(std::memory_order_relaxed);
  while (c > 0) {
    if (g_fault_countdown.compare_exchange_weak(c, c - 1, std::memory_order_relaxed)) {
      if (c == 1) throw std::bad_alloc();
      return;
    }
  }
}

void arm_fault(int64_t n) { g_fault_countdown.store(n > 0 ? n : 0); }

int64_t fault_armed() { return g_fault_countdown.load(); }

#define HIPCHK(expr, what)                                                            \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      set_error(SVDB_EDEVICE, std::string(what) + ": " + hipGetErrorString(e_));      \
      return SVDB_EDEVICE;                                                            \
    }                                                                                 \
  } while (0)

static inline size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

// ------------------------------------------------------------------ device

static std::once_flag g_dev_once;
static int g_dev = -1;

int device_init() {
  std::call_once(g_dev_once, [] {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return;
    const char* e = std::getenv("SVDB_DEVICE");
    int d = e ? std::atoi(e) : 0;
    g_dev = (d >= 0 && d < n) ? d : 0;
  });
  if (g_dev < 0) {
    set_error(SVDB_EDEVICE, "no usable HIP device (libsvdb needs an MI355X / gfx950 GPU)");
    return SVDB_EDEVICE;
  }
  if (hipSetDevice(g_dev) != hipSuccess) {
    set_error(SVDB_EDEVICE, "hipSetDevice failed");
    return SVDB_EDEVICE;
  }
  return g_dev;
}

// ------------------------------------------------------------------ legacy metrics

namespace {
struct RefCtx {
  std::mutex mu;
  hipStream_t stream = nullptr;
  double* buf = nullptr;
  size_t cap = 0;  // doubles
  uint32_t* slots = nullptr;
  float* out = nullptr;
  uint64_t* host = nullptr;  // pinned: the one-launch form's (result, seq)
  uint64_t seq = 0;
};
RefCtx g_ref;
}  // namespace

int pair_ref32_host(const double* a_in, const double* b_in, size_t n, int metric, float* out) {
  if (device_init() < 0) return SVDB_EDEVICE;
  std::lock_guard<std::mutex> g(g_ref.mu);
  if (!g_ref.stream) {
    HIPCHK(hipStreamCreateWithFlags(&g_ref.stream, hipStreamNonBlocking), "stream");
    HIPCHK(hipMalloc(&g_ref.slots, 2 * sizeof(uint32_t)), "hipMalloc");
    HIPCHK(hipMalloc(&g_ref.out, sizeof(float)), "hipMalloc");
    uint32_t s[2] = {0, 1};
    HIPCHK(hipMemcpy(g_ref.slots, s, sizeof(s), hipMemcpyHostToDevice), "copy");
    HIPCHK(hipHostMalloc((void**)&g_ref.host, 64), "hipHostMalloc");
    g_ref.host[0] = g_ref.host[1] = 0;
  }
  if (n <= kRef32ArgsMax) {
    // one launch: the pair in the kernel arguments, the result polled from
    // pinned host memory (a stream wait after 50 ms reports a failed launch)
    PairRef32Args a;
    a.n = (uint32_t)n;
    a.metric = metric;
    a.out = g_ref.host;
    a.seq = ++g_ref.seq;
    for (size_t i = 0; i < n; i++) {
      a.ab[i] = a_in[i];
      a.ab[n + i] = b_in[i];
    }
    HIPCHK(launch_pair_ref32_args(a, g_ref.stream), "k_pair_ref32_args");
    volatile uint64_t* hv = g_ref.host;
    const auto t0 = std::chrono::steady_clock::now();
    while (hv[1] != a.seq) {
      if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) {
        HIPCHK(hipStreamSynchronize(g_ref.stream), "k_pair_ref32_args");
        if (hv[1] != a.seq) {
          set_error(SVDB_EDEVICE, "k_pair_ref32_args: no answer");
          return SVDB_EDEVICE;
        }
        break;
      }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    const uint32_t bits = (uint32_t)hv[0];
    std::memcpy(out, &bits, sizeof(float));
    return SVDB_OK;
  }
  size_t ld = std::max<size_t>(n, 1);
  if (2 * ld > g_ref.cap) {
    if (g_ref.buf) (void)hipFree(g_ref.buf);
    g_ref.buf = nullptr;
    HIPCHK(hipMalloc(&g_ref.buf, 2 * ld * sizeof(double)), "hipMalloc");
    g_ref.cap = 2 * ld;
  }
  if (n) {
    HIPCHK(hipMemcpyAsync(g_ref.buf, a_in, n * sizeof(double), hipMemcpyHostToDevice, g_ref.stream), "copy");
    HIPCHK(hipMemcpyAsync(g_ref.buf + ld, b_in, n * sizeof(double), hipMemcpyHostToDevice, g_ref.stream), "copy");
  }
  HIPCHK(launch_pair_ref32(g_ref.buf, ld, g_ref.slots, g_ref.slots + 1, 1, (uint32_t)n, metric,
                           g_ref.out, g_ref.stream), "k_pair_ref32");
  HIPCHK(hipMemcpyAsync(out, g_ref.out, sizeof(float), hipMemcpyDeviceToHost, g_ref.stream), "copy");
  HIPCHK(hipStreamSynchronize(g_ref.stream), "sync");
  return SVDB_OK;
}

// ------------------------------------------------------------------ engine

Engine::Engine(size_t initial_slots) {
  if (device_init() < 0) return;
  device_ = g_dev;
  if (hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) != hipSuccess) {
    set_error(SVDB_EDEVICE, "hipStreamCreate failed");
    return;
  }
  for (auto& c : ctx_)
    if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) {
      set_error(SVDB_EDEVICE, "hipStreamCreate failed");
      return;
    }
  init_slots_ = round_up(std::max<size_t>(initial_slots, 1024), 64);
  // Graph replay of repeated exact searches is opt-in: on ROCm 7.2 one
  // replay of the one-query exact call measured slower than its direct
  // launches (62.5 against 57.4 µs per kdtree_nearest call, 10k rows).
  const char* gv = getenv("SVDB_GRAPHS");
  graphs_on_ = gv && gv[0] == '1';
  ok_ = true;
}

Engine::~Engine() {
  if (!ok_) return;
  (void)hipSetDevice(device_);
  (void)hipStreamSynchronize(stream_);
  void* bufs[] = {X_, nx2_, xn_, alive_, rank_, sh_.Xb, sh_.nx, sh_.xn, sh_.inv, sh_.pen,
                  sh_.rho_bits /* degen_count and degen live inside this block */, sh8_.X8,
                  sh8_.sx, sh8_.nx, sh8_.xn, sh8_.inv, sh8_.pen, sh8_.rho_bits};
  if (near_dev_) (void)hipFree(near_dev_);
  if (near_host_) (void)hipHostFree(near_host_);
  for (auto& c : ctx_) {
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    void* cb[] = {c.s_q_, c.s_qn_, c.s_k_, c.s_pk_, c.s_ps_, c.s_ids_, c.s_sc_, c.s_pair_,
                  c.s_pout_, c.s_qb_, c.s_qp_, c.s_ck_, c.s_cs_, c.s_sk_, c.s_ss_, c.s_beta_,
                  c.s_fail_, c.s_err_, c.s_lk_, c.s_ls_, c.s_km_, c.s_fq_, c.s_fid_, c.s_fsc_, c.s_fk_, c.s_sk2_, c.s_ss2_,
                  c.s_beta2_, c.s_fail2_, c.s_cidx_, c.s_cq_, c.s_cid_, c.s_csc_};
    for (void* p : cb)
      if (p) (void)hipFree(p);
    if (c.h_fail_) (void)hipHostFree(c.h_fail_);
    if (c.h_io_) (void)hipHostFree(c.h_io_);
    if (c.h_stage_) (void)hipHostFree(c.h_stage_);
    for (auto& ps : c.prof_pending_) {
      (void)hipEventDestroy(ps.a);
      (void)hipEventDestroy(ps.b);
    }
    for (auto& gr : c.graphs)
      if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
    if (c.stream) (void)hipStreamDestroy(c.stream);
  }
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (h_stage_) (void)hipHostFree(h_stage_);
  for (auto& e : ev_pool_) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(stream_);
}

int Engine::dev_alloc(void** p, size_t bytes) {
  if (dev_bytes_ + bytes > budget_) {
    set_error(SVDB_ENOMEM, "device-memory budget exceeded (" + std::to_string(dev_bytes_ + bytes) +
                               " > " + std::to_string(budget_) + " bytes)");
    return SVDB_ENOMEM;
  }
  if (hipMalloc(p, std::max<size_t>(bytes, 256)) != hipSuccess) {
    *p = nullptr;
    set_error(SVDB_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    return SVDB_ENOMEM;
  }
  dev_bytes_ += bytes;
  return SVDB_OK;
}

void Engine::dev_free(void* p, size_t bytes) {
  if (!p) return;
  (void)hipFree(p);
  dev_bytes_ -= std::min(bytes, dev_bytes_);
}

// Grow-only device scratch of the current search context: a buffer being
// replaced may still be read by that context's queued kernels.
int Engine::scratch(void** p, size_t* have, size_t need) {
  if (need <= *have && *p) return SVDB_OK;
  HIPCHK(hipStreamSynchronize(c_->stream), "sync");
  scratch_gen_++;  // captured launch sequences that used the old buffer are stale
  dev_free(*p, *have);
  *p = nullptr;
  *have = 0;
  size_t n = round_up(std::max<size_t>(need, 4096), 4096);
  int rc = dev_alloc(p, n);
  if (rc) return rc;
  *have = n;
  return SVDB_OK;
}

// Pinned host staging shared by row uploads and searches (both run under
// mu_ and synchronise the stream before returning, so the buffer is free on
// entry).  Searches stage their query rows, k values, ids and scores here:
// copies from pinned memory are plain DMAs, where pageable ones go through
// the runtime's bounce buffers and cost tens of microseconds each, most of a
// one-query call.  Requests above kMaxStage return nullptr and the caller
// uses pageable vectors instead.
static constexpr size_t kMaxStage = 256ull << 20;

void* Engine::host_stage(void** buf, size_t* have, size_t bytes) {
  if (bytes > kMaxStage) return nullptr;
  if (*buf && *have >= bytes) return *buf;
  scratch_gen_++;
  if (*buf) (void)hipHostFree(*buf);
  *buf = nullptr;
  *have = 0;
  const size_t n = round_up(std::max<size_t>(bytes, 1u << 20), 1u << 20);
  if (hipHostMalloc(buf, n) != hipSuccess) {
    *buf = nullptr;
    (void)hipGetLastError();
    return nullptr;
  }
  *have = n;
  return *buf;
}

int Engine::ensure_dim(size_t d) {
  if (d == 0) {
    set_error(SVDB_EINVAL, "vector dimension 0");
    return SVDB_EINVAL;
  }
  if (d_ == 0) {
    d_ = d;
    ld_ = round_up(d, 8);
    // BF16 shadow for the MFMA first pass: rows padded to 32, 64, 128 or
    // 256 elements (the scan keeps a wavefront's queries in registers).
    size_t ks = 1;
    while (32 * ks < d) ks *= 2;
    // wider rows: the wide kernel's padded widths 512, 768, 1536 (0: none)
    sh_.dp = d <= 256 ? 32 * ks : d <= 512 ? 512 : d <= 768 ? 768 : d <= 1536 ? 1536 : 0;
    const char* off = std::getenv("SVDB_NO_SHADOW");
    lowp_ = lowp_wanted_ && sh_.dp != 0 && !(off && *off && *off != '0');
    // FP8 rows: the smallest supported padded length (0: no FP8 first pass)
    for (size_t c : {256, 512, 768, 1024, 1536})
      if (d <= c) {
        sh8_.dp = c;
        break;
      }
    const char* pp = std::getenv("SVDB_PREPASS");
    prepass_ = !(pp && *pp == '0');
    return ensure_capacity(init_slots_);
  }
  if (d != d_) {
    set_error(SVDB_EINVAL, "vector dimension " + std::to_string(d) + " differs from the store's " +
                               std::to_string(d_));
    return SVDB_EINVAL;
  }
  return SVDB_OK;
}

int Engine::ensure_capacity(size_t need) {
  if (need <= cap_) return SVDB_OK;
  if (need >= (size_t)UINT32_MAX - 64) {
    set_error(SVDB_EINVAL, "store is limited to 2^32-65 slots");
    return SVDB_EINVAL;
  }
  size_t nc = round_up(std::max({need, cap_ * 2, init_slots_}), 128);
  nc = std::min(nc, (size_t)UINT32_MAX / 128 * 128 - 128);
  const size_t nw = nc / 64, ow = cap_ / 64;
  // Host bitmap first: if its growth throws, nothing has changed yet (cap_
  // never runs ahead of h_alive_ / h_rank_, which append indexes by slot).
  fault_point();
  h_alive_.resize(nw, 0);
  h_rank_.resize(nw, 0);
  if (lowp_) {
    int rs = alloc_shadow(nc);
    if (rs) return rs;
  }
  if (lowp8_) {
    int rs = alloc_shadow8(nc);
    if (rs) return rs;
  }
  double *X = nullptr, *nx2 = nullptr, *xn = nullptr;
  uint64_t* al = nullptr;
  uint32_t* rk = nullptr;
  int rc;
  if ((rc = dev_alloc((void**)&X, nc * ld_ * sizeof(double)))) return rc;
  if ((rc = dev_alloc((void**)&nx2, nc * sizeof(double))) ||
      (rc = dev_alloc((void**)&xn, nc * sizeof(double))) ||
      (rc = dev_alloc((void**)&al, nw * sizeof(uint64_t))) ||
      (rc = dev_alloc((void**)&rk, nw * sizeof(uint32_t)))) {
    dev_free(X, nc * ld_ * sizeof(double));
    dev_free(nx2, nc * sizeof(double));
    dev_free(xn, nc * sizeof(double));
    dev_free(al, nw * sizeof(uint64_t));
    return rc;
  }
  if (nphys_) {
    HIPCHK(hipMemcpyAsync(X, X_, nphys_ * ld_ * sizeof(double), hipMemcpyDeviceToDevice, stream_), "grow");
    HIPCHK(hipMemcpyAsync(nx2, nx2_, nphys_ * sizeof(double), hipMemcpyDeviceToDevice, stream_), "grow");
    HIPCHK(hipMemcpyAsync(xn, xn_, nphys_ * sizeof(double), hipMemcpyDeviceToDevice, stream_), "grow");
  }
  HIPCHK(hipStreamSynchronize(stream_), "grow");
  dev_free(X_, cap_ * ld_ * sizeof(double));
  dev_free(nx2_, cap_ * sizeof(double));
  dev_free(xn_, cap_ * sizeof(double));
  dev_free(alive_, ow * sizeof(uint64_t));
  dev_free(rank_, ow * sizeof(uint32_t));
  X_ = X;
  nx2_ = nx2;
  xn_ = xn;
  alive_ = al;
  rank_ = rk;
  cap_ = nc;
  dirty_word_ = 0;  // the new bitmap arrays hold nothing yet
  return SVDB_OK;
}

// (Re)allocates the BF16 shadow for `nc` slots: new slots start "dead"
// (zero row, aux values that make every ordering key +inf); the first
// nphys_ slots are copied from the old shadow.
int Engine::alloc_shadow(size_t nc) {
  Shadow n = sh_;
  int rc;
  const size_t xb = nc * sh_.dp * sizeof(uint16_t), fb = nc * sizeof(float);
  if ((rc = dev_alloc((void**)&n.Xb, xb))) return rc;
  if ((rc = dev_alloc((void**)&n.nx, fb)) || (rc = dev_alloc((void**)&n.xn, fb)) ||
      (rc = dev_alloc((void**)&n.inv, fb)) || (rc = dev_alloc((void**)&n.pen, fb))) {
    dev_free(n.Xb, xb);
    if (n.nx != sh_.nx) dev_free(n.nx, fb);
    if (n.xn != sh_.xn) dev_free(n.xn, fb);
    if (n.inv != sh_.inv) dev_free(n.inv, fb);
    return rc;
  }
  if (!sh_.rho_bits) {
    if ((rc = dev_alloc((void**)&n.rho_bits, 3 * sizeof(uint32_t) + kDegenCap * sizeof(uint32_t))))
      return rc;
    n.degen_count = n.rho_bits + 1;
    n.degen = n.rho_bits + 3;
    HIPCHK(hipMemsetAsync(n.rho_bits, 0, 3 * sizeof(uint32_t), stream_), "memset");
  }
  HIPCHK(hipMemsetAsync(n.Xb, 0, xb, stream_), "memset");
  HIPCHK(launch_shadow_init(n, 0, (uint32_t)nc, stream_), "k_shadow_init");
  if (sh_.Xb && nphys_) {
    HIPCHK(hipMemcpyAsync(n.Xb, sh_.Xb, nphys_ * sh_.dp * sizeof(uint16_t), hipMemcpyDeviceToDevice, stream_), "grow");
    HIPCHK(hipMemcpyAsync(n.nx, sh_.nx, nphys_ * sizeof(float), hipMemcpyDeviceToDevice, stream_), "grow");
    HIPCHK(hipMemcpyAsync(n.xn, sh_.xn, nphys_ * sizeof(float), hipMemcpyDeviceToDevice, stream_), "grow");
    HIPCHK(hipMemcpyAsync(n.inv, sh_.inv, nphys_ * sizeof(float), hipMemcpyDeviceToDevice, stream_), "grow");
    HIPCHK(hipMemcpyAsync(n.pen, sh_.pen, nphys_ * sizeof(float), hipMemcpyDeviceToDevice, stream_), "grow");
  }
  HIPCHK(hipStreamSynchronize(stream_), "grow");
  if (sh_.Xb) {
    const size_t oxb = cap_ * sh_.dp * sizeof(uint16_t), ofb = cap_ * sizeof(float);
    dev_free(sh_.Xb, oxb);
    dev_free(sh_.nx, ofb);
    dev_free(sh_.xn, ofb);
    dev_free(sh_.inv, ofb);
    dev_free(sh_.pen, ofb);
  }
  sh_ = n;
  return SVDB_OK;
}

// Quantizes slots [first, first+n) (or the device slot list) into the shadow.
int Engine::shadow_rows(uint32_t first, uint32_t n, const uint32_t* dslots) {
  if (n == 0) return SVDB_OK;
  if (lowp_)
    HIPCHK(launch_quantize_rows(sh_, X_, ld_, nx2_, xn_, dslots, first, n, (uint32_t)d_, stream_),
           "k_quantize_rows");
  if (lowp8_)
    HIPCHK(launch_quantize_rows8(sh8_, X_, ld_, nx2_, xn_, dslots, first, n, (uint32_t)d_, stream_),
           "k_quantize_rows8");
  return SVDB_OK;
}

// (Re)allocates the FP8 shadow for `nc` slots, as alloc_shadow does the BF16 one.
int Engine::alloc_shadow8(size_t nc) {
  Shadow8 n = sh8_;
  int rc;
  const size_t xb = nc * sh8_.dp, fb = nc * sizeof(float);
  float** arrs[5] = {&n.sx, &n.nx, &n.xn, &n.inv, &n.pen};
  if ((rc = dev_alloc((void**)&n.X8, xb))) return rc;
  for (int i = 0; i < 5; i++)
    if ((rc = dev_alloc((void**)arrs[i], fb))) {
      dev_free(n.X8, xb);
      for (int j = 0; j < i; j++) dev_free(*arrs[j], fb);
      return rc;
    }
  if (!sh8_.rho_bits) {
    if ((rc = dev_alloc((void**)&n.rho_bits, 3 * sizeof(uint32_t) + kDegenCap * sizeof(uint32_t))))
      return rc;
    n.degen_count = n.rho_bits + 1;
    n.degen = n.rho_bits + 3;
    HIPCHK(hipMemsetAsync(n.rho_bits, 0, 3 * sizeof(uint32_t), stream_), "memset");
  }
  HIPCHK(hipMemsetAsync(n.X8, 0, xb, stream_), "memset");
  HIPCHK(launch_shadow_init8(n, 0, (uint32_t)nc, stream_), "k_shadow_init8");
  if (sh8_.X8 && nphys_) {
    HIPCHK(hipMemcpyAsync(n.X8, sh8_.X8, nphys_ * sh8_.dp, hipMemcpyDeviceToDevice, stream_), "grow");
    const float* olds[5] = {sh8_.sx, sh8_.nx, sh8_.xn, sh8_.inv, sh8_.pen};
    for (int i = 0; i < 5; i++)
      HIPCHK(hipMemcpyAsync(*arrs[i], olds[i], nphys_ * sizeof(float), hipMemcpyDeviceToDevice,
                            stream_), "grow");
  }
  HIPCHK(hipStreamSynchronize(stream_), "grow");
  if (sh8_.X8) {
    const size_t ofb = cap_ * sizeof(float);
    dev_free(sh8_.X8, cap_ * sh8_.dp);
    float* olds[5] = {sh8_.sx, sh8_.nx, sh8_.xn, sh8_.inv, sh8_.pen};
    for (int i = 0; i < 5; i++) dev_free(olds[i], ofb);
  }
  sh8_ = n;
  return SVDB_OK;
}

int Engine::enable_shadow(int precision) {
  std::lock_guard<std::mutex> g(mu_);
  store_ver_++;
  if (int rc0 = complete_pending()) return rc0;
  if (precision != SVDB_PREC_BF16 && precision != SVDB_PREC_FP8 && precision != SVDB_PREC_AUTO) {
    set_error(SVDB_EINVAL, "enable_shadow takes SVDB_PREC_BF16, SVDB_PREC_FP8 or SVDB_PREC_AUTO");
    return SVDB_EINVAL;
  }
  const bool fp8 = precision == SVDB_PREC_FP8;
  if (!fp8) lowp_wanted_ = true;
  if (d_ == 0) {
    if (fp8) {
      set_error(SVDB_EINVAL, "the FP8 shadow is built once the store has rows");
      return SVDB_EINVAL;
    }
    return SVDB_OK;
  }
  if (fp8 ? lowp8_ : lowp_) return SVDB_OK;
  if (!fp8 && sh_.dp == 0) {
    set_error(SVDB_EINVAL, "BF16 first pass supports dimensions up to 1536");
    return SVDB_EINVAL;
  }
  if (fp8 && !fp8_dp_supported(sh8_.dp)) {
    set_error(SVDB_EINVAL, "FP8 first pass supports dimensions up to 1536");
    return SVDB_EINVAL;
  }
  int rc = fp8 ? alloc_shadow8(cap_) : alloc_shadow(cap_);
  if (rc) return rc;
  // quantize the existing rows into this shadow only
  const bool had_b = lowp_, had_8 = lowp8_;
  lowp_ = !fp8;
  lowp8_ = fp8;
  rc = shadow_rows(0, (uint32_t)nphys_, nullptr);
  lowp_ = had_b || !fp8;
  lowp8_ = had_8 || fp8;
  if (rc) return rc;
  for (size_t s = 0; s < nphys_; s++)
    if (!((h_alive_[s >> 6] >> (s & 63)) & 1ull)) {
      if (fp8) {
        HIPCHK(hipMemsetAsync(sh8_.X8 + s * sh8_.dp, 0, sh8_.dp, stream_), "memset");
        HIPCHK(launch_shadow_init8(sh8_, (uint32_t)s, 1, stream_), "k_shadow_init8");
      } else {
        HIPCHK(hipMemsetAsync(sh_.Xb + s * sh_.dp, 0, sh_.dp * sizeof(uint16_t), stream_), "memset");
        HIPCHK(launch_shadow_init(sh_, (uint32_t)s, 1, stream_), "k_shadow_init");
      }
    }
  HIPCHK(hipStreamSynchronize(stream_), "enable_shadow");
  return SVDB_OK;
}

int Engine::sync_alive() {
  const size_t nw = (nphys_ + 63) / 64;
  if (dirty_word_ >= nw) {
    dirty_word_ = SIZE_MAX;
    return SVDB_OK;
  }
  const size_t w0 = dirty_word_, n = nw - w0;
  HIPCHK(hipMemcpyAsync(alive_ + w0, h_alive_.data() + w0, n * sizeof(uint64_t),
                        hipMemcpyHostToDevice, c_->stream), "alive upload");
  HIPCHK(hipMemcpyAsync(rank_ + w0, h_rank_.data() + w0, n * sizeof(uint32_t),
                        hipMemcpyHostToDevice, c_->stream), "rank upload");
  dirty_word_ = SIZE_MAX;
  return SVDB_OK;
}

// After deletes or rewrites, the shadows' degenerate lists are rebuilt
// from the live slots before the next low-precision search (call after
// sync_alive: the kernel reads the device bitmap).
int Engine::refresh_degen() {
  if (!degen_dirty_) return SVDB_OK;
  if (lowp_)
    HIPCHK(launch_rebuild_degen(sh_.pen, alive_, (uint32_t)nphys_, sh_.degen_count, sh_.degen,
                                c_->stream), "degenerate list");
  if (lowp8_)
    HIPCHK(launch_rebuild_degen(sh8_.pen, alive_, (uint32_t)nphys_, sh8_.degen_count, sh8_.degen,
                                c_->stream), "degenerate list");
  degen_dirty_ = false;
  return SVDB_OK;
}

uint32_t Engine::select_slot(size_t logical) const {
  if (!has_deletes_) return (uint32_t)logical;
  const size_t nw = (nphys_ + 63) / 64;
  size_t w = (size_t)(std::upper_bound(h_rank_.begin(), h_rank_.begin() + nw, (uint32_t)logical) -
                      h_rank_.begin()) - 1;
  uint64_t word = h_alive_[w];
  size_t j = logical - h_rank_[w];
  for (size_t t = 0; t < j; t++) word &= word - 1;
  return (uint32_t)(w * 64 + __builtin_ctzll(word));
}

int Engine::slot_of(size_t logical, uint32_t* slot) {
  std::lock_guard<std::mutex> g(mu_);
  if (logical >= nalive_) {
    set_error(SVDB_EINVAL, "index out of range");
    return SVDB_EINVAL;
  }
  *slot = select_slot(logical);
  return SVDB_OK;
}

int Engine::append(const double* rows, size_t n, size_t lds, size_t dim, uint32_t* first_slot) {
  std::lock_guard<std::mutex> g(mu_);
  store_ver_++;
  if (!ok_) return SVDB_EDEVICE;
  if (hipSetDevice(device_) != hipSuccess) return SVDB_EDEVICE;
  // searches in flight were submitted before this mutation: they finish
  // against the store as it was
  if (int rc0 = complete_pending()) return rc0;
  int rc = ensure_dim(dim);
  if (rc) return rc;
  if ((rc = ensure_capacity(nphys_ + n))) return rc;
  const size_t s0 = nphys_;
  if (n) {
    // Rows go through a pinned staging buffer in large chunks, already in
    // the device layout (stride ld, zero padding), one DMA per chunk.
    const size_t row_b = ld_ * sizeof(double);
    const size_t chunk = std::max<size_t>(1, std::min<size_t>(n, (64u << 20) / row_b));
    double* st = static_cast<double*>(host_stage(&h_stage_, &h_stage_n_, chunk * row_b));
    if (!st) {
      set_error(SVDB_EDEVICE, "hipHostMalloc failed");
      return SVDB_EDEVICE;
    }
    for (size_t r0 = 0; r0 < n; r0 += chunk) {
      const size_t m = std::min(chunk, n - r0);
      for (size_t r = 0; r < m; r++) {
        std::memcpy(st + r * ld_, rows + (r0 + r) * lds, d_ * sizeof(double));
        if (ld_ != d_) std::memset(st + r * ld_ + d_, 0, (ld_ - d_) * sizeof(double));
      }
      HIPCHK(hipMemcpyAsync(X_ + (s0 + r0) * ld_, st, m * row_b, hipMemcpyHostToDevice, stream_),
             "row upload");
      HIPCHK(hipStreamSynchronize(stream_), "row upload");
    }
    HIPCHK(launch_row_norms(X_, ld_, nullptr, (uint32_t)s0, (uint32_t)n, (uint32_t)d_, nx2_, xn_,
                            stream_), "k_row_norms");
    if ((rc = shadow_rows((uint32_t)s0, (uint32_t)n, nullptr))) return rc;
  }
  for (size_t s = s0; s < s0 + n; s++) {
    if ((s & 63) == 0) h_rank_[s >> 6] = (uint32_t)nalive_;
    h_alive_[s >> 6] |= 1ull << (s & 63);
    nalive_++;
  }
  dirty_word_ = std::min(dirty_word_, s0 / 64);
  nphys_ += n;
  HIPCHK(hipStreamSynchronize(stream_), "append");
  if (first_slot) *first_slot = (uint32_t)s0;
  return SVDB_OK;
}

int Engine::overwrite(size_t logical, const double* row, size_t dim) {
  std::lock_guard<std::mutex> g(mu_);
  store_ver_++;
  if (!ok_) return SVDB_EDEVICE;
  if (hipSetDevice(device_) != hipSuccess) return SVDB_EDEVICE;
  // searches in flight were submitted before this mutation: they finish
  // against the store as it was
  if (int rc0 = complete_pending()) return rc0;
  if (logical >= nalive_) {
    set_error(SVDB_EINVAL, "index out of range");
    return SVDB_EINVAL;
  }
  if (dim != d_) {
    set_error(SVDB_EINVAL, "vector dimension differs from the store's");
    return SVDB_EINVAL;
  }
  const uint32_t s = select_slot(logical);
  // a rewritten row may leave (or join) the degenerate list
  degen_dirty_ = lowp_ || lowp8_;
  HIPCHK(hipMemcpyAsync(X_ + (size_t)s * ld_, row, d_ * sizeof(double), hipMemcpyHostToDevice,
                        stream_), "row upload");
  HIPCHK(launch_row_norms(X_, ld_, nullptr, s, 1, (uint32_t)d_, nx2_, xn_, stream_), "k_row_norms");
  int rc = shadow_rows(s, 1, nullptr);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(stream_), "overwrite");
  return SVDB_OK;
}

int Engine::erase(size_t logical) {
  std::lock_guard<std::mutex> g(mu_);
  store_ver_++;
  if (!ok_) return SVDB_EDEVICE;
  if (hipSetDevice(device_) != hipSuccess) return SVDB_EDEVICE;
  // searches in flight were submitted before this mutation: they finish
  // against the store as it was
  if (int rc0 = complete_pending()) return rc0;
  if (logical >= nalive_) {
    set_error(SVDB_EINVAL, "index out of range");
    return SVDB_EINVAL;
  }
  const uint32_t s = select_slot(logical);
  const size_t w = s >> 6, nw = (nphys_ + 63) / 64;
  // the first pass must never pick a tombstone
  if (lowp_) {
    HIPCHK(hipMemsetAsync(sh_.Xb + (size_t)s * sh_.dp, 0, sh_.dp * sizeof(uint16_t), stream_), "memset");
    HIPCHK(launch_shadow_init(sh_, s, 1, stream_), "k_shadow_init");
  }
  if (lowp8_) {
    HIPCHK(hipMemsetAsync(sh8_.X8 + (size_t)s * sh8_.dp, 0, sh8_.dp, stream_), "memset");
    HIPCHK(launch_shadow_init8(sh8_, s, 1, stream_), "k_shadow_init8");
  }
  if (lowp_ || lowp8_) HIPCHK(hipStreamSynchronize(stream_), "erase");
  h_alive_[w] &= ~(1ull << (s & 63));
  for (size_t v = w + 1; v < nw; v++) h_rank_[v]--;
  nalive_--;
  has_deletes_ = true;
  dirty_word_ = std::min(dirty_word_, w);
  degen_dirty_ = lowp_ || lowp8_;
  // Tombstones past a quarter of the slots (and at least kCompactMin):
  // squeeze them out.  Compaction is an optimisation — if it cannot get the
  // device memory it leaves the store as it is.
  const size_t dead = nphys_ - nalive_;
  if (dead >= kCompactMin && 4 * dead > nphys_) {
    // the delete is done whatever happens here
    int rc;
    try {
      rc = compact();
    } catch (const std::bad_alloc&) {
      set_error(SVDB_ENOMEM, "host allocation failed");
      rc = SVDB_ENOMEM;
    }
    if (rc != SVDB_OK) fprintf(stderr, "svdb: compaction skipped: %s\n", last_error());
  }
  return SVDB_OK;
}

// Moves the live rows to slots 0..nalive-1 in order (slot order is logical
// order, so this is a stable gather), recomputes their norms, and rebuilds
// the shadows from the compacted master: requantized rows, and the
// degenerate list and ρ from the live rows only (both otherwise only grow).
int Engine::compact() {
  if (nalive_ == nphys_) return SVDB_OK;
  fault_point();
  std::vector<uint32_t> keep;
  keep.reserve(nalive_);
  for (size_t s = 0; s < nphys_; s++)
    if ((h_alive_[s >> 6] >> (s & 63)) & 1ull) keep.push_back((uint32_t)s);
  const size_t xb = cap_ * ld_ * sizeof(double), nb = cap_ * sizeof(double);
  double *X = nullptr, *nx2 = nullptr, *xn = nullptr;
  uint32_t* didx = nullptr;
  const size_t ib = std::max<size_t>(keep.size(), 1) * sizeof(uint32_t);
  if (dev_alloc((void**)&X, xb) || dev_alloc((void**)&nx2, nb) || dev_alloc((void**)&xn, nb) ||
      dev_alloc((void**)&didx, ib)) {
    dev_free(X, xb);
    dev_free(nx2, nb);
    dev_free(xn, nb);
    (void)hipGetLastError();
    return SVDB_OK;  // not enough device memory: keep the tombstones
  }
  const uint32_t n = (uint32_t)keep.size();
  int rc = SVDB_OK;
  auto fail = [&](const char* what) {
    set_error(SVDB_EDEVICE, std::string("compaction: ") + what);
    rc = SVDB_EDEVICE;
  };
  if (hipMemcpyAsync(didx, keep.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, stream_) !=
          hipSuccess ||
      launch_gather_rows(X_, ld_, didx, n, (uint32_t)ld_, X, ld_, stream_) != hipSuccess ||
      launch_row_norms(X, ld_, nullptr, 0, n, (uint32_t)d_, nx2, xn, stream_) != hipSuccess ||
      hipStreamSynchronize(stream_) != hipSuccess)
    fail("row gather");
  dev_free(didx, ib);
  if (rc) {
    dev_free(X, xb);
    dev_free(nx2, nb);
    dev_free(xn, nb);
    return rc;
  }
  dev_free(X_, xb);
  dev_free(nx2_, nb);
  dev_free(xn_, nb);
  X_ = X;
  nx2_ = nx2;
  xn_ = xn;
  // host bitmap: slots 0..n-1 alive
  const size_t ow = (nphys_ + 63) / 64;
  for (size_t w = 0; w < ow; w++) {
    const size_t lo = w * 64;
    h_alive_[w] = lo + 64 <= n ? ~0ull : lo >= n ? 0ull : ((1ull << (n - lo)) - 1ull);
    h_rank_[w] = (uint32_t)std::min<size_t>(lo, n);
  }
  nphys_ = n;
  has_deletes_ = false;
  dirty_word_ = 0;
  // shadows: every slot dead and zeroed, ρ and the degenerate list reset,
  // then the live rows requantized.  Should a step fail, the shadows are
  // switched off (a half-rebuilt shadow could hide rows from the first
  // pass) and searches use the FP64 scan.
  bool ok = true;
  if (lowp_)
    ok = ok && hipMemsetAsync(sh_.Xb, 0, cap_ * sh_.dp * sizeof(uint16_t), stream_) == hipSuccess &&
         hipMemsetAsync(sh_.rho_bits, 0, 3 * sizeof(uint32_t), stream_) == hipSuccess &&
         launch_shadow_init(sh_, 0, (uint32_t)cap_, stream_) == hipSuccess;
  if (lowp8_)
    ok = ok && hipMemsetAsync(sh8_.X8, 0, cap_ * sh8_.dp, stream_) == hipSuccess &&
         hipMemsetAsync(sh8_.rho_bits, 0, 3 * sizeof(uint32_t), stream_) == hipSuccess &&
         launch_shadow_init8(sh8_, 0, (uint32_t)cap_, stream_) == hipSuccess;
  ok = ok && shadow_rows(0, n, nullptr) == SVDB_OK &&
       hipStreamSynchronize(stream_) == hipSuccess;
  if (!ok) {
    lowp_ = lowp8_ = false;
    lowp_wanted_ = false;
    fprintf(stderr, "svdb: shadow rebuild after compaction failed; FP64 scan only\n");
  }
  degen_dirty_ = false;
  st_.compactions++;
  return SVDB_OK;
}

// ------------------------------------------------------------------ profiling

hipEvent_t Engine::take_event() {
  if (!ev_pool_.empty()) {
    hipEvent_t e = ev_pool_.back();
    ev_pool_.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

void Engine::prof_begin(int which) {
  if (!prof_ || prof_mute_) return;
  ProfSlot ps{which, take_event(), take_event()};
  (void)hipEventRecord(ps.a, c_->stream);
  c_->prof_pending_.push_back(ps);
}

void Engine::prof_end(int which) {
  if (!prof_ || prof_mute_) return;
  for (auto it = c_->prof_pending_.rbegin(); it != c_->prof_pending_.rend(); ++it)
    if (it->kind == which) {
      (void)hipEventRecord(it->b, c_->stream);
      return;
    }
}

void Engine::prof_collect() {
  for (auto& ps : c_->prof_pending_) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ps.a, ps.b) == hipSuccess) {
      kt_[ps.kind].launches++;
      kt_[ps.kind].total_ms += ms;
    }
    ev_pool_.push_back(ps.a);
    ev_pool_.push_back(ps.b);
  }
  c_->prof_pending_.clear();
}

int Engine::set_profiling(bool on) {
  std::lock_guard<std::mutex> g(mu_);
  prof_ = on;
  for (auto& k : kt_) k = KernelTime{};
  return SVDB_OK;
}

void Engine::kernel_times(KernelTime* out) {
  std::lock_guard<std::mutex> g(mu_);
  for (int i = 0; i < SVDB_K_COUNT; i++) out[i] = kt_[i];
}

EngineStats Engine::stats() {
  std::lock_guard<std::mutex> g(mu_);
  return st_;
}

int Engine::set_budget(size_t bytes) {
  std::lock_guard<std::mutex> g(mu_);
  budget_ = bytes ? bytes : SIZE_MAX;
  return SVDB_OK;
}

int Engine::set_shortlist(size_t L) {
  std::lock_guard<std::mutex> g(mu_);
  store_ver_++;
  if (L < 1 || L > SVDB_MAX_K * 2) {
    set_error(SVDB_EINVAL, "shortlist width out of range");
    return SVDB_EINVAL;
  }
  shortlist_ = L;
  return SVDB_OK;
}

// ------------------------------------------------------------------ search

// Queries per launch of the scan: bounds the partial-list scratch.
static constexpr size_t kMaxBatch = 8192;
// Queries per launch of the low-precision pipeline.
static constexpr size_t kMaxBatchLowp = 16384;

// Row slices of an exact scan of nq queries (a multiple of 8, one set per
// XCD; 0 for an empty store), with at most max_entries partial-list entries
// (nq × slices × waves × kmax); 0 with rows present when even 8 slices
// would need more.
uint32_t Engine::exact_slices(uint32_t nq, int kmax, size_t max_entries) const {
  if (nphys_ == 0) return 0;
  // About 8,192 blocks (query groups × row slices): short slices keep the
  // query groups of one slice, which run back to back on one XCD, close
  // enough together for that XCD's L2 to serve them all the slice's rows
  // (1M × 128, 1,024 queries: 15.6 → 12.7 ms Euclidean, 12.0 → 9.9 ms dot
  // against 1,024 blocks; 16,384 lost again in the merge).  At most 1,024
  // slices, so a few queries do not leave the merge thousands of lists.
  const size_t groups = (nq + kScanTQ - 1) / kScanTQ;
  size_t want = std::min<size_t>(1024, (8192 + groups - 1) / groups);
  size_t maxs = std::max<size_t>(1, nphys_ / 1024);
  size_t nslices = std::max<size_t>(1, std::min(want, maxs));
  // large batches at large k take fewer, longer slices
  while (nslices > 8 && (size_t)nq * nslices * kWavesPerBlock * kmax > max_entries)
    nslices = (nslices + 1) / 2;
  nslices = round_up(nslices, 8);
  if ((size_t)nq * nslices * kWavesPerBlock * kmax > max_entries) return 0;
  return (uint32_t)nslices;
}

int Engine::run_exact(const double* dQ, size_t ldq, uint32_t nq, uint32_t D, const int* dk,
                      int metric, int kmax, uint64_t* dids, double* dscores, uint32_t ostride) {
  int rc;
  if (nq == 0) return SVDB_OK;
  // partial lists stay under 2^25 entries (~400 MB)
  const uint32_t nslices = exact_slices(nq, kmax, std::max<size_t>((size_t)1 << 25,
                                                                    (size_t)nq * 32 * kmax));
  const size_t nlist = (size_t)nq * nslices * kWavesPerBlock * kmax;
  if ((rc = scratch(&c_->s_pk_, &c_->s_pk_n_, std::max<size_t>(nlist, 1) * sizeof(uint64_t)))) return rc;
  if ((rc = scratch(&c_->s_ps_, &c_->s_ps_n_, std::max<size_t>(nlist, 1) * sizeof(uint32_t)))) return rc;
  if ((rc = scratch(&c_->s_qn_, &c_->s_qn_n_, (size_t)nq * 3 * sizeof(double)))) return rc;
  return run_exact_with(dQ, ldq, nq, D, dk, metric, kmax, dids, dscores, ostride, nslices,
                        (uint64_t*)c_->s_pk_, (uint32_t*)c_->s_ps_, (double*)c_->s_qn_);
}

// The exact scan and merge into caller-provided scratch: partial lists
// pk / ps of nq × nslices × kWavesPerBlock × kmax entries, and 3·nq
// doubles at qn: the query norms, their squares as fma chains, and the
// per-query global bounds of the scan (ScanF64Args.qq, .gtk).  Allocates
// nothing.
int Engine::run_exact_with(const double* dQ, size_t ldq, uint32_t nq, uint32_t D, const int* dk,
                           int metric, int kmax, uint64_t* dids, double* dscores,
                           uint32_t ostride, uint32_t nslices, uint64_t* pk, uint32_t* ps,
                           double* qn) {
  int rc;
  if (nq == 0) return SVDB_OK;
  // the squared row norms cover every column: usable when the scan does
  StoreView sv{X_, ld_, xn_, alive_, rank_, (uint32_t)nphys_, has_deletes_ ? 1 : 0,
               D == d_ ? nx2_ : nullptr};
  const uint32_t rps = nslices ? (uint32_t)((nphys_ + nslices - 1) / nslices) : 0;
  const uint32_t P = scan_f64_lists(metric, D, sv, kmax, nq, nslices);
  const bool mfma = scan_f64_uses_mfma(metric, D, sv, kmax, nq);
  double* qq = qn + nq;
  uint64_t* gtk = reinterpret_cast<uint64_t*>(qn + 2 * (size_t)nq);
  if ((rc = sync_alive())) return rc;
  if (metric == kCos || (metric == kL2 && mfma))
    HIPCHK(launch_query_norms(dQ, ldq, nq, D, qn, metric == kL2 ? qq : nullptr, c_->stream),
           "k_query_norms");
  if (nslices) {
    HIPCHK(launch_fill_i32(reinterpret_cast<int*>(gtk), 2 * nq, -1, c_->stream), "k_fill_i32");
    ScanF64Args a{sv, dQ, ldq, qn, dk, nq, D, rps, P, (uint32_t)kmax, pk, ps, gtk, qq};
    prof_begin(SVDB_K_SCAN_FP64);
    HIPCHK(launch_scan_f64(a, metric, kmax, nslices, c_->stream), "k_scan_f64");
    prof_end(SVDB_K_SCAN_FP64);
  }
  MergeArgs m{sv, dk, nq, P, (uint32_t)kmax, pk, ps, ostride, dids, dscores};
  prof_begin(SVDB_K_MERGE);
  HIPCHK(launch_merge(m, metric, kmax, c_->stream), "k_merge");
  prof_end(SVDB_K_MERGE);
  return SVDB_OK;
}

// Low-precision first pass + shortlist + FP64 rerank for nq device queries
// (all columns, one metric and k).  Queries the rerank cannot certify are
// gathered and answered by the FP64 scan.
// Shortlist: the configured width (64 by default) for small k; for BF16
// 256 once k passes 32, where the k-th and L-th neighbours need more room
// than the first-pass error bound leaves in a 2k-wide list.  FP8 starts at
// the configured width (or k) and relies on the wider retries.
uint32_t Engine::lowp_width(int kk, bool fp8, size_t shortlist) {
  return (uint32_t)std::min<size_t>(
      kMaxShort, std::max<size_t>({shortlist, (size_t)kk, (!fp8 && kk > 32) ? (size_t)256 : 0}));
}

int Engine::run_lowp(const double* dQ, size_t ldq, uint32_t nq, int metric, int k,
                     uint64_t* dids, double* dscores, uint32_t ostride, bool fp8) {
  LowpRange r{metric, 0, nq, k, 0, 0, 0, 0, 0};
  return run_lowp_ranges(dQ, ldq, &r, 1, k, dids, dscores, ostride, fp8);
}

// Low-precision search of nr metric ranges: range i holds R[i].n queries of
// metric R[i].metric at query rows R[i].qoff.. (dQ, dids, dscores).  Several
// ranges (BF16 rows <= 256 wide only) each take a run of slots starting on a
// wavefront boundary (padding slots are flagged and never scanned).  The
// first pass runs one single-metric launch per metric, or one launch of the
// mixed-metric kernel when all ranges fit one query block; the pre-pass runs
// per range; the shortlists and the reranks of all ranges share a launch;
// widening and fallback run per range, each with its own k and shortlist
// width (k is the largest over ranges).
int Engine::lowp_enqueue(const double* dQ, size_t ldq, LowpRange* R, int nr, int k,
                         uint64_t* dids, double* dscores, uint32_t ostride, bool fp8) {
  int rc;
  if (nr < 1 || nr > 3) {
    set_error(SVDB_EINVAL, "one to three query ranges");
    return SVDB_EINVAL;
  }
  const size_t shortlist = shortlist_;
  const bool mixed = nr > 1;  // several ranges: slots aligned per range
  bool multi_metric = false;  // ... of different metrics: the mixed-metric scan
  for (int i = 1; i < nr; i++) multi_metric |= R[i].metric != R[0].metric;
  if (multi_metric && (fp8 || sh_.dp > 256)) {
    set_error(SVDB_EINVAL, "mixed-metric first pass needs the BF16 shadow of rows <= 256 wide");
    return SVDB_EINVAL;
  }
  // (the FP8 pass hands leftovers to a nested BF16 pass, which reuses the
  // candidate and fail-list buffers: one range only)
  if (fp8 && nr > 1) {
    set_error(SVDB_EINVAL, "FP8 first pass takes one query range");
    return SVDB_EINVAL;
  }
  uint32_t nq = 0;  // slots
  for (int i = 0; i < nr; i++) {
    R[i].soff = nq;
    nq += mixed ? (uint32_t)round_up(R[i].n, kSlotAlign) : R[i].n;
    if (R[i].k <= 0 || R[i].k > k) R[i].k = k;
  }
  const size_t dp = fp8 ? sh8_.dp : sh_.dp;
  const size_t qbytes = fp8 ? 1 : sizeof(uint16_t);  // bytes per query element
  const uint32_t qpb = (uint32_t)(fp8 ? fp8_queries_per_block(dp) : lowp_queries_per_block(dp));
  const uint32_t nq_pad = (uint32_t)round_up(nq, qpb);
  auto width = [&](int kk) { return lowp_width(kk, fp8, shortlist); };
  const uint32_t L = width(k);  // the widest range's
  const uint32_t nrows_pad = (uint32_t)round_up(nphys_, kLowpTile);
  const uint32_t groups = nq_pad / qpb;
  const uint32_t tiles = nrows_pad / kLowpTile;
  // Row slices for a scan of grp query blocks wanting kk neighbours: at
  // least 8 (each query then has >= 32 candidate lists, so the lists' own
  // bound sits well past the shortlist's), more when there are few query
  // blocks, each at least 4 tiles long.  One block per CU fits at a time,
  // so the grid is kept at or just under two rounds of the 256 CUs (a few
  // blocks past a round cost a whole round).
  auto slicing = [&](uint32_t grp, int kk, LowpRange& r) {
    uint32_t ns = std::max<uint32_t>(8, (512 / grp) & ~7u);
    // Past k = 32 the lists' own bound (the smallest 8th entry) must stay
    // well beyond the k-th neighbour: at least 512 lists per query.
    if (kk > 32) ns = std::max<uint32_t>(ns, 128);
    // FP8: the widening retries take up to kMaxShort candidates from these
    // lists; 2048 lists per query keep their own bound (the smallest 8th
    // entry) clear of the 4096th candidate.
    if (fp8) ns = std::max<uint32_t>(ns, 512);
    ns = std::max<uint32_t>(1, std::min<uint32_t>(ns, std::max<uint32_t>(1, tiles / 4)));
    r.rps = (uint32_t)((tiles + ns - 1) / ns) * kLowpTile;
    ns = (nrows_pad + r.rps - 1) / r.rps;
    // one set of slices per XCD: a multiple of 8 (trailing slices may be empty)
    r.ns = (uint32_t)round_up(ns, 8);
    r.NL = r.ns * kLowpLanesPerQuery;
  };
  // Scan launches: consecutive ranges of one metric share one (their slots
  // are contiguous; the padding slots between them are flagged), sliced for
  // its own query count.  When the metric ranges fit one block of queries,
  // one launch of the mixed-metric kernel serves them all instead (one pass
  // over the rows rather than one per metric); past that, the single-metric
  // kernels win: they skip most subtiles with one test, which the mixed
  // kernel has no registers for.
  const bool mixed_scan = multi_metric && groups <= 1;
  struct ScanGroup {
    int r0, r1, metric;
    uint32_t soff, n;  // slots soff .. soff+n-1
  };
  std::vector<ScanGroup> sgs;
  for (int i = 0; i < nr; i++) {
    if (!sgs.empty() && (mixed_scan || R[i].metric == sgs.back().metric))
      sgs.back().r1 = i + 1;
    else
      sgs.push_back({i, i + 1, mixed_scan ? kMixed : R[i].metric, R[i].soff, 0});
  }
  size_t cand_end = 0;  // candidate-list entries, all groups
  for (auto& g : sgs) {
    const LowpRange& last = R[g.r1 - 1];
    g.n = last.soff + last.n - g.soff;
    const uint32_t send = g.r1 < nr ? R[g.r1].soff : nq;
    int kg = 0;
    for (int i = g.r0; i < g.r1; i++) kg = std::max(kg, R[i].k);
    LowpRange sl{};
    slicing((g.n + qpb - 1) / qpb, kg, sl);
    for (int i = g.r0; i < g.r1; i++) {
      R[i].ns = sl.ns;
      R[i].rps = sl.rps;
      R[i].NL = sl.NL;
      R[i].coff = cand_end + (size_t)(R[i].soff - g.soff) * sl.NL * kLowpLP;
    }
    cand_end += (size_t)(send - g.soff) * sl.NL * kLowpLP;
  }
  // Sample pre-pass: the same scan over ns_pre evenly spaced runs of tp
  // tiles (at most 1/16 of the store), each lane keeping only the smallest
  // key of its rows.  The lanes' rows are disjoint, so the L-th smallest of
  // those minima is the ordering key of a real row with L-1 others at or
  // below it: the global L-th key lies at or below it, and the main scan's
  // lists start from it.  They then skip nearly every row from the first
  // tile on, where otherwise a short slice spends most of its subtiles
  // filling its lists (the insertion path costs several times the MFMAs).
  const uint32_t ns_pre = tiles >= 16 * kPrepassSlices ? kPrepassSlices : kPrepassSlices / 4;
  const bool pre = prepass_ && tiles >= 16 * ns_pre && ns_pre * kLowpLanesPerQuery >= L;
  // BF16: twice the sample for shortlists past 64 (their threshold sits
  // further out, and a sparser sample leaves more rows below it to insert)
  // and for batches of up to 4096 queries (the scan's saving outweighs the
  // larger pre-pass there; at 16,384 queries and L = 64 they break even).
  // Not FP8: a list the threshold leaves short bounds β by the threshold,
  // and FP8's wider error band then certifies fewer queries (measured on
  // 5M x 768 cosine: 861 instead of 4,293 of 12,288 before the BF16 pass).
  const uint32_t tpmax =
      (!fp8 && (L > 64 || nq <= 4096)) ? 2 * kPrepassMaxTiles : kPrepassMaxTiles;
  const uint32_t tp = pre ? std::min<uint32_t>(tpmax, tiles / (16 * ns_pre)) : 0;
  const uint32_t NL_pre = ns_pre * kLowpLanesPerQuery;
  const size_t ncand = std::max(cand_end, pre ? (size_t)nq * NL_pre : (size_t)0);
  // query rows: nq_pad slots plus one block, since a launch over one range
  // (at slot offset soff) reads whole blocks of rows from soff on
  if ((rc = scratch(&c_->s_qb_, &c_->s_qb_n_, ((size_t)nq_pad + qpb) * dp * qbytes))) return rc;
  if ((rc = scratch(&c_->s_qp_, &c_->s_qp_n_, (size_t)nq_pad * sizeof(QParams)))) return rc;
  if ((rc = scratch(&c_->s_ck_, &c_->s_ck_n_, ncand * sizeof(float)))) return rc;
  if ((rc = scratch(&c_->s_cs_, &c_->s_cs_n_, ncand * sizeof(uint32_t)))) return rc;
  if ((rc = scratch(&c_->s_sk_, &c_->s_sk_n_, (size_t)nq * L * sizeof(float)))) return rc;
  if ((rc = scratch(&c_->s_ss_, &c_->s_ss_n_, (size_t)nq * L * sizeof(uint32_t)))) return rc;
  if ((rc = scratch(&c_->s_beta_, &c_->s_beta_n_, (size_t)nq * sizeof(float)))) return rc;
  // fail counts (one per range) then each range's fail list at nr + soff
  if ((rc = scratch(&c_->s_fail_, &c_->s_fail_n_, ((size_t)nq + nr) * sizeof(uint32_t)))) return rc;
  if ((rc = scratch(&c_->s_err_, &c_->s_err_n_, (size_t)nq * sizeof(double)))) return rc;
  // (a range's lists sit at soff * k, k the largest over ranges)
  if ((rc = scratch(&c_->s_lk_, &c_->s_lk_n_, (size_t)nq * k * sizeof(uint64_t)))) return rc;
  if ((rc = scratch(&c_->s_ls_, &c_->s_ls_n_, (size_t)nq * k * sizeof(uint32_t)))) return rc;
  if ((rc = scratch(&c_->s_km_, &c_->s_km_n_, (size_t)nq * sizeof(float)))) return rc;
  if (!c_->h_fail_) HIPCHK(hipHostMalloc((void**)&c_->h_fail_, sizeof(uint32_t)), "hipHostMalloc");
  const size_t io_b = round_up(nr * sizeof(uint32_t), 64) + (size_t)nq * sizeof(double);
  if (c_->h_io_n_ < io_b) {
    if (c_->h_io_) (void)hipHostFree(c_->h_io_);
    c_->h_io_ = nullptr;
    c_->h_io_n_ = 0;
    HIPCHK(hipHostMalloc(&c_->h_io_, round_up(io_b, 4096)), "hipHostMalloc");
    c_->h_io_n_ = round_up(io_b, 4096);
  }
  uint32_t* h_counts = static_cast<uint32_t*>(c_->h_io_);
  double* h_err = reinterpret_cast<double*>(static_cast<char*>(c_->h_io_) + round_up(nr * sizeof(uint32_t), 64));
  if ((rc = sync_alive())) return rc;
  if ((rc = refresh_degen())) return rc;

  HIPCHK(hipMemsetAsync(c_->s_err_, 0, (size_t)nq * sizeof(double), c_->stream), "memset");
  // query prep per range; each range's slots run to the next range (the last
  // to nq_pad), the slots past its queries flagged as padding
  for (int i = 0; i < nr; i++) {
    const uint32_t send = i + 1 < nr ? R[i + 1].soff : nq_pad;
    const double* q = dQ + (size_t)R[i].qoff * ldq;
    unsigned char* qb = (unsigned char*)c_->s_qb_ + (size_t)R[i].soff * dp * qbytes;
    QParams* qp = (QParams*)c_->s_qp_ + R[i].soff;
    if (fp8)
      HIPCHK(launch_query_prep8(q, ldq, R[i].n, send - R[i].soff, (uint32_t)d_, dp, R[i].metric,
                                sh8_.rho_bits, (uint8_t*)qb, qp, c_->stream), "k_query_prep8");
    else
      HIPCHK(launch_query_prep(q, ldq, R[i].n, send - R[i].soff, (uint32_t)d_, dp, R[i].metric,
                               sh_.rho_bits, (uint16_t*)qb, qp, c_->stream), "k_query_prep");
  }
  // per-row ordering-key arrays: BF16 L2 (nx, pen), dot (xn, pen), cosine
  // (inv, pen); FP8 L2 (nx, sx), dot (xn, sx), cosine (sx/‖x‖, pen)
  auto w0_of = [&](int m) {
    return fp8 ? (m == kL2 ? sh8_.nx : m == kDot ? sh8_.xn : sh8_.inv)
               : (m == kL2 ? sh_.nx : m == kDot ? sh_.xn : sh_.inv);
  };
  auto w1_of = [&](int m) { return fp8 ? (m == kCos ? sh8_.pen : sh8_.sx) : sh_.pen; };
  const uint32_t* degen = fp8 ? sh8_.degen : sh_.degen;
  const uint32_t* degen_count = fp8 ? sh8_.degen_count : sh_.degen_count;
  auto scan = [&](const LowpScanArgs& x, int m, uint32_t nsl, bool minonly) {
    return fp8 ? launch_scan_fp8(x, m, nsl, minonly, c_->stream)
               : launch_scan_bf16(x, m, nsl, minonly, c_->stream);
  };
  // the fields every launch shares; each sets its own rows, queries and lists
  LowpScanArgs sa{fp8 ? (const uint16_t*)sh8_.X8 : sh_.Xb, dp, nullptr, nullptr, nullptr,
                  nullptr, (const uint16_t*)c_->s_qb_, (const QParams*)c_->s_qp_, nq, 0, 0,
                  nrows_pad, 0, (float*)c_->s_ck_, (uint32_t*)c_->s_cs_, nullptr, 0};
  if (pre) {
    prof_begin(SVDB_K_PREPASS);
    for (int i = 0; i < nr; i++) {
      LowpScanArgs pa = sa;
      pa.w0 = w0_of(R[i].metric);
      pa.w1 = w1_of(R[i].metric);
      pa.w2 = pa.w3 = nullptr;
      pa.Qb = (const uint16_t*)((const unsigned char*)c_->s_qb_ + (size_t)R[i].soff * dp * qbytes);
      pa.qp = (const QParams*)c_->s_qp_ + R[i].soff;
      pa.nq = R[i].n;
      pa.rows_per_slice = tp * kLowpTile;
      pa.slice_stride = (tiles / ns_pre) * kLowpTile;
      pa.NL = NL_pre;
      pa.cand_key = (float*)c_->s_ck_ + (size_t)R[i].soff * NL_pre;
      HIPCHK(scan(pa, R[i].metric, ns_pre, true), "first-pass scan (sample)");
      // thresholds in c_->s_beta_, read by the main scan before the shortlist
      // below overwrites them (padding slots' values are never used)
      HIPCHK(launch_kth_key((const float*)c_->s_ck_ + (size_t)R[i].soff * NL_pre, R[i].n, NL_pre,
                            width(R[i].k), (float*)c_->s_beta_ + R[i].soff, c_->stream), "k_kth_key");
    }
    prof_end(SVDB_K_PREPASS);
    sa.tau0 = (const float*)c_->s_beta_;
    sa.tau0_stride = 1;
  }
  prof_begin(SVDB_K_SCAN_LOWP);
  for (const ScanGroup& g : sgs) {
    if (g.n == 0) continue;
    const LowpRange& r = R[g.r0];
    const bool mk = g.metric == kMixed;
    LowpScanArgs ma = sa;
    ma.w0 = mk ? sh_.nx : w0_of(g.metric);
    ma.w1 = mk ? sh_.pen : w1_of(g.metric);
    ma.w2 = mk ? sh_.xn : nullptr;
    ma.w3 = mk ? sh_.inv : nullptr;
    ma.Qb = (const uint16_t*)((const unsigned char*)c_->s_qb_ + (size_t)g.soff * dp * qbytes);
    ma.qp = (const QParams*)c_->s_qp_ + g.soff;
    ma.nq = g.n;
    ma.rows_per_slice = ma.slice_stride = r.rps;
    ma.NL = r.NL;
    ma.cand_key = (float*)c_->s_ck_ + r.coff;
    ma.cand_slot = (uint32_t*)c_->s_cs_ + r.coff;
    if (ma.tau0) ma.tau0 += g.soff;
    HIPCHK(scan(ma, g.metric, r.ns, false), "first-pass scan");
  }
  prof_end(SVDB_K_SCAN_LOWP);
  // range i's shortlist: width Li = width(k_i), at slot offset soff (row
  // stride Li inside the range's block of soff*L .. )
  // every range's shortlist, in as few launches as the kernels allow
  std::vector<ShortlistArgs> sls(nr);
  for (int i = 0; i < nr; i++)
    sls[i] = ShortlistArgs{(const float*)c_->s_ck_ + R[i].coff, (const uint32_t*)c_->s_cs_ + R[i].coff,
                           R[i].n, R[i].NL, width(R[i].k),
                           (uint32_t*)c_->s_ss_ + (size_t)R[i].soff * L,
                           (float*)c_->s_sk_ + (size_t)R[i].soff * L, (float*)c_->s_beta_ + R[i].soff};
  prof_begin(SVDB_K_SHORTLIST);
  HIPCHK(launch_shortlist_ranges(sls.data(), nr, c_->stream), "k_shortlist");
  prof_end(SVDB_K_SHORTLIST);
  StoreView sv{X_, ld_, xn_, alive_, rank_, (uint32_t)nphys_, has_deletes_ ? 1 : 0, nullptr};
  const uint32_t Lmax = L;  // shortlist block stride per slot
  // FP64 rerank of every range's shortlist; fail counts and the first-pass
  // error report come back together after one wait
  HIPCHK(hipMemsetAsync(c_->s_fail_, 0, nr * sizeof(uint32_t), c_->stream), "memset");
  auto rerank_args = [&](int ri) {
    const LowpRange& rg = R[ri];
    const uint32_t* fl = (uint32_t*)c_->s_fail_ + nr + rg.soff;
    return RerankArgs{sv, dQ + (size_t)rg.qoff * ldq, ldq, (uint32_t)d_,
                      (const QParams*)c_->s_qp_ + rg.soff,
                      (const uint32_t*)c_->s_ss_ + (size_t)rg.soff * Lmax,
                      (const float*)c_->s_sk_ + (size_t)rg.soff * Lmax,
                      (const float*)c_->s_beta_ + rg.soff, degen, degen_count, w0_of(rg.metric),
                      w1_of(rg.metric), rg.n, width(rg.k), (uint32_t)rg.k, ostride,
                      dids + (size_t)rg.qoff * ostride, dscores + (size_t)rg.qoff * ostride,
                      (uint32_t*)c_->s_fail_ + ri, const_cast<uint32_t*>(fl),
                      (double*)c_->s_err_ + rg.soff, nullptr,
                      (uint64_t*)c_->s_lk_ + (size_t)rg.soff * k,
                      (uint32_t*)c_->s_ls_ + (size_t)rg.soff * k, (float*)c_->s_km_ + rg.soff, 0};
  };
  std::vector<RerankArgs> ras(nr);
  std::vector<int> rms(nr);
  for (int ri = 0; ri < nr; ri++) {
    ras[ri] = rerank_args(ri);
    rms[ri] = R[ri].metric;
  }
  prof_begin(SVDB_K_RERANK);
  HIPCHK(launch_rerank_ranges(ras.data(), rms.data(), nr, c_->stream), "k_rerank");
  prof_end(SVDB_K_RERANK);
  HIPCHK(hipMemcpyAsync(h_counts, c_->s_fail_, nr * sizeof(uint32_t), hipMemcpyDeviceToHost, c_->stream),
         "copy");
  HIPCHK(hipMemcpyAsync(h_err, c_->s_err_, (size_t)nq * sizeof(double), hipMemcpyDeviceToHost, c_->stream),
         "copy");
  // what lowp_complete needs once the counts are back
  SearchCtx::LowpPending& lp = c_->lp;
  lp.on = true;
  for (int i = 0; i < nr; i++) lp.R[i] = R[i];
  lp.nr = nr;
  lp.k = k;
  lp.dQ = dQ;
  lp.ldq = ldq;
  lp.dids = dids;
  lp.dscores = dscores;
  lp.ostride = ostride;
  lp.fp8 = fp8;
  lp.nq = nq;
  lp.shortlist = shortlist;
  lp.failed = false;
  return SVDB_OK;
}

// Waits for lowp_enqueue's work and answers what the first pass could not
// certify: the widening retries, the FP8 queries' BF16 second pass, and the
// FP64 rescan.
int Engine::lowp_complete() {
  int rc;
  SearchCtx::LowpPending lp = c_->lp;  // a nested pass below reuses c_->lp
  c_->lp.on = false;
  if (!lp.on) return SVDB_OK;
  LowpRange* R = lp.R;
  const int nr = lp.nr;
  const double* dQ = lp.dQ;
  const size_t ldq = lp.ldq;
  uint64_t* dids = lp.dids;
  double* dscores = lp.dscores;
  const uint32_t ostride = lp.ostride;
  const bool fp8 = lp.fp8;
  const uint32_t nq = lp.nq;
  auto width = [&](int kk) { return lowp_width(kk, fp8, lp.shortlist); };
  auto w0_of = [&](int m) {
    return fp8 ? (m == kL2 ? sh8_.nx : m == kDot ? sh8_.xn : sh8_.inv)
               : (m == kL2 ? sh_.nx : m == kDot ? sh_.xn : sh_.inv);
  };
  auto w1_of = [&](int m) { return fp8 ? (m == kCos ? sh8_.pen : sh8_.sx) : sh_.pen; };
  const uint32_t* degen = fp8 ? sh8_.degen : sh_.degen;
  const uint32_t* degen_count = fp8 ? sh8_.degen_count : sh_.degen_count;
  StoreView sv{X_, ld_, xn_, alive_, rank_, (uint32_t)nphys_, has_deletes_ ? 1 : 0, nullptr};
  const uint32_t Lmax = width(lp.k);  // shortlist block stride per slot
  auto rerank_args = [&](int ri) {
    const LowpRange& rg = R[ri];
    const uint32_t* fl = (uint32_t*)c_->s_fail_ + nr + rg.soff;
    return RerankArgs{sv, dQ + (size_t)rg.qoff * ldq, ldq, (uint32_t)d_,
                      (const QParams*)c_->s_qp_ + rg.soff,
                      (const uint32_t*)c_->s_ss_ + (size_t)rg.soff * Lmax,
                      (const float*)c_->s_sk_ + (size_t)rg.soff * Lmax,
                      (const float*)c_->s_beta_ + rg.soff, degen, degen_count, w0_of(rg.metric),
                      w1_of(rg.metric), rg.n, width(rg.k), (uint32_t)rg.k, ostride,
                      dids + (size_t)rg.qoff * ostride, dscores + (size_t)rg.qoff * ostride,
                      (uint32_t*)c_->s_fail_ + ri, const_cast<uint32_t*>(fl),
                      (double*)c_->s_err_ + rg.soff, nullptr,
                      (uint64_t*)c_->s_lk_ + (size_t)rg.soff * lp.k,
                      (uint32_t*)c_->s_ls_ + (size_t)rg.soff * lp.k, (float*)c_->s_km_ + rg.soff,
                      0};
  };
  uint32_t* h_counts = static_cast<uint32_t*>(c_->h_io_);
  double* h_err = reinterpret_cast<double*>(static_cast<char*>(c_->h_io_) +
                                            round_up(nr * sizeof(uint32_t), 64));
  HIPCHK(hipStreamSynchronize(c_->stream), "lowp");
  double mx = 0.0;
  for (uint32_t i = 0; i < nq; i++) mx = std::max(mx, h_err[i]);
  std::vector<uint32_t> nfail(h_counts, h_counts + nr);
  for (int ri = 0; ri < nr; ri++) {
    const LowpRange& rg = R[ri];
    const int metric = rg.metric;
    const uint32_t nqr = rg.n;
    const int k = rg.k;                // this range's k (shadows the call's)
    const uint32_t L = width(k);       // and its shortlist width
    const uint32_t NL = rg.NL;         // and its candidate lists per query
    const uint32_t nf1 = nfail[ri];
    st_.certified += nqr - nf1;
    if (nf1 == 0) continue;
    c_->lp.failed = true;  // (restored below for the caller's record)
    lp.failed = true;
    // this range's views: slot-space arrays at soff, query-space at qoff
    const double* rQ = dQ + (size_t)rg.qoff * ldq;
    uint64_t* rids = dids + (size_t)rg.qoff * ostride;
    double* rsc = dscores + (size_t)rg.qoff * ostride;
    const float* rck = (const float*)c_->s_ck_ + rg.coff;
    const uint32_t* rcs = (const uint32_t*)c_->s_cs_ + rg.coff;
    const RerankArgs ra = rerank_args(ri);
    uint32_t* fail_list = ra.fail_list;  // may be redirected to a widening retry's list
    uint32_t nf = nf1;
    // Widening retries: queries not certified at L are shortlisted again at
    // 1024, then 4096, from the same candidate lists (when they hold at
    // least twice that many entries) and reranked; only those still
    // uncertified go on.
    for (int wi = 0; wi < 2; wi++) {
      const uint32_t Lw = wi == 0 ? 1024u : (uint32_t)kMaxShort;
      if (nf == 0 || Lw <= L || (size_t)NL * kLowpLP < 2 * (size_t)Lw) continue;
      // A retry that has lately certified almost none of its queries for
      // this precision and metric is skipped (its queries go on to the next
      // stage), except every 16th time, which keeps the estimate current.
      // Which stage certifies a query never changes its ids.
      float& miss = retry_miss_[fp8 ? 1 : 0][metric][wi];
      if (miss > 0.95f && (++retry_seen_[fp8 ? 1 : 0][metric][wi] % 16) != 0) continue;
      const uint32_t nf_in = nf;
      if ((rc = scratch(&c_->s_sk2_, &c_->s_sk2_n_, (size_t)nf * Lw * sizeof(float)))) return rc;
      if ((rc = scratch(&c_->s_ss2_, &c_->s_ss2_n_, (size_t)nf * Lw * sizeof(uint32_t)))) return rc;
      if ((rc = scratch(&c_->s_beta2_, &c_->s_beta2_n_, (size_t)nf * sizeof(float)))) return rc;
      if ((rc = scratch(&c_->s_fail2_, &c_->s_fail2_n_, 2 * ((size_t)nf + 1) * sizeof(uint32_t)))) return rc;
      // two fail lists alternate: this stage reads `fail_list`, writes the other
      uint32_t* out_count =
          (uint32_t*)c_->s_fail2_ + (fail_list == (uint32_t*)c_->s_fail2_ + 1 ? nf + 1 : 0);
      uint32_t* out_list = out_count + 1;
      HIPCHK(hipMemsetAsync(out_count, 0, sizeof(uint32_t), c_->stream), "memset");
      ShortlistArgs sw{rck, rcs, nf, NL, Lw, (uint32_t*)c_->s_ss2_, (float*)c_->s_sk2_,
                       (float*)c_->s_beta2_, fail_list};
      prof_begin(SVDB_K_SHORTLIST);
      HIPCHK(launch_shortlist(sw, c_->stream), "k_shortlist (widened)");
      prof_end(SVDB_K_SHORTLIST);
      RerankArgs rw = ra;
      rw.short_slot = (const uint32_t*)c_->s_ss2_;
      rw.short_key = (const float*)c_->s_sk2_;
      rw.beta = (const float*)c_->s_beta2_;
      rw.nq = nf;
      rw.L = Lw;
      rw.fail_count = out_count;
      rw.fail_list = out_list;
      rw.qidx = fail_list;
      // the list starts from the earlier stage's; its candidates are not
      // read again
      rw.carry = 1;
      prof_begin(SVDB_K_RERANK);
      HIPCHK(launch_rerank(rw, metric, c_->stream), "k_rerank (widened)");
      prof_end(SVDB_K_RERANK);
      HIPCHK(hipMemcpyAsync(c_->h_fail_, out_count, sizeof(uint32_t), hipMemcpyDeviceToHost, c_->stream),
             "copy");
      HIPCHK(hipStreamSynchronize(c_->stream), "lowp widened");
      nf = *c_->h_fail_;
      fail_list = out_list;
      // (a batch of 256 or more failing queries is a sample on its own)
      if (nf_in >= 32) {
        const float w = nf_in >= 256 ? 1.0f : 0.5f;
        miss = (1.0f - w) * miss + w * (float)nf / (float)nf_in;
      }
    }
    // the error report again, now with the widened queries certified
    if (nf != nf1) {
      double* he = reinterpret_cast<double*>(static_cast<char*>(c_->h_io_) +
                                             round_up(nr * sizeof(uint32_t), 64));
      HIPCHK(hipMemcpyAsync(he, (double*)c_->s_err_ + rg.soff, nqr * sizeof(double),
                            hipMemcpyDeviceToHost, c_->stream), "copy");
      HIPCHK(hipStreamSynchronize(c_->stream), "copy");
      for (uint32_t i = 0; i < nqr; i++) mx = std::max(mx, he[i]);
    }
    st_.certified += nf1 - nf;
    st_.widened += nf1 - nf;
    if (nf == 0) continue;
    // Second chance for FP8: queries its stages leave uncertified go through
    // the BF16 first pass when the store has a BF16 shadow (a bound about
    // 16x tighter); that pass rescans in FP64 only what it cannot certify.
    const bool bf16_ok = lowp_ && (sh_.dp <= 256 || bf16_wide_dp_supported(sh_.dp));
    if (fp8 && bf16_ok) {
      if ((rc = scratch(&c_->s_cidx_, &c_->s_cidx_n_, (size_t)nf * sizeof(uint32_t)))) return rc;
      if ((rc = scratch(&c_->s_cq_, &c_->s_cq_n_, (size_t)nf * d_ * sizeof(double)))) return rc;
      if ((rc = scratch(&c_->s_cid_, &c_->s_cid_n_, (size_t)nf * k * sizeof(uint64_t)))) return rc;
      if ((rc = scratch(&c_->s_csc_, &c_->s_csc_n_, (size_t)nf * k * sizeof(double)))) return rc;
      // the BF16 pass reuses the fail-list buffers: keep this list apart
      HIPCHK(hipMemcpyAsync(c_->s_cidx_, fail_list, nf * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                            c_->stream), "copy");
      HIPCHK(launch_gather_rows(rQ, ldq, (const uint32_t*)c_->s_cidx_, nf, (uint32_t)d_,
                                (double*)c_->s_cq_, d_, c_->stream), "k_gather_rows");
      st_.second_pass += nf;
      prof_begin(SVDB_K_SECOND_PASS);
      prof_mute_ = true;
      const double keep = mx;
      rc = run_lowp((const double*)c_->s_cq_, d_, nf, metric, k, (uint64_t*)c_->s_cid_, (double*)c_->s_csc_,
                    (uint32_t)k, false);
      prof_mute_ = false;
      if (rc) return rc;
      prof_end(SVDB_K_SECOND_PASS);
      mx = std::max(keep, st_.last_max_rel_err);
      HIPCHK(launch_scatter_results((const uint64_t*)c_->s_cid_, (const double*)c_->s_csc_, (uint32_t)k,
                                    (const uint32_t*)c_->s_cidx_, nf, rids, rsc, ostride, c_->stream),
             "k_scatter_results");
      HIPCHK(hipStreamSynchronize(c_->stream), "second pass");
      continue;
    }
    if ((rc = scratch(&c_->s_fq_, &c_->s_fq_n_, (size_t)nf * d_ * sizeof(double)))) return rc;
    if ((rc = scratch(&c_->s_fid_, &c_->s_fid_n_, (size_t)nf * k * sizeof(uint64_t)))) return rc;
    if ((rc = scratch(&c_->s_fsc_, &c_->s_fsc_n_, (size_t)nf * k * sizeof(double)))) return rc;
    if ((rc = scratch(&c_->s_fk_, &c_->s_fk_n_, (size_t)nf * sizeof(int)))) return rc;
    std::vector<int> hk(nf, k);
    HIPCHK(hipMemcpyAsync(c_->s_fk_, hk.data(), nf * sizeof(int), hipMemcpyHostToDevice, c_->stream), "copy");
    HIPCHK(launch_gather_rows(rQ, ldq, fail_list, nf, (uint32_t)d_, (double*)c_->s_fq_, d_, c_->stream),
           "k_gather_rows");
    rc = run_exact((const double*)c_->s_fq_, d_, nf, (uint32_t)d_, (const int*)c_->s_fk_, metric, k,
                   (uint64_t*)c_->s_fid_, (double*)c_->s_fsc_, (uint32_t)k);
    if (rc) return rc;
    HIPCHK(launch_scatter_results((const uint64_t*)c_->s_fid_, (const double*)c_->s_fsc_, (uint32_t)k,
                                  fail_list, nf, rids, rsc, ostride, c_->stream),
           "k_scatter_results");
    HIPCHK(hipStreamSynchronize(c_->stream), "fallback");
    st_.fallbacks += nf;
  }
  st_.last_max_rel_err = mx;
  c_->lp.failed = lp.failed;
  return SVDB_OK;
}

int Engine::run_lowp_ranges(const double* dQ, size_t ldq, LowpRange* R, int nr, int k,
                            uint64_t* dids, double* dscores, uint32_t ostride, bool fp8) {
  int rc = lowp_enqueue(dQ, ldq, R, nr, k, dids, dscores, ostride, fp8);
  if (rc) {
    c_->lp.on = false;
    return rc;
  }
  return lowp_complete();
}



int Engine::use_lowp(int precision, size_t D, int* rc) {
  *rc = SVDB_OK;
  if (precision == SVDB_PREC_FP64) return 0;
  const bool bf = lowp_ && D == d_, f8 = lowp8_ && D == d_;
  if (precision == SVDB_PREC_BF16 && !bf) {
    set_error(SVDB_EINVAL, "BF16 first pass needs the shadow (dimension <= 1536) and all columns");
    *rc = SVDB_EINVAL;
    return 0;
  }
  if (precision == SVDB_PREC_FP8 && !f8) {
    set_error(SVDB_EINVAL, "FP8 first pass needs the FP8 shadow (svdb_enable_shadow) and all columns");
    *rc = SVDB_EINVAL;
    return 0;
  }
  if (precision == SVDB_PREC_FP8) return 2;
  if (bf) return 1;
  return (precision == SVDB_PREC_AUTO && f8) ? 2 : 0;
}

static int check_search(size_t d, size_t D, int metric, int precision) {
  if (metric < 0 || metric > 2) {
    set_error(SVDB_EINVAL, "unknown metric");
    return SVDB_EINVAL;
  }
  if (precision < SVDB_PREC_FP64 || precision > SVDB_PREC_AUTO) {
    set_error(SVDB_EINVAL, "unknown precision");
    return SVDB_EINVAL;
  }
  if (D == 0 || D > d) {
    set_error(SVDB_EINVAL, "compared columns must be in [1, store dimension]");
    return SVDB_EINVAL;
  }
  if (metric != kL2 && D != d) {
    set_error(SVDB_EINVAL, "cosine and dot compare all columns");
    return SVDB_EINVAL;
  }
  return SVDB_OK;
}

int Engine::search_device(const double* dQ, size_t nq, size_t ldq, size_t D, size_t k,
                          int metric, int precision, uint64_t* dids, double* dscores) {
  std::lock_guard<std::mutex> g(mu_);
  c_ = &ctx_[0];
  if (!ok_) return SVDB_EDEVICE;
  if (hipSetDevice(device_) != hipSuccess) return SVDB_EDEVICE;
  if (d_ == 0) {
    set_error(SVDB_EEMPTY, "store has no rows");
    return SVDB_EEMPTY;
  }
  int rc = check_search(d_, D, metric, precision);
  if (rc) return rc;
  if (k < 1 || k > SVDB_MAX_K || ldq < D) {
    set_error(SVDB_EINVAL, "k must be in [1, 128] and ldq >= compared columns");
    return SVDB_EINVAL;
  }
  const size_t nb = std::min(nq, std::max(kMaxBatch, kMaxBatchLowp));
  if ((rc = scratch(&c_->s_k_, &c_->s_k_n_, nb * sizeof(int)))) return rc;
  std::vector<int> hk(nb, (int)k);
  HIPCHK(hipMemcpyAsync(c_->s_k_, hk.data(), nb * sizeof(int), hipMemcpyHostToDevice, c_->stream), "k upload");
  const int which = nphys_ > 0 ? use_lowp(precision, D, &rc) : 0;
  if (rc) return rc;
  const bool lowp = which != 0;
  const size_t step = lowp ? kMaxBatchLowp : kMaxBatch;
  for (size_t q0 = 0; q0 < nq; q0 += step) {
    const uint32_t n = (uint32_t)std::min(step, nq - q0);
    rc = lowp ? run_lowp(dQ + q0 * ldq, ldq, n, metric, (int)k, dids + q0 * k, dscores + q0 * k,
                         (uint32_t)k, which == 2)
              : run_exact(dQ + q0 * ldq, ldq, n, (uint32_t)D, (const int*)c_->s_k_, metric, (int)k,
                          dids + q0 * k, dscores + q0 * k, (uint32_t)k);
    if (rc) return rc;
  }
  HIPCHK(hipStreamSynchronize(c_->stream), "search");
  prof_collect();
  st_.searches++;
  st_.queries += nq;
  return SVDB_OK;
}

int Engine::search_device_noalloc(const double* dQ, size_t nq, size_t ldq, size_t k, int metric,
                                  uint64_t* dids, double* dscores, void* scratch, size_t bytes) {
  std::lock_guard<std::mutex> g(mu_);
  c_ = &ctx_[0];
  if (!ok_) return SVDB_EDEVICE;
  if (hipSetDevice(device_) != hipSuccess) return SVDB_EDEVICE;
  if (d_ == 0) {
    set_error_static(SVDB_EEMPTY, "store has no rows");
    return SVDB_EEMPTY;
  }
  if (metric < 0 || metric > 2 || k < 1 || k > SVDB_MAX_K || ldq < d_) {
    set_error_static(SVDB_EINVAL, "bad argument: metric, k in [1, 128], or ldq below the "
                                  "store's dimension");
    return SVDB_EINVAL;
  }
  if (nq == 0) return SVDB_OK;
  // scratch: k per query | query norms, squared norms, global bounds |
  // partial keys | partial slots
  char* base = static_cast<char*>(scratch);
  const size_t kb = round_up(nq * sizeof(int), 256), qb = round_up(nq * 3 * sizeof(double), 256);
  if (nq > kMaxBatch || kb + qb >= bytes) {
    set_error_static(SVDB_EINVAL, "batch too large for the allocation-free search");
    return SVDB_EINVAL;
  }
  const size_t entries = (bytes - kb - qb - 256) / (sizeof(uint64_t) + sizeof(uint32_t));
  const uint32_t nslices = exact_slices((uint32_t)nq, (int)k, entries);
  if (nphys_ > 0 && nslices == 0) {
    set_error_static(SVDB_EINVAL, "queries x k too large for the allocation-free scratch");
    return SVDB_EINVAL;
  }
  const size_t nlist = (size_t)nq * nslices * kWavesPerBlock * k;
  int* dk = reinterpret_cast<int*>(base);
  double* qn = reinterpret_cast<double*>(base + kb);
  uint64_t* pk = reinterpret_cast<uint64_t*>(base + kb + qb);
  uint32_t* ps = reinterpret_cast<uint32_t*>(base + kb + qb + round_up(nlist * 8, 256));
  if (launch_fill_i32(dk, (uint32_t)nq, (int)k, c_->stream) != hipSuccess) {
    set_error_static(SVDB_EDEVICE, "k fill failed");
    return SVDB_EDEVICE;
  }
  const bool prof = prof_;
  prof_ = false;  // the profiling slots create events
  int rc = run_exact_with(dQ, ldq, (uint32_t)nq, (uint32_t)d_, dk, metric, (int)k, dids, dscores,
                          (uint32_t)k, nslices, pk, ps, qn);
  prof_ = prof;
  if (rc) return rc;
  if (hipStreamSynchronize(c_->stream) != hipSuccess) {
    set_error_static(SVDB_EDEVICE, "search failed on the device");
    return SVDB_EDEVICE;
  }
  st_.searches++;
  st_.queries += nq;
  return SVDB_OK;
}

int Engine::nearest1(const double* q, size_t D, size_t* id, bool* handled) {
  std::lock_guard<std::mutex> g(mu_);
  *handled = false;
  if (graphs_on_ || D == 0 || D > kNearestMaxD || D > d_ || nphys_ == 0) return SVDB_OK;
  c_ = &ctx_[0];
  int rc;
  const size_t part_b = kNearestMaxBlocks * (sizeof(uint64_t) + sizeof(uint32_t));
  if (!near_dev_) {
    // the ticket starts at 0 and every launch leaves it at 0
    HIPCHK(hipMalloc(&near_dev_, part_b + 256), "hipMalloc");
    HIPCHK(hipMemset(static_cast<char*>(near_dev_) + part_b, 0, 256), "hipMemset");
  }
  if (!near_host_) {
    HIPCHK(hipHostMalloc((void**)&near_host_, 64), "hipHostMalloc");
    near_host_[0] = near_host_[1] = 0;  // seq starts at 1
  }
  if ((rc = sync_alive())) return rc;
  Nearest1Args a{};
  a.st = StoreView{X_, ld_, xn_, alive_, rank_, (uint32_t)nphys_, has_deletes_ ? 1 : 0, nullptr};
  a.D = (uint32_t)D;
  a.part_key = static_cast<uint64_t*>(near_dev_);
  a.part_slot = reinterpret_cast<uint32_t*>(a.part_key + kNearestMaxBlocks);
  a.ticket = reinterpret_cast<uint32_t*>(static_cast<char*>(near_dev_) + part_b);
  a.out_id = near_host_;
  a.seq = ++near_seq_;
  for (size_t i = 0; i < D; i++) a.q[i] = q[i];
  HIPCHK(launch_nearest1(a, c_->stream), "k_nearest1");
  // The answer is in host memory as soon as seq is: poll it (a stream wait
  // costs several microseconds more than the kernel); after 50 ms without
  // it, wait for the stream, which also reports a failed launch.
  volatile uint64_t* hv = near_host_;
  const auto t0 = std::chrono::steady_clock::now();
  while (hv[1] != a.seq) {
    if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) {
      HIPCHK(hipStreamSynchronize(c_->stream), "k_nearest1");
      if (hv[1] != a.seq) {
        set_error(SVDB_EDEVICE, "k_nearest1: no answer");
        return SVDB_EDEVICE;
      }
      break;
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  const uint64_t v = hv[0];
  *id = v == ~0ull ? (size_t)-1 : (size_t)v;
  *handled = true;
  return SVDB_OK;
}

int Engine::search_host(const double* Q, size_t nq, size_t ldq, size_t D, const size_t* k,
                        const int* metric, int precision, size_t ostride, size_t* ids,
                        double* scores) {
  std::lock_guard<std::mutex> g(mu_);
  SearchCtx& c = ctx_[0];
  int rc = search_begin_locked(c, Q, nq, ldq, D, k, metric, precision, ostride, ids, scores);
  if (rc == SVDB_OK) rc = search_complete(c);
  c.state = 0;
  c_ = &ctx_[0];
  return rc;
}

int Engine::search_begin(int slot, const double* Q, size_t nq, size_t ldq, size_t D,
                         const size_t* k, const int* metric, int precision, size_t ostride,
                         size_t* ids, double* scores) {
  std::lock_guard<std::mutex> g(mu_);
  if (slot < 1 || slot >= kCtx || ctx_[slot].state != 0) {
    set_error(SVDB_EINVAL, "search slot busy or out of range");
    return SVDB_EINVAL;
  }
  SearchCtx& c = ctx_[slot];
  int rc = search_begin_locked(c, Q, nq, ldq, D, k, metric, precision, ostride, ids, scores);
  if (rc != SVDB_OK) c.state = 0;
  c_ = &ctx_[0];
  return rc;
}

int Engine::search_end(int slot) {
  std::lock_guard<std::mutex> g(mu_);
  if (slot < 1 || slot >= kCtx || ctx_[slot].state == 0) {
    set_error(SVDB_EINVAL, "no search in flight in this slot");
    return SVDB_EINVAL;
  }
  SearchCtx& c = ctx_[slot];
  const int rc = c.state == 1 ? search_complete(c) : c.rc;
  c.state = 0;
  c_ = &ctx_[0];
  return rc;
}

// (under mu_) completes the searches begun and not yet ended, so that a
// store mutation cannot change what they see; their results wait in the
// callers' arrays for search_end.
int Engine::complete_pending() {
  for (int i = 1; i < kCtx; i++)
    if (ctx_[i].state == 1) {
      ctx_[i].rc = search_complete(ctx_[i]);
      ctx_[i].state = 2;
    }
  c_ = &ctx_[0];
  return SVDB_OK;
}

// Runs enqueue — work recorded on c's stream with no host synchronisation
// inside — eagerly the first time its key is seen in this context, and
// from the second time on as one launch of a HIP graph captured from it.
// The key holds every value the sequence bakes into its launches that is
// not covered by store_ver_ and scratch_gen_ (which are part of it too), so
// a key match replays the same launches on the same buffers; the copies in
// it read and write pinned host staging at replay time, so new query rows
// and k values flow in and results out.  A capture that fails (an operation
// the stream cannot record) falls back to running the work eagerly.  At
// most kGraphs graphs per context, least recently used replaced.
int Engine::run_graphed(SearchCtx& c, const std::vector<uint64_t>& key,
                        const std::function<int()>& enqueue) {
  constexpr size_t kGraphs = 16, kSeen = 64;
  if (!graphs_on_ || prof_) return enqueue();
  for (auto& gr : c.graphs)
    if (gr.key == key) {
      gr.used = ++graph_clock_;
      HIPCHK(hipGraphLaunch(gr.exec, c.stream), "hipGraphLaunch");
      st_.graph_launches++;
      return SVDB_OK;
    }
  bool seen = false;
  for (const auto& sk : c.seen)
    if (sk == key) {
      seen = true;
      break;
    }
  if (!seen) {
    if (c.seen.size() >= kSeen) c.seen.erase(c.seen.begin());
    c.seen.push_back(key);
    return enqueue();
  }
  if (hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return enqueue();
  }
  const uint64_t gen = scratch_gen_;
  const int rc = enqueue();
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(c.stream, &graph);
  if (e != hipSuccess || !graph || scratch_gen_ != gen) {
    // the capture broke (nothing it recorded ran): run the work eagerly
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    return enqueue();
  }
  if (rc != SVDB_OK) {
    (void)hipGraphDestroy(graph);
    return rc;
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    (void)hipGetLastError();
    return enqueue();
  }
  if (c.graphs.size() >= kGraphs) {
    size_t lru = 0;
    for (size_t i = 1; i < c.graphs.size(); i++)
      if (c.graphs[i].used < c.graphs[lru].used) lru = i;
    (void)hipGraphExecDestroy(c.graphs[lru].exec);
    c.graphs.erase(c.graphs.begin() + lru);
  }
  c.graphs.push_back(SearchCtx::Graph{key, exec, ++graph_clock_});
  st_.graph_captures++;
  HIPCHK(hipGraphLaunch(exec, c.stream), "hipGraphLaunch");
  st_.graph_launches++;
  return SVDB_OK;
}

// Starts a host-query search in context c.  The BF16 form over rows <= 256
// wide (the serving case) is only enqueued: query upload, the first pass,
// shortlists, reranks and the result download run on c's stream, and
// search_complete finishes.  Every other form runs to completion here.
// c.state: 1 enqueued, 2 complete (c.rc).
int Engine::search_begin_locked(SearchCtx& c, const double* Q, size_t nq, size_t ldq, size_t D,
                                const size_t* k, const int* metric, int precision,
                                size_t ostride, size_t* ids, double* scores) {
  c_ = &c;
  c.state = 2;
  c.rc = SVDB_OK;
  if (!ok_) return SVDB_EDEVICE;
  if (hipSetDevice(device_) != hipSuccess) return SVDB_EDEVICE;
  if (nq == 0) return SVDB_OK;
  size_t kmax_all = 0;
  for (size_t i = 0; i < nq; i++) {
    if (k[i] < 1 || k[i] > SVDB_MAX_K) {
      set_error(SVDB_EINVAL, "k must be in [1, 128]");
      return SVDB_EINVAL;
    }
    kmax_all = std::max(kmax_all, k[i]);
  }
  if (ostride < kmax_all) {
    set_error(SVDB_EINVAL, "out_stride smaller than the largest k");
    return SVDB_EINVAL;
  }
  if (d_ == 0) {
    // Empty store: nothing to rank.
    for (size_t i = 0; i < nq; i++)
      for (size_t j = 0; j < k[i]; j++) {
        ids[i * ostride + j] = (size_t)-1;
        scores[i * ostride + j] = NAN;
      }
    return SVDB_OK;
  }
  int rc;
  for (size_t i = 0; i < nq; i++)
    if ((rc = check_search(d_, D, metric[i], precision))) return rc;
  fault_point();
  const int which = nphys_ > 0 ? use_lowp(precision, D, &rc) : 0;
  if (rc) return rc;
  const bool lowp = which != 0;
  const size_t step = lowp ? kMaxBatchLowp : kMaxBatch;

  std::vector<size_t>& sel = c.sel;
  // Host side of one chunk (query rows, k values, ids, scores): pinned
  // staging when it fits under its cap, else pageable vectors.
  std::vector<double> hq, hsc;
  std::vector<int> hk;
  std::vector<uint64_t> hid;
  double* pq = nullptr;
  int* pk = nullptr;
  uint64_t* pid = nullptr;
  double* psc = nullptr;
  auto host_bufs = [&](size_t n, size_t kk) {
    const size_t qb = round_up(n * D * sizeof(double), 64), kb = round_up(n * sizeof(int), 64),
                 ob = round_up(n * kk * sizeof(uint64_t), 64);
    if (char* st = static_cast<char*>(host_stage(&c.h_stage_, &c.h_stage_n_, qb + kb + 2 * ob))) {
      pq = reinterpret_cast<double*>(st);
      pk = reinterpret_cast<int*>(st + qb);
      pid = reinterpret_cast<uint64_t*>(st + qb + kb);
      psc = reinterpret_cast<double*>(st + qb + kb + ob);
      return true;
    }
    hq.resize(n * D);
    hk.resize(n);
    hid.resize(n * kk);
    hsc.resize(n * kk);
    pq = hq.data();
    pk = hk.data();
    pid = hid.data();
    psc = hsc.data();
    return false;
  };
  // The BF16 first pass over rows <= 256 wide: one pass for every metric of
  // the batch, queries grouped by metric into ranges, all at the batch's
  // largest k (a wavefront of k > 32 queries in a block sets that block's
  // pace anyway, and extra ranges cost more in latency-bound shortlist and
  // rerank launches than their narrower lists save).  Enqueued only.
  if (which == 1 && sh_.dp <= 256 && nq <= step) {
    LowpRange R[3];
    int nr = 0;
    sel.clear();
    for (int m = 0; m < 3; m++) {
      const size_t first = sel.size();
      for (size_t i = 0; i < nq; i++)
        if (metric[i] == m) sel.push_back(i);
      if (sel.size() > first)
        R[nr++] = LowpRange{m, (uint32_t)first, (uint32_t)(sel.size() - first), (int)kmax_all,
                            0, 0, 0, 0, 0};
    }
    const size_t n = sel.size();
    if ((rc = scratch(&c.s_q_, &c.s_q_n_, n * D * sizeof(double)))) return rc;
    if ((rc = scratch(&c.s_ids_, &c.s_ids_n_, n * kmax_all * sizeof(uint64_t)))) return rc;
    if ((rc = scratch(&c.s_sc_, &c.s_sc_n_, n * kmax_all * sizeof(double)))) return rc;
    if (!host_bufs(n, kmax_all)) {
      set_error(SVDB_ENOMEM, "pinned staging unavailable");
      return SVDB_ENOMEM;
    }
    c.kq.assign(k, k + nq);
    c.nq = nq;
    c.kmax = kmax_all;
    c.ostride = ostride;
    c.ids = ids;
    c.scores = scores;
    c.pid = pid;
    c.psc = psc;
    for (size_t t = 0; t < n; t++) std::memcpy(pq + t * D, Q + sel[t] * ldq, D * sizeof(double));
    HIPCHK(hipMemcpyAsync(c.s_q_, pq, n * D * sizeof(double), hipMemcpyHostToDevice, c.stream),
           "q upload");
    rc = lowp_enqueue((const double*)c.s_q_, D, R, nr, (int)kmax_all, (uint64_t*)c.s_ids_,
                      (double*)c.s_sc_, (uint32_t)kmax_all, false);
    if (rc) {
      c.lp.on = false;
      (void)hipStreamSynchronize(c.stream);
      return rc;
    }
    // the results of every query the first pass certifies (in practice all
    // of them): downloaded now, before anyone waits
    HIPCHK(hipMemcpyAsync(pid, c.s_ids_, n * kmax_all * sizeof(uint64_t), hipMemcpyDeviceToHost,
                          c.stream), "ids download");
    HIPCHK(hipMemcpyAsync(psc, c.s_sc_, n * kmax_all * sizeof(double), hipMemcpyDeviceToHost,
                          c.stream), "scores download");
    c.state = 1;
    return SVDB_OK;
  }
  for (int m = 0; m < 3; m++) {
    sel.clear();
    for (size_t i = 0; i < nq; i++)
      if (metric[i] == m) sel.push_back(i);
    for (size_t b0 = 0; b0 < sel.size(); b0 += step) {
      const size_t n = std::min(step, sel.size() - b0);
      size_t kmax = 0;
      for (size_t t = 0; t < n; t++) kmax = std::max(kmax, k[sel[b0 + t]]);
      if ((rc = scratch(&c.s_q_, &c.s_q_n_, n * D * sizeof(double)))) return rc;
      if ((rc = scratch(&c.s_k_, &c.s_k_n_, n * sizeof(int)))) return rc;
      if ((rc = scratch(&c.s_ids_, &c.s_ids_n_, n * kmax * sizeof(uint64_t)))) return rc;
      if ((rc = scratch(&c.s_sc_, &c.s_sc_n_, n * kmax * sizeof(double)))) return rc;
      const bool pinned = host_bufs(n, kmax);
      for (size_t t = 0; t < n; t++) {
        const size_t i = sel[b0 + t];
        std::memcpy(pq + t * D, Q + i * ldq, D * sizeof(double));
        pk[t] = (int)k[i];
      }
      auto enqueue = [&]() -> int {
        HIPCHK(hipMemcpyAsync(c.s_q_, pq, n * D * sizeof(double), hipMemcpyHostToDevice,
                              c.stream), "q upload");
        HIPCHK(hipMemcpyAsync(c.s_k_, pk, n * sizeof(int), hipMemcpyHostToDevice, c.stream),
               "k upload");
        const int r = lowp ? run_lowp((const double*)c.s_q_, D, (uint32_t)n, m, (int)kmax,
                                      (uint64_t*)c.s_ids_, (double*)c.s_sc_, (uint32_t)kmax,
                                      which == 2)
                           : run_exact((const double*)c.s_q_, D, (uint32_t)n, (uint32_t)D,
                                       (const int*)c.s_k_, m, (int)kmax, (uint64_t*)c.s_ids_,
                                       (double*)c.s_sc_, (uint32_t)kmax);
        if (r) return r;
        HIPCHK(hipMemcpyAsync(pid, c.s_ids_, n * kmax * sizeof(uint64_t),
                              hipMemcpyDeviceToHost, c.stream), "ids download");
        HIPCHK(hipMemcpyAsync(psc, c.s_sc_, n * kmax * sizeof(double), hipMemcpyDeviceToHost,
                              c.stream), "scores download");
        return SVDB_OK;
      };
      if (!lowp && pinned) {
        // The exact path as one graph launch once its shape repeats (a
        // handler answering one query per call): the tombstone upload a
        // mutation left pending runs first, outside any capture.
        if ((rc = sync_alive())) return rc;
        const std::vector<uint64_t> key = {
            1, (uint64_t)m, n, kmax, D, store_ver_, scratch_gen_, nphys_,
            (uint64_t)(uintptr_t)pq, (uint64_t)(uintptr_t)pk, (uint64_t)(uintptr_t)pid,
            (uint64_t)(uintptr_t)psc};
        rc = run_graphed(c, key, enqueue);
      } else {
        rc = enqueue();
      }
      if (rc) return rc;
      HIPCHK(hipStreamSynchronize(c.stream), "search");
      prof_collect();
      for (size_t t = 0; t < n; t++) {
        const size_t i = sel[b0 + t];
        for (size_t j = 0; j < k[i]; j++) {
          ids[i * ostride + j] = (size_t)pid[t * kmax + j];
          scores[i * ostride + j] = psc[t * kmax + j];
        }
      }
    }
  }
  st_.searches++;
  st_.queries += nq;
  return SVDB_OK;
}

// Finishes a search search_begin_locked enqueued in c: waits for it, runs
// the retries of the queries the first pass could not certify, and writes
// the caller's ids and scores.
int Engine::search_complete(SearchCtx& c) {
  if (c.state != 1) return c.rc;
  c_ = &c;
  c.state = 2;
  int rc = lowp_complete();
  if (rc) return c.rc = rc;
  const size_t n = c.sel.size(), kk = c.kmax;
  if (c.lp.failed) {
    // retries rewrote some queries' results on the device
    HIPCHK(hipMemcpyAsync(c.pid, c.s_ids_, n * kk * sizeof(uint64_t), hipMemcpyDeviceToHost,
                          c.stream), "ids download");
    HIPCHK(hipMemcpyAsync(c.psc, c.s_sc_, n * kk * sizeof(double), hipMemcpyDeviceToHost,
                          c.stream), "scores download");
  }
  HIPCHK(hipStreamSynchronize(c.stream), "search");
  prof_collect();
  for (size_t t = 0; t < n; t++) {
    const size_t i = c.sel[t];
    for (size_t j = 0; j < c.kq[i]; j++) {
      c.ids[i * c.ostride + j] = (size_t)c.pid[t * kk + j];
      c.scores[i * c.ostride + j] = c.psc[t * kk + j];
    }
  }
  st_.searches++;
  st_.queries += c.nq;
  return c.rc = SVDB_OK;
}

// ------------------------------------------------------------------ pairs

int Engine::pairs_f64(const size_t* a, const size_t* b, size_t np, int metric, double* out) {
  std::lock_guard<std::mutex> g(mu_);
  c_ = &ctx_[0];
  if (!ok_) return SVDB_EDEVICE;
  if (hipSetDevice(device_) != hipSuccess) return SVDB_EDEVICE;
  if (metric < 0 || metric > 2) {
    set_error(SVDB_EINVAL, "unknown metric");
    return SVDB_EINVAL;
  }
  if (np == 0) return SVDB_OK;
  fault_point();
  std::vector<uint32_t> hs(2 * np);
  for (size_t p = 0; p < np; p++) {
    if (a[p] >= nalive_ || b[p] >= nalive_) {
      set_error(SVDB_EINVAL, "index out of range");
      return SVDB_EINVAL;
    }
    hs[p] = select_slot(a[p]);
    hs[np + p] = select_slot(b[p]);
  }
  int rc;
  if ((rc = scratch(&c_->s_pair_, &c_->s_pair_n_, 2 * np * sizeof(uint32_t)))) return rc;
  if ((rc = scratch(&c_->s_pout_, &c_->s_pout_n_, np * sizeof(double)))) return rc;
  HIPCHK(hipMemcpyAsync(c_->s_pair_, hs.data(), 2 * np * sizeof(uint32_t), hipMemcpyHostToDevice, c_->stream), "pairs upload");
  const uint32_t* sa = (const uint32_t*)c_->s_pair_;
  HIPCHK(launch_pair_f64(X_, ld_, sa, sa + np, (uint32_t)np, (uint32_t)d_, metric,
                         (double*)c_->s_pout_, c_->stream), "k_pair_f64");
  HIPCHK(hipMemcpyAsync(out, c_->s_pout_, np * sizeof(double), hipMemcpyDeviceToHost, c_->stream), "pairs download");
  HIPCHK(hipStreamSynchronize(c_->stream), "pairs");
  return SVDB_OK;
}

}  // namespace svdb
