// svdb_engine.h — device store + search engine behind the C ABI (internal).
//
// One Engine per VectorDatabase (and per stand-alone KDTree).  It owns, in
// HBM:
//   X       FP64 master rows, slot-major, row stride ld = round_up(d, 8)
//   nx2, xn FP64 Σ fma(x,x) and its sqrt per slot
//   alive   1 bit per slot; rank_base[w] = alive slots before word w
// Slots are append-only: an insert takes the next slot, an update rewrites
// its slot in place, a delete clears the slot's alive bit (a tombstone).
// Slot order therefore equals logical (positional) order, so the lowest
// slot on a tie is the lowest logical index, and a logical index is the
// rank of its slot among the alive slots.  Compaction squeezes tombstones
// out when they pass a quarter of the slots (and number at least 256), and
// rebuilds the shadows, their degenerate lists and ρ from the live rows.
//
// Every public method takes the engine mutex: callers may enter from any
// thread (the reference server runs one thread per connection).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "svdb.h"
#include "svdb_kernels.h"
#include "svdb_lowp.h"

namespace svdb {

void set_error(int code, const std::string& msg);
void set_error_static(int code, const char* msg);  // allocates nothing
const char* last_error();

// Host-allocation fault injection for tests (svdb_debug_fail_host_alloc):
// placed before host allocations that may fail; when armed, the n-th call
// throws std::bad_alloc, as a failed allocation there would.
void fault_point();
void arm_fault(int64_t n);  // 0 disarms
int64_t fault_armed();     // calls left before the fault (0: disarmed)

// The reference's float arithmetic for one pair of host vectors of length n
// (vector_database.c:301-352), computed on the GPU: both vectors are staged
// into device memory and k_pair_ref32 runs on a process-wide stream.
int pair_ref32_host(const double* a, const double* b, size_t n, int metric, float* out);

// Host-mirror (db->vectors) halves of insert / update / delete, applied
// after the device store's half by the serving loop, in mutation order.
// mirror_reserve grows the mirror's capacity to `rows` ahead of time.
int mirror_reserve(VectorDatabase* db, size_t rows);
void mirror_insert(VectorDatabase* db, Vector vec);
void mirror_update(VectorDatabase* db, size_t index, double* data, size_t dim);
void mirror_delete(VectorDatabase* db, size_t index);

// Device used by every engine in this process (SVDB_DEVICE, default 0).
int device_init();

struct KernelTime {
  uint64_t launches = 0;
  double total_ms = 0.0;
};

struct EngineStats {
  uint64_t searches = 0, queries = 0, certified = 0, fallbacks = 0, widened = 0, second_pass = 0;
  uint64_t compactions = 0;
  uint64_t graph_launches = 0, graph_captures = 0;
  double last_max_rel_err = 0.0;
};

class Engine {
 public:
  explicit Engine(size_t initial_slots);
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  bool ok() const { return ok_; }
  size_t dim() const { return d_; }
  size_t rows() const { return nalive_; }
  size_t phys_rows() const { return nphys_; }
  size_t device_bytes() const { return dev_bytes_; }
  size_t budget() const { return budget_; }

  // ---- store (host rows; row i at rows + i*lds) ----
  int append(const double* rows, size_t n, size_t lds, size_t dim, uint32_t* first_slot);
  int overwrite(size_t logical, const double* row, size_t dim);
  int erase(size_t logical);
  int slot_of(size_t logical, uint32_t* slot);

  // ---- search ----
  // Host queries; per-query k and metric; ids/scores row stride ostride.
  int search_host(const double* Q, size_t nq, size_t ldq, size_t D, const size_t* k,
                  const int* metric, int precision, size_t ostride, size_t* ids, double* scores);
  // kdtree_nearest's one-query form: the nearest live row over the first D
  // columns in one launch (k_nearest1), the id read back from host memory.
  // Sets *handled = false, and does nothing, where it does not apply (D
  // above kNearestMaxD or the row width, or graph replay switched on); the
  // caller then takes search_host.
  int nearest1(const double* q, size_t D, size_t* id, bool* handled);
  // Split form of search_host for the serving loop: begin enqueues the
  // search on context `slot` (1 or 2) and returns; end completes it (waits,
  // runs any retries) and writes ids / scores, which must stay valid from
  // begin to end.  Two searches may be in flight at once; a store mutation
  // completes them first.
  int search_begin(int slot, const double* Q, size_t nq, size_t ldq, size_t D, const size_t* k,
                   const int* metric, int precision, size_t ostride, size_t* ids,
                   double* scores);
  int search_end(int slot);
  // Device queries and outputs; one k and metric.
  int search_device(const double* dQ, size_t nq, size_t ldq, size_t D, size_t k, int metric,
                    int precision, uint64_t* dids, double* dscores);
  // The FP64 exact scan of search_device with every intermediate in the
  // caller's device scratch (`bytes` at `scratch`, 256-byte aligned):
  // allocates nothing, host or device.
  int search_device_noalloc(const double* dQ, size_t nq, size_t ldq, size_t k, int metric,
                            uint64_t* dids, double* dscores, void* scratch, size_t bytes);

  // ---- pair metrics over stored rows (logical indices) ----
  int pairs_f64(const size_t* a, const size_t* b, size_t np, int metric, double* out);

  // ---- knobs / stats ----
  int set_budget(size_t bytes);
  int set_shortlist(size_t L);
  int set_profiling(bool on);
  int enable_shadow(int precision);
  bool shadow_on() const { return lowp_; }
  bool shadow8_on() const { return lowp8_; }
  void kernel_times(KernelTime* out);
  EngineStats stats();

  std::mutex& mutex() { return mu_; }

 private:
  int ensure_dim(size_t d);
  int ensure_capacity(size_t slots);
  int sync_alive();
  int refresh_degen();
  int compact();
  static constexpr size_t kCompactMin = 256;  // fewest tombstones worth a compaction
  int dev_alloc(void** p, size_t bytes);
  void dev_free(void* p, size_t bytes);
  int scratch(void** p, size_t* have, size_t need);
  // pinned staging, grow-only; nullptr past its cap
  void* host_stage(void** buf, size_t* have, size_t bytes);
  int run_exact(const double* dQ, size_t ldq, uint32_t nq, uint32_t D, const int* dk, int metric,
                int kmax, uint64_t* dids, double* dscores, uint32_t ostride);
  int run_exact_with(const double* dQ, size_t ldq, uint32_t nq, uint32_t D, const int* dk,
                     int metric, int kmax, uint64_t* dids, double* dscores, uint32_t ostride,
                     uint32_t nslices, uint64_t* pk, uint32_t* ps, double* qn);
  uint32_t exact_slices(uint32_t nq, int kmax, size_t max_entries) const;
  struct LowpRange {
    int metric;
    uint32_t qoff, n;  // query rows qoff .. qoff+n-1
    int k;             // neighbours wanted (0: the call's k)
    uint32_t soff;     // first slot (set by run_lowp_ranges)
    // first-pass row slicing of this range's scan and where its candidate
    // lists start (set by run_lowp_ranges)
    uint32_t ns, rps, NL;
    size_t coff;
  };
  static uint32_t lowp_width(int kk, bool fp8, size_t shortlist);
  int run_lowp(const double* dQ, size_t ldq, uint32_t nq, int metric, int k, uint64_t* dids,
               double* dscores, uint32_t ostride, bool fp8);
  int run_lowp_ranges(const double* dQ, size_t ldq, LowpRange* R, int nr, int k, uint64_t* dids,
                      double* dscores, uint32_t ostride, bool fp8);
  // 0: FP64 scan, 1: BF16 first pass, 2: FP8 first pass
  int use_lowp(int precision, size_t D, int* rc);
  int alloc_shadow(size_t cap);
  int alloc_shadow8(size_t cap);
  int shadow_rows(uint32_t first, uint32_t n, const uint32_t* dslots);
  uint32_t select_slot(size_t logical) const;
  void prof_begin(int which);
  void prof_end(int which);
  void prof_collect();
  hipEvent_t take_event();

  bool ok_ = false;
  int device_ = 0;
  hipStream_t stream_ = nullptr;
  std::mutex mu_;

  size_t d_ = 0, ld_ = 0;
  size_t cap_ = 0, nphys_ = 0, nalive_ = 0, init_slots_ = 0;
  double* X_ = nullptr;
  double* nx2_ = nullptr;
  double* xn_ = nullptr;
  uint64_t* alive_ = nullptr;
  uint32_t* rank_ = nullptr;
  std::vector<uint64_t> h_alive_;
  std::vector<uint32_t> h_rank_;
  bool has_deletes_ = false;
  bool degen_dirty_ = false;  // degenerate lists may hold deleted / rewritten slots
  // The widening retries' recent share of queries left uncertified, per
  // precision (BF16, FP8), metric and width (1024, 4096); see lowp_complete.
  float retry_miss_[2][3][2] = {};
  uint32_t retry_seen_[2][3][2] = {};
  size_t dirty_word_ = SIZE_MAX;  // first bitmap word not yet on the device

  // low-precision shadow
  bool lowp_ = false;
  bool lowp_wanted_ = true;
  bool prepass_ = true;  // sample pre-pass before the first-pass scan (SVDB_PREPASS=0: off)
  static constexpr uint32_t kPrepassSlices = 256;  // 64 on stores of under 16*256 tiles
  static constexpr uint32_t kPrepassMaxTiles = 8;
  Shadow sh_;
  bool lowp8_ = false;  // FP8 shadow present (built on request)
  Shadow8 sh8_;

  size_t dev_bytes_ = 0;
  size_t budget_ = SIZE_MAX;
  size_t shortlist_ = 64;

  struct ProfSlot {
    int kind;
    hipEvent_t a, b;
  };

  // Everything one search call owns while it runs: its stream, its
  // grow-only device scratch and pinned staging, its profiling slots, and —
  // for a search split into begin / end (search_begin) — what end needs.
  // Context 0 serves the synchronous entry points; the serving loop keeps
  // up to two searches in flight on contexts 1 and 2, whose kernels overlap
  // on the device.  c_ is the context the current call works in (under mu_).
  struct SearchCtx {
    hipStream_t stream = nullptr;
    void* s_q_ = nullptr;      size_t s_q_n_ = 0;
    void* s_qn_ = nullptr;     size_t s_qn_n_ = 0;
    void* s_k_ = nullptr;      size_t s_k_n_ = 0;
    void* s_pk_ = nullptr;     size_t s_pk_n_ = 0;
    void* s_ps_ = nullptr;     size_t s_ps_n_ = 0;
    void* s_ids_ = nullptr;    size_t s_ids_n_ = 0;
    void* s_sc_ = nullptr;     size_t s_sc_n_ = 0;
    void* s_pair_ = nullptr;   size_t s_pair_n_ = 0;
    void* s_pout_ = nullptr;   size_t s_pout_n_ = 0;
    void* s_qb_ = nullptr;     size_t s_qb_n_ = 0;
    void* s_qp_ = nullptr;     size_t s_qp_n_ = 0;
    void* s_ck_ = nullptr;     size_t s_ck_n_ = 0;
    void* s_cs_ = nullptr;     size_t s_cs_n_ = 0;
    void* s_sk_ = nullptr;     size_t s_sk_n_ = 0;
    void* s_ss_ = nullptr;     size_t s_ss_n_ = 0;
    void* s_beta_ = nullptr;   size_t s_beta_n_ = 0;
    void* s_fail_ = nullptr;   size_t s_fail_n_ = 0;
    void* s_err_ = nullptr;    size_t s_err_n_ = 0;
    void* s_lk_ = nullptr;     size_t s_lk_n_ = 0;    // uncertified queries' lists: keys
    void* s_ls_ = nullptr;     size_t s_ls_n_ = 0;    //   slots
    void* s_km_ = nullptr;     size_t s_km_n_ = 0;    //   shortlists' largest key
    void* s_fq_ = nullptr;     size_t s_fq_n_ = 0;
    void* s_fid_ = nullptr;    size_t s_fid_n_ = 0;
    void* s_fsc_ = nullptr;    size_t s_fsc_n_ = 0;
    void* s_fk_ = nullptr;     size_t s_fk_n_ = 0;
    void* s_sk2_ = nullptr;    size_t s_sk2_n_ = 0;   // widening retry: shortlist keys
    void* s_ss2_ = nullptr;    size_t s_ss2_n_ = 0;   //   slots
    void* s_beta2_ = nullptr;  size_t s_beta2_n_ = 0; //   bounds
    void* s_fail2_ = nullptr;  size_t s_fail2_n_ = 0; //   queries still uncertified
    void* s_cidx_ = nullptr;   size_t s_cidx_n_ = 0;  // FP8 -> BF16 second pass: query indices
    void* s_cq_ = nullptr;     size_t s_cq_n_ = 0;    //   gathered queries
    void* s_cid_ = nullptr;    size_t s_cid_n_ = 0;   //   their ids
    void* s_csc_ = nullptr;    size_t s_csc_n_ = 0;   //   and scores
    uint32_t* h_fail_ = nullptr;  // pinned: fail count
    void* h_io_ = nullptr;        // pinned: per-range fail counts + first-pass errors
    size_t h_io_n_ = 0;
    void* h_stage_ = nullptr;     // pinned: search staging
    size_t h_stage_n_ = 0;
    std::vector<ProfSlot> prof_pending_;
    // a low-precision pass enqueued and not yet completed (lowp_enqueue)
    struct LowpPending {
      bool on = false;
      LowpRange R[3];
      int nr = 0, k = 0;
      const double* dQ = nullptr;
      size_t ldq = 0;
      uint64_t* dids = nullptr;
      double* dscores = nullptr;
      uint32_t ostride = 0;
      bool fp8 = false;
      uint32_t nq = 0;  // slots
      size_t shortlist = 0;
      bool failed = false;  // some query went through a retry (set by lowp_complete)
    } lp;
    // an asynchronous search (search_begin): 0 idle, 1 enqueued, 2 completed
    int state = 0;
    int rc = 0;
    size_t nq = 0, kmax = 0, ostride = 0;
    std::vector<size_t> sel;   // slot order -> query index
    std::vector<size_t> kq;    // per query k
    size_t* ids = nullptr;     // the caller's outputs
    double* scores = nullptr;
    uint64_t* pid = nullptr;   // pinned results (in h_stage_)
    double* psc = nullptr;
    // Launch sequences of this context captured as HIP graphs, by key
    // (run_graphed), and the keys run once eagerly so far.
    struct Graph {
      std::vector<uint64_t> key;
      hipGraphExec_t exec = nullptr;
      uint64_t used = 0;
    };
    std::vector<Graph> graphs;
    std::vector<std::vector<uint64_t>> seen;
  };
  static constexpr int kCtx = 3;
  SearchCtx ctx_[kCtx];
  SearchCtx* c_ = &ctx_[0];
  void* h_stage_ = nullptr;  // pinned: row upload staging
  size_t h_stage_n_ = 0;
  void* near_dev_ = nullptr;        // k_nearest1: block results and ticket
  uint64_t* near_host_ = nullptr;   // pinned: its answer (id, seq)
  uint64_t near_seq_ = 0;
  int lowp_enqueue(const double* dQ, size_t ldq, LowpRange* R, int nr, int k, uint64_t* dids,
                   double* dscores, uint32_t ostride, bool fp8);
  int lowp_complete();
  int search_begin_locked(SearchCtx& c, const double* Q, size_t nq, size_t ldq, size_t D,
                          const size_t* k, const int* metric, int precision, size_t ostride,
                          size_t* ids, double* scores);
  int search_complete(SearchCtx& c);
  int complete_pending();  // completes every begun search (before a store mutation)
  int run_graphed(SearchCtx& c, const std::vector<uint64_t>& key,
                  const std::function<int()>& enqueue);
  // Everything a captured launch sequence bakes in besides its key's own
  // values: the store (rows, pointers, tombstones, shadows, shortlist) and
  // the scratch and staging buffers.  Any change bumps one of these.
  uint64_t store_ver_ = 0;
  uint64_t scratch_gen_ = 0;
  uint64_t graph_clock_ = 0;
  bool graphs_on_ = false;  // SVDB_GRAPHS=1

  bool prof_ = false;
  bool prof_mute_ = false;  // inside the BF16 second pass: timed as a whole
  std::vector<hipEvent_t> ev_pool_;
  KernelTime kt_[SVDB_K_COUNT];
  EngineStats st_;
};

}  // namespace svdb
