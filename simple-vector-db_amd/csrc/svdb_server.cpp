// svdb_server.cpp — continuous-batching serving loop over one store.
//
// Requests (searches and mutations) enter one FIFO.  A device-owner thread
// runs ticks: it applies mutations in arrival order, and gathers every
// search queued behind them (up to max_batch queries, and while the
// estimated device memory of the tick stays within the budget) into one
// batch — any mix of sizes, k and metrics — answered by one call into the
// engine (one first pass over the store for the whole tick, by metric).
// Requests that arrive while a tick is on the GPU wait for the next tick;
// the queries of the
// next tick are staged (copied and grouped on the host) by a second thread
// while the current tick runs, and searches keep joining it until the
// device thread takes it.  A mutation is a barrier: searches queued
// before it see the store without it, searches after it see it.  The host
// mirror (db->vectors) follows the device store on a third thread, in the
// same order; a mutation completes once both halves are applied, and a
// search completes only once the mirror holds every mutation submitted
// before it, so the ids it returns name the same rows in vector_db_read.
// No exception leaves a worker thread: a failed host allocation fails the
// request it was made for (SVDB_ENOMEM).
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <memory>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "svdb_engine.h"

namespace svdb {

static int64_t now_us() {
  return std::chrono::duration_cast<std::chrono::microseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct Request {
  int64_t ticket = 0;
  int kind = 0;  // 0 search, 1 insert, 2 update, 3 delete
  std::vector<double> data;
  size_t nq = 0, k = 0, index = 0, dim = 0;
  int metric = 0;
  size_t est_bytes = 0;
  std::vector<size_t> ids;
  std::vector<double> scores;
  int rc = 0;
  bool done = false;
  int64_t t_submit = 0, t_start = 0, t_done = 0;
  uint64_t tick = 0;
  size_t tick_queries = 0;
};

// A staged tick: concatenated queries with per-query k and metric.
struct Tick {
  std::vector<std::shared_ptr<Request>> reqs;
  std::vector<double> q;
  std::vector<size_t> k;
  std::vector<int> metric;
  size_t kmax = 0, nq = 0, bytes = 0;
  // A tick is either a run of consecutive mutations (applied in order) or
  // a batch of searches.
  std::vector<std::shared_ptr<Request>> mutations;
  // searches: mutations handed to the mirror before this tick ran; the
  // tick completes once the mirror has applied that many
  uint64_t mirror_seq = 0;
};

class Server {
 public:
  static constexpr size_t kMaxMutationsPerTick = 1024;

  Server(VectorDatabase* db, const svdb_server_config& cfg) : db_(db), cfg_(cfg) {
    eng_ = static_cast<Engine*>(db->svdb_engine);
    stager_ = std::thread([this] { stage_loop(); });
    exec_ = std::thread([this] { exec_loop(); });
    mirror_ = std::thread([this] { mirror_loop(); });
  }

  ~Server() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_q_.notify_all();
    cv_tick_.notify_all();
    stager_.join();
    exec_.join();
    {
      std::lock_guard<std::mutex> g(mu_);
      mirror_stop_ = true;  // after the executor: no more mirror work comes
    }
    cv_mirror_.notify_all();
    mirror_.join();
    // requests still queued are failed, not dropped
    std::lock_guard<std::mutex> g(mu_);
    for (auto& r : queue_) {
      r->rc = SVDB_EINVAL;
      r->done = true;
    }
    cv_done_.notify_all();
  }

  int64_t submit(std::shared_ptr<Request> r) {
    std::lock_guard<std::mutex> g(mu_);
    if (stop_) return SVDB_EINVAL;
    r->ticket = next_++;
    r->t_submit = now_us();
    all_[r->ticket] = r;  // may throw (host memory): nothing queued then
    try {
      queue_.push_back(r);
    } catch (...) {
      all_.erase(r->ticket);
      throw;
    }
    cv_q_.notify_all();
    return r->ticket;
  }

  int wait(int64_t ticket, size_t* ids, double* scores, svdb_request_info* info) {
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> g(mu_);
      auto it = all_.find(ticket);
      if (it == all_.end()) {
        set_error(SVDB_EINVAL, "unknown ticket");
        return SVDB_EINVAL;
      }
      r = it->second;
      cv_done_.wait(g, [&] { return r->done; });
      all_.erase(ticket);
    }
    if (r->kind == 0 && r->rc == SVDB_OK) {
      if (ids) std::memcpy(ids, r->ids.data(), r->ids.size() * sizeof(size_t));
      if (scores) std::memcpy(scores, r->scores.data(), r->scores.size() * sizeof(double));
    }
    if (r->kind == 1 && ids) ids[0] = r->index;
    if (info) {
      info->submit_us = r->t_submit;
      info->start_us = r->t_start;
      info->done_us = r->t_done;
      info->tick = r->tick;
      info->tick_queries = r->tick_queries;
    }
    return r->rc;
  }

  size_t query_bytes(size_t nq, size_t k) const {
    // FP64 query + BF16 query + parameters + results + first-pass candidate
    // lists (at most 512 slices × 4 lists × 8 entries of 8 bytes)
    const size_t d = eng_->dim() ? eng_->dim() : 1;
    return nq * (d * 8 + d * 2 + 64 + k * 16 + 512 * 4 * 8 * 8);
  }

  const svdb_server_config& cfg() const { return cfg_; }
  size_t dim() const { return eng_->dim(); }

 private:
  // Stager: forms the next tick while the executor runs the current one,
  // and keeps adding the searches that arrive meanwhile to it (while they
  // fit) until the executor takes it; a search then waits at most for the
  // tick that is running when it arrives.
  void stage_loop() {
    std::unique_lock<std::mutex> g(mu_);
    for (;;) {
      cv_q_.wait(g, [&] { return stop_ || can_stage(); });
      if (stop_) return;
      if (!staged_ && queue_.front()->kind != 0) {
        std::unique_ptr<Tick> t(new (std::nothrow) Tick);
        bool room = t != nullptr;
        if (room) {
          try {
            fault_point();
            t->mutations.reserve(kMaxMutationsPerTick);
          } catch (const std::bad_alloc&) {
            room = false;
          }
        }
        if (!room) {  // no host memory for the tick: fail the head request
          fail_head(SVDB_ENOMEM);
          continue;
        }
        while (!queue_.empty() && queue_.front()->kind != 0 &&
               t->mutations.size() < kMaxMutationsPerTick) {
          t->mutations.push_back(queue_.front());  // capacity reserved
          queue_.pop_front();
        }
        staged_ = std::move(t);
      } else {
        std::unique_ptr<Tick> fresh;
        Tick* t = staged_.get();
        if (!t) {
          fresh.reset(new (std::nothrow) Tick);
          t = fresh.get();
          if (!t) {
            fail_head(SVDB_ENOMEM);
            continue;
          }
        }
        const size_t d = eng_->dim();
        while (!queue_.empty() && queue_.front()->kind == 0) {
          auto r = queue_.front();
          if (r->dim != d) {  // store dimension changed under it
            queue_.pop_front();
            r->rc = SVDB_EINVAL;
            r->done = true;
            cv_done_.notify_all();
            continue;
          }
          if (!t->reqs.empty() &&
              (t->nq + r->nq > cfg_.max_batch || t->bytes + r->est_bytes > cfg_.budget_bytes))
            break;
          queue_.pop_front();
          const size_t o = t->nq;
          try {  // room first: a host allocation failure fails this request only
            fault_point();
            t->q.resize((o + r->nq) * d);
            t->k.resize(o + r->nq, r->k);
            t->metric.resize(o + r->nq, r->metric);
            t->reqs.reserve(t->reqs.size() + 1);
          } catch (const std::bad_alloc&) {
            r->rc = SVDB_ENOMEM;
            r->done = true;
            cv_done_.notify_all();
            continue;
          }
          std::memcpy(&t->q[o * d], r->data.data(), r->nq * d * sizeof(double));
          t->kmax = std::max(t->kmax, r->k);
          t->nq += r->nq;
          t->bytes += r->est_bytes;
          t->reqs.push_back(r);
        }
        if (fresh && !fresh->reqs.empty()) staged_ = std::move(fresh);
      }
      cv_tick_.notify_all();
    }
  }

  // (under mu_) fails the request at the head of the queue
  void fail_head(int rc) {
    auto r = queue_.front();
    queue_.pop_front();
    r->rc = rc;
    r->t_start = r->t_done = now_us();
    r->done = true;
    cv_done_.notify_all();
  }

  // (under mu_) whether the stager has work: a queue head that starts a new
  // tick, or a search that joins the staged search tick
  bool can_stage() const {
    if (queue_.empty()) return false;
    if (!staged_) return true;
    const Tick& t = *staged_;
    const Request& r = *queue_.front();
    if (!t.mutations.empty() || r.kind != 0) return false;
    if (r.dim != eng_->dim()) return true;
    return t.nq + r.nq <= cfg_.max_batch && t.bytes + r.est_bytes <= cfg_.budget_bytes;
  }

  // A search tick on the device: begun in an engine slot (1 or 2), its
  // results land in ids / sc when it is ended.
  struct Flight {
    std::unique_ptr<Tick> t;
    int slot = 0;
    int64_t t0 = 0;
    uint64_t tick = 0;
    int rc = SVDB_OK;
    std::vector<size_t> ids;
    std::vector<double> sc;
  };

  // Executor: keeps up to two search ticks in flight, so one tick's first
  // pass runs on the device while the previous one's shortlists, reranks
  // and result copies finish and its requests are answered.  A mutation
  // tick waits until no search is in flight (those searches were submitted
  // before it), then applies its mutations in order.
  void exec_loop() {
    std::deque<std::unique_ptr<Flight>> fl;
    for (;;) {
      std::unique_ptr<Tick> t;
      bool finish_oldest = false;
      {
        std::unique_lock<std::mutex> g(mu_);
        if (fl.empty()) cv_tick_.wait(g, [&] { return stop_ || staged_; });
        const bool can_take =
            staged_ && (staged_->mutations.empty() ? fl.size() < 2 : fl.empty());
        if (can_take) {
          t = std::move(staged_);
          t->mirror_seq = mirror_issued_;
        } else if (!fl.empty()) {
          finish_oldest = true;
        } else if (stop_) {
          return;
        } else {
          continue;
        }
      }
      if (finish_oldest) {
        end_flight(*fl.front());
        fl.pop_front();
        continue;
      }
      cv_q_.notify_all();  // the stager may form the next tick now
      const int64_t t0 = now_us();
      const uint64_t tick = ++ticks_;
      if (!t->mutations.empty()) {
        // Device halves now (the next tick's searches must see them); host
        // mirror halves, in the same order, on the mirror thread, which
        // completes the requests.  The mirror queue's node is allocated
        // before the device half runs, so a mutation applied on the device
        // always reaches the mirror.
        for (auto& m : t->mutations) {
          std::list<std::shared_ptr<Request>> node;
          try {
            fault_point();
            node.push_back(m);
            run_mutation(*m);
          } catch (const std::bad_alloc&) {
            m->rc = SVDB_ENOMEM;
          } catch (...) {
            m->rc = SVDB_EINVAL;
          }
          if (m->rc != SVDB_OK || node.empty()) {
            if (m->rc == SVDB_OK) m->rc = SVDB_ENOMEM;
            finish(m, t0, tick, 0);
          } else {
            std::lock_guard<std::mutex> g(mu_);
            m->t_start = t0;
            m->tick = tick;
            mirror_q_.splice(mirror_q_.end(), node);  // allocates nothing
            mirror_issued_++;
          }
        }
        cv_mirror_.notify_all();
        continue;
      }
      // a search tick: begin it in the free slot
      std::unique_ptr<Flight> f(new (std::nothrow) Flight);
      if (!f) {
        fail_tick(*t, SVDB_ENOMEM, t0, tick);
        continue;
      }
      f->slot = (!fl.empty() && fl.front()->slot == 1) ? 2 : 1;
      f->t0 = t0;
      f->tick = tick;
      const size_t d = eng_->dim();
      try {  // a host allocation failure fails the tick's searches
        fault_point();
        f->ids.resize(t->nq * t->kmax);
        f->sc.resize(t->nq * t->kmax);
        f->rc = eng_->search_begin(f->slot, t->q.data(), t->nq, d, d, t->k.data(),
                                   t->metric.data(), cfg_.precision, t->kmax, f->ids.data(),
                                   f->sc.data());
      } catch (const std::bad_alloc&) {
        f->rc = SVDB_ENOMEM;
      } catch (...) {
        f->rc = SVDB_EINVAL;
      }
      f->t = std::move(t);
      if (f->rc != SVDB_OK) {  // nothing in flight for this tick
        fail_tick(*f->t, f->rc, t0, tick);
        continue;
      }
      try {
        fl.push_back(std::move(f));
      } catch (...) {
        // (the deque could not grow: answer this tick now)
        end_flight(*f);
      }
    }
  }

  // Ends a flight: waits for its search, hands each request its rows of the
  // results, and completes them once the mirror shows the store they saw.
  void end_flight(Flight& f) {
    int rc = eng_->search_end(f.slot);
    Tick& t = *f.t;
    size_t o = 0;
    for (auto& r : t.reqs) {
      r->rc = rc;
      if (rc == SVDB_OK) {
        try {
          fault_point();
          r->ids.resize(r->nq * r->k);
          r->scores.resize(r->nq * r->k);
        } catch (const std::bad_alloc&) {
          r->rc = SVDB_ENOMEM;
          o += r->nq;
          continue;
        }
        for (size_t i = 0; i < r->nq; i++)
          for (size_t j = 0; j < r->k; j++) {
            r->ids[i * r->k + j] = f.ids[(o + i) * t.kmax + j];
            r->scores[i * r->k + j] = f.sc[(o + i) * t.kmax + j];
          }
      }
      o += r->nq;
    }
    // the ids name rows of the store as the mutations before this tick
    // left it: the mirror (vector_db_read) must show that state first.  If
    // it does not yet, the mirror thread completes the requests once it
    // does, and the executor goes on with the next tick.
    {
      std::unique_lock<std::mutex> g(mu_);
      if (mirror_applied_ < t.mirror_seq) {
        bool handed = false;
        try {
          fault_point();
          after_mirror_.push_back(Waiting{t.mirror_seq, f.t0, f.tick, t.nq, std::move(t.reqs)});
          handed = true;
        } catch (...) {
        }
        if (handed) {
          cv_mirror_.notify_all();
          return;
        }
        // (no room to hand over: wait here instead)
        cv_applied_.wait(g, [&] { return mirror_applied_ >= t.mirror_seq; });
      }
    }
    finish(t.reqs, f.t0, f.tick, t.nq);
  }

  // Search requests answered, waiting for the mirror to reach `seq`.
  struct Waiting {
    uint64_t seq;
    int64_t t0;
    uint64_t tick;
    size_t nq;
    std::vector<std::shared_ptr<Request>> reqs;
  };

  void fail_tick(Tick& t, int rc, int64_t t0, uint64_t tick) {
    for (auto& r : t.reqs) r->rc = rc;
    finish(t.reqs, t0, tick, t.nq);
  }

  // Device half of a mutation (the engine store and its shadow).  Indices
  // are logical positions at this point of the mutation order, which the
  // engine tracks itself (the host mirror may lag behind).
  void run_mutation(Request& r) {
    const size_t rows = eng_->rows();
    if (r.kind == 1) {
      r.index = rows;
      if ((r.rc = mirror_reserve(db_, rows + 1)) == SVDB_OK)
        r.rc = eng_->append(r.data.data(), 1, r.dim, r.dim, nullptr);
    } else if (r.index >= rows) {
      set_error(SVDB_EINVAL, "index out of range");
      r.rc = SVDB_EINVAL;
    } else if (r.kind == 2) {
      r.rc = eng_->overwrite(r.index, r.data.data(), r.dim);
    } else {
      r.rc = eng_->erase(r.index);
    }
  }

  // Host-mirror half, in mutation order; completes the request.  The
  // mirror functions allocate nothing that can throw (malloc / realloc,
  // reserved ahead by mirror_reserve); a row whose copy cannot be
  // allocated is mirrored with a NULL data pointer, which the reference's
  // handlers treat as an invalid row.
  void mirror_loop() {
    for (;;) {
      std::shared_ptr<Request> r;
      {
        std::unique_lock<std::mutex> g(mu_);
        cv_mirror_.wait(g, [&] {
          return mirror_stop_ || !mirror_q_.empty() || ready_after_mirror();
        });
        if (ready_after_mirror()) {
          // searches handed over by the executor whose mutations are mirrored
          Waiting w = std::move(after_mirror_.front());
          after_mirror_.pop_front();
          g.unlock();
          finish(w.reqs, w.t0, w.tick, w.nq);
          continue;
        }
        if (mirror_q_.empty()) return;
        r = mirror_q_.front();
        mirror_q_.pop_front();
      }
      if (r->kind == 1 || r->kind == 2) {
        double* data = (double*)malloc(r->dim * sizeof(double));
        if (data) std::memcpy(data, r->data.data(), r->dim * sizeof(double));
        if (r->kind == 1) {
          Vector v;
          std::memset(&v, 0, sizeof(v));
          v.dimension = r->dim;
          v.data = data;
          mirror_insert(db_, v);
        } else {
          mirror_update(db_, r->index, data, r->dim);
        }
      } else {
        mirror_delete(db_, r->index);
      }
      {
        std::lock_guard<std::mutex> g(mu_);
        mirror_applied_++;
      }
      cv_applied_.notify_all();
      finish(r, r->t_start, r->tick, 0);
    }
  }

  // (under mu_) the oldest handed-over search tick can complete
  bool ready_after_mirror() const {
    return !after_mirror_.empty() && after_mirror_.front().seq <= mirror_applied_;
  }

  // Completes requests (allocates nothing).
  void finish(const std::vector<std::shared_ptr<Request>>& rs, int64_t t0, uint64_t tick,
              size_t nq) {
    finish(rs.data(), rs.size(), t0, tick, nq);
  }
  void finish(const std::shared_ptr<Request>& r, int64_t t0, uint64_t tick, size_t nq) {
    finish(&r, 1, t0, tick, nq);
  }
  void finish(const std::shared_ptr<Request>* rs, size_t n, int64_t t0, uint64_t tick,
              size_t nq) {
    const int64_t t1 = now_us();
    {
      std::lock_guard<std::mutex> g(mu_);
      for (size_t i = 0; i < n; i++) {
        Request& r = *rs[i];
        r.t_start = t0;
        r.t_done = t1;
        r.tick = tick;
        r.tick_queries = nq;
        r.done = true;
      }
    }
    cv_done_.notify_all();
  }

  VectorDatabase* db_;
  Engine* eng_;
  svdb_server_config cfg_;
  std::mutex mu_;
  std::condition_variable cv_q_, cv_tick_, cv_done_, cv_mirror_, cv_applied_;
  std::deque<std::shared_ptr<Request>> queue_;
  std::list<std::shared_ptr<Request>> mirror_q_;
  uint64_t mirror_issued_ = 0;   // mutations handed to the mirror thread
  uint64_t mirror_applied_ = 0;  // ... and applied by it
  std::deque<Waiting> after_mirror_;  // answered searches waiting for the mirror
  bool mirror_stop_ = false;
  std::unique_ptr<Tick> staged_;
  std::unordered_map<int64_t, std::shared_ptr<Request>> all_;
  int64_t next_ = 0;
  uint64_t ticks_ = 0;
  bool stop_ = false;
  std::thread stager_, exec_, mirror_;
};

}  // namespace svdb

using svdb::Request;
using svdb::Server;

namespace {
// The C entry points report a host allocation failure as SVDB_ENOMEM (any
// other exception as SVDB_EINVAL) and let no exception cross the C boundary.
template <class F>
int64_t c_call(F&& f) {
  int code = SVDB_EINVAL;
  const char* msg = "internal error";
  try {
    return f();
  } catch (const std::bad_alloc&) {
    code = SVDB_ENOMEM;
    msg = "host allocation failed";
  } catch (...) {
  }
  try {
    svdb::set_error(code, msg);
  } catch (...) {
  }
  return code;
}
}  // namespace

extern "C" {

void svdb_server_default_config(svdb_server_config* cfg) {
  if (!cfg) return;
  cfg->max_batch = 8192;
  cfg->budget_bytes = (size_t)64 << 30;
  cfg->precision = SVDB_PREC_AUTO;
}

svdb_server* svdb_server_start(VectorDatabase* db, const svdb_server_config* cfg) {
  if (!db || !db->svdb_engine) {
    svdb::set_error(SVDB_EINVAL, "NULL database");
    return nullptr;
  }
  svdb_server_config c;
  svdb_server_default_config(&c);
  if (cfg) c = *cfg;
  if (c.max_batch == 0 || c.budget_bytes == 0) {
    svdb::set_error(SVDB_EINVAL, "max_batch and budget_bytes must be positive");
    return nullptr;
  }
  Server* sv = nullptr;
  if (c_call([&]() -> int64_t {
        sv = new Server(db, c);
        return SVDB_OK;
      }) != SVDB_OK)
    return nullptr;
  return reinterpret_cast<svdb_server*>(sv);
}

void svdb_server_stop(svdb_server* s) { delete reinterpret_cast<Server*>(s); }

int64_t svdb_server_submit_search(svdb_server* s, const double* queries, size_t nq, size_t dim,
                                  size_t k, int metric) {
  Server* sv = reinterpret_cast<Server*>(s);
  if (!sv || !queries || nq == 0 || k < 1 || k > SVDB_MAX_K || metric < 0 || metric > 2) {
    svdb::set_error(SVDB_EINVAL, "bad search request");
    return SVDB_EINVAL;
  }
  return c_call([&]() -> int64_t {
    auto r = std::make_shared<Request>();
    r->kind = 0;
    r->nq = nq;
    r->k = k;
    r->metric = metric;
    r->dim = dim;
    if (sv->dim() != 0 && dim != sv->dim()) {
      svdb::set_error(SVDB_EINVAL, "query dimension differs from the store's");
      return SVDB_EINVAL;
    }
    r->est_bytes = sv->query_bytes(nq, k);
    if (r->est_bytes > sv->cfg().budget_bytes || nq > sv->cfg().max_batch) {
      svdb::set_error(SVDB_ENOMEM, "request larger than the server's batch or memory budget");
      return SVDB_ENOMEM;
    }
    r->data.assign(queries, queries + nq * dim);
    return sv->submit(r);
  });
}

int64_t svdb_server_submit_insert(svdb_server* s, const double* row, size_t dim) {
  Server* sv = reinterpret_cast<Server*>(s);
  if (!sv || !row || dim == 0) {
    svdb::set_error(SVDB_EINVAL, "bad insert request");
    return SVDB_EINVAL;
  }
  return c_call([&]() -> int64_t {
    auto r = std::make_shared<Request>();
    r->kind = 1;
    r->dim = dim;
    r->data.assign(row, row + dim);
    return sv->submit(r);
  });
}

int64_t svdb_server_submit_update(svdb_server* s, size_t index, const double* row, size_t dim) {
  Server* sv = reinterpret_cast<Server*>(s);
  if (!sv || !row || dim == 0) {
    svdb::set_error(SVDB_EINVAL, "bad update request");
    return SVDB_EINVAL;
  }
  return c_call([&]() -> int64_t {
    auto r = std::make_shared<Request>();
    r->kind = 2;
    r->index = index;
    r->dim = dim;
    r->data.assign(row, row + dim);
    return sv->submit(r);
  });
}

int64_t svdb_server_submit_delete(svdb_server* s, size_t index) {
  Server* sv = reinterpret_cast<Server*>(s);
  if (!sv) {
    svdb::set_error(SVDB_EINVAL, "NULL server");
    return SVDB_EINVAL;
  }
  return c_call([&]() -> int64_t {
    auto r = std::make_shared<Request>();
    r->kind = 3;
    r->index = index;
    return sv->submit(r);
  });
}

int svdb_server_wait(svdb_server* s, int64_t ticket, size_t* ids, double* scores,
                     svdb_request_info* info) {
  Server* sv = reinterpret_cast<Server*>(s);
  if (!sv) {
    svdb::set_error(SVDB_EINVAL, "NULL server");
    return SVDB_EINVAL;
  }
  return (int)c_call([&]() -> int64_t { return sv->wait(ticket, ids, scores, info); });
}

}  // extern "C"
