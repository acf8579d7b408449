// svdb_uuid.h — uuid → position index for vector_db_read_by_uuid, kept in
// step with the host mirror (db->vectors), whose positions shift down on a
// delete.  Replaces the front-to-back strncmp scan of vector_database.c:145-158.
//
// Rows get ids in insertion order, never reused.  A row's position is the
// number of live ids below its id, kept by a Fenwick tree over live flags:
// append, delete and position queries are O(log n) while the mirror itself
// shifts.  Each non-empty uuid maps to its live ids in ascending order; the
// lowest is the first match in position order, which is what the scan
// returns.  Empty uuids are not indexed (a lookup of "" scans).  The index
// stores no uuid text per row: callers pass a row's old uuid, read from the
// mirror, when it is deleted or renamed.  No method throws: an allocation
// failure marks the index stale, a stale index ignores mutations and answers
// no lookup, and clear() followed by append() of the whole mirror renews it.
#pragma once

#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <unordered_map>
#include <vector>

namespace svdb {

class UuidIndex {
 public:
  static constexpr size_t kNone = (size_t)-1;

  // n rows appended at the end; uuids[i] (may be NULL: no uuid) of at most
  // len bytes each
  void append(size_t n, const char* const* uuids, size_t len) noexcept {
    guarded([&] { append_(n, uuids, len); });
  }

  // the row at position pos (whose uuid is old_uuid) is removed
  void erase(size_t pos, const char* old_uuid, size_t len) noexcept {
    guarded([&] { erase_(pos, old_uuid, len); });
  }

  // the row at position pos changes its uuid from old_uuid to new_uuid
  void rename(size_t pos, const char* old_uuid, const char* new_uuid, size_t len) noexcept {
    guarded([&] { rename_(pos, old_uuid, new_uuid, len); });
  }

  // position of the first row with this uuid; kNone when none is indexed or
  // the index is stale
  size_t find(const char* uuid, size_t len) const noexcept {
    if (stale_) return kNone;
    try {
      auto it = by_uuid_.find(key(uuid, len));
      if (it == by_uuid_.end() || it->second.empty()) return kNone;
      return prefix(it->second.front());
    } catch (...) {
      return kNone;
    }
  }

  size_t size() const noexcept { return live_; }
  bool stale() const noexcept { return stale_; }
  void mark_stale() noexcept { stale_ = true; }

  void clear() noexcept {
    flag_.clear();
    bit_.clear();
    by_uuid_.clear();
    cap_ = 0;
    live_ = 0;
    stale_ = false;
  }

 private:
  template <class F>
  void guarded(F&& f) noexcept {
    if (stale_) return;
    try {
      f();
    } catch (...) {
      stale_ = true;
    }
  }

  void append_(size_t n, const char* const* uuids, size_t len) {
    const size_t id0 = flag_.size();
    if (id0 + n > cap_) {
      size_t c = cap_ ? cap_ : 1024;
      while (c < id0 + n) c *= 2;
      flag_.reserve(c);
      cap_ = c;
      flag_.resize(id0 + n, 1);
      rebuild();
    } else {
      flag_.resize(id0 + n, 1);
      for (size_t i = 0; i < n; i++) add(id0 + i, 1);
    }
    live_ += n;
    if (uuids)
      for (size_t i = 0; i < n; i++) {
        const std::string k = key(uuids[i], len);
        if (!k.empty()) by_uuid_[k].push_back(id0 + i);  // ids ascend
      }
  }

  void erase_(size_t pos, const char* old_uuid, size_t len) {
    if (pos >= live_) return;
    const size_t id = select(pos);
    drop(key(old_uuid, len), id);
    flag_[id] = 0;
    add(id, -1);
    live_--;
    if (flag_.size() > 2 * live_ + 4096) compact();
  }

  void rename_(size_t pos, const char* old_uuid, const char* new_uuid, size_t len) {
    if (pos >= live_) return;
    const std::string ko = key(old_uuid, len), kn = key(new_uuid, len);
    if (ko == kn) return;
    const size_t id = select(pos);
    drop(ko, id);
    if (!kn.empty()) {
      auto& v = by_uuid_[kn];
      v.insert(std::lower_bound(v.begin(), v.end(), id), id);
    }
  }

  static std::string key(const char* u, size_t len) {
    return u ? std::string(u, strnlen(u, len)) : std::string();
  }

  void drop(const std::string& k, size_t id) {
    if (k.empty()) return;
    auto it = by_uuid_.find(k);
    if (it == by_uuid_.end()) return;
    auto& v = it->second;
    auto p = std::lower_bound(v.begin(), v.end(), id);
    if (p != v.end() && *p == id) v.erase(p);
    if (v.empty()) by_uuid_.erase(it);
  }

  // Fenwick tree over ids (1-based inside): prefix(i) = live ids < i
  void add(size_t id, int d) {
    for (size_t i = id + 1; i <= cap_; i += i & (~i + 1)) bit_[i] += d;
  }
  size_t prefix(size_t id) const {
    int64_t s = 0;
    for (size_t i = id; i > 0; i -= i & (~i + 1)) s += bit_[i];
    return (size_t)s;
  }
  // id of the live row at position pos (0-based): binary lifting
  size_t select(size_t pos) const {
    size_t i = 0;
    int64_t rem = (int64_t)pos + 1;
    for (size_t step = cap_; step; step >>= 1)
      if (i + step <= cap_ && bit_[i + step] < rem) {
        i += step;
        rem -= bit_[i];
      }
    return i;  // (1-based i+1) -> id i
  }
  void rebuild() {
    bit_.assign(cap_ + 1, 0);
    for (size_t id = 0; id < flag_.size(); id++) bit_[id + 1] = flag_[id];
    for (size_t i = 1; i <= cap_; i++) {
      const size_t j = i + (i & (~i + 1));
      if (j <= cap_) bit_[j] += bit_[i];
    }
  }
  // renumber the live ids 0..live-1 in order (ids of deleted rows dropped)
  void compact() {
    std::vector<size_t> newid(flag_.size(), kNone);
    size_t n = 0;
    for (size_t id = 0; id < flag_.size(); id++)
      if (flag_[id]) newid[id] = n++;
    for (auto& kv : by_uuid_)
      for (auto& id : kv.second) id = newid[id];
    flag_.assign(n, 1);
    size_t c = 1024;
    while (c < n) c *= 2;
    cap_ = c;
    rebuild();
  }

  std::vector<char> flag_;     // per id: live
  std::vector<int32_t> bit_;   // Fenwick tree, cap_ + 1 entries
  size_t cap_ = 0;             // ids the tree covers (a power of two)
  size_t live_ = 0;
  bool stale_ = false;
  std::unordered_map<std::string, std::vector<size_t>> by_uuid_;
};

}  // namespace svdb
