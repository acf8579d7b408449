// svdb_noalloc.hip — the allocation-free entry point (svdb_search_device_noalloc).
//
// For callers that must not allocate (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc): the FP64 exact top-k of
// svdb_search_device with every intermediate — k per query, query norms and
// the partial top-k lists of the row slices — in `scratch`, a file-scope
// device array of kNoAllocScratchBytes that each GPU holds for the life of
// the process.  The caller passes queries and outputs in device memory.
//
// In host code `scratch` names a host-side stand-in, not a device address:
// the device address is taken once per device with hipGetSymbolAddress
// (with that device current) and only that address is used.  Calls are
// serialised by a process-wide lock (they share the array); the array keeps
// the last call's contents, and nothing in it has to start at zero (every
// partial-list entry the merge reads is written by the scan of the same
// call).
#include <hip/hip_runtime.h>

#include <mutex>

#include "svdb.h"
#include "svdb_engine.h"

namespace {

constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;  // 64 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_mu;               // one call at a time: the array is shared
void* g_ptr[kMaxDevices] = {};  // its device address, per device

// (under g_mu) the array's address on `dev` (current), looked up once
void* scratch_address(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  if (!g_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    g_ptr[dev] = p;
  }
  return g_ptr[dev];
}

}  // namespace

extern "C" int svdb_search_device_noalloc(VectorDatabase* db, const double* d_queries, size_t nq,
                                          size_t ldq, size_t k, int metric, uint64_t* d_ids,
                                          double* d_scores) {
  svdb::Engine* e = db ? static_cast<svdb::Engine*>(db->svdb_engine) : nullptr;
  if (!e || !d_queries || !d_ids || !d_scores) {
    svdb::set_error_static(SVDB_EINVAL, "NULL argument");
    return SVDB_EINVAL;
  }
  const int dev = svdb::device_init();
  if (dev < 0) return SVDB_EDEVICE;
  std::lock_guard<std::mutex> g(g_mu);
  void* p = scratch_address(dev);
  if (!p) {
    svdb::set_error_static(SVDB_EDEVICE, "hipGetSymbolAddress of the scratch array failed");
    return SVDB_EDEVICE;
  }
  return e->search_device_noalloc(d_queries, nq, ldq, k, metric, d_ids, d_scores, p,
                                  kNoAllocScratchBytes);
}
