// svdb_abi.cpp — the C ABI: the reference's store / kd-tree / metric
// functions (vector_database.h, kdtree.h) and the batch entry points
// (svdb.h), all served by the device engine.
//
// Host-side semantics follow vector_database.c / kdtree.c: positional
// indices, delete shifts later rows down, insert takes ownership of
// vec.data, read returns a borrowed pointer into the host mirror, sentinel
// returns ((size_t)-1, NULL, -1.0f) with a message on stderr.  Every
// similarity computation runs on the GPU; there is no host compute path.
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "kdtree.h"
#include "svdb.h"
#include "svdb_engine.h"
#include "svdb_mirror.h"
#include "svdb_uuid.h"
#include "vector_database.h"

using svdb::Engine;
using svdb::UuidIndex;

static_assert(sizeof(Vector) == 56, "Vector layout must match the reference");
static_assert(offsetof(Vector, dimension) == 40 && offsetof(Vector, data) == 48, "Vector layout");
static_assert(offsetof(VectorDatabase, vectors) == 0 && offsetof(VectorDatabase, size) == 8 &&
                  offsetof(VectorDatabase, capacity) == 16 && offsetof(VectorDatabase, kdtree) == 24 &&
                  offsetof(VectorDatabase, mutex) == 32,
              "VectorDatabase layout must match the reference");
static_assert(offsetof(KDTree, root) == 0 && offsetof(KDTree, dimension) == 8, "KDTree layout");

namespace {

// What KDTree::svdb_index points to.
struct TreeIndex {
  Engine* eng = nullptr;   // the database's engine, or the tree's own
  bool owned = false;      // stand-alone tree: eng and labels are ours
  std::vector<size_t> labels;                    // slot → label (stand-alone)
  std::unordered_map<size_t, size_t> by_label;   // label → logical row (stand-alone)
};

Engine* eng_of(VectorDatabase* db) { return db ? static_cast<Engine*>(db->svdb_engine) : nullptr; }
UuidIndex* uuids_of(VectorDatabase* db) {
  return db ? static_cast<UuidIndex*>(db->svdb_uuid) : nullptr;
}
// Runs an engine call from a C entry point: a host allocation failure is
// reported as SVDB_ENOMEM (any other exception as SVDB_EINVAL), and no
// exception crosses the C boundary (the engine's own lock is released by
// its guard on the way out).
template <class F>
int c_call(F&& f) {
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

// the index's record of one new row at the end of the mirror
void index_append(VectorDatabase* db, const Vector& v) {
  const char* u = v.uuid;
  uuids_of(db)->append(1, &u, UUID_SIZE);
}

KDTree* make_tree(size_t dimension, Engine* attached) {
  KDTree* t = (KDTree*)calloc(1, sizeof(KDTree));
  if (!t) return nullptr;
  TreeIndex* ix = new (std::nothrow) TreeIndex();
  if (!ix) {
    free(t);
    return nullptr;
  }
  if (attached) {
    ix->eng = attached;
  } else {
    ix->eng = new (std::nothrow) Engine(1024);
    ix->owned = true;
    if (!ix->eng || !ix->eng->ok()) {
      fprintf(stderr, "kdtree_create: %s\n", svdb::last_error());
      delete ix->eng;
      delete ix;
      free(t);
      return nullptr;
    }
  }
  t->root = nullptr;
  t->dimension = dimension;
  t->svdb_index = ix;
  return t;
}

int check_metric_int(int m) { return m >= 0 && m <= 2; }

}  // namespace

extern "C" {

// ======================================================================
// kdtree.h
// ======================================================================

KDTree* kdtree_create(size_t dimension) {
  if (dimension == 0) return nullptr;
  return make_tree(dimension, nullptr);
}

void kdtree_insert(KDTree* tree, const double* point, size_t index) {
  if (!tree || !tree->svdb_index || !point) return;
  TreeIndex* ix = static_cast<TreeIndex*>(tree->svdb_index);
  if (!ix->owned) return;  // the database store already holds the row
  auto it = ix->by_label.find(index);
  if (it != ix->by_label.end()) {
    if (c_call([&] { return ix->eng->overwrite(it->second, point, tree->dimension); }) != SVDB_OK)
      fprintf(stderr, "kdtree_insert: %s\n", svdb::last_error());
    return;
  }
  // label bookkeeping is allocated before the row is appended, so an
  // allocation failure leaves the tree unchanged
  try {
    ix->labels.reserve(ix->labels.size() + 1);
    ix->by_label.emplace(index, ix->labels.size());
  } catch (...) {
    fprintf(stderr, "kdtree_insert: out of host memory\n");
    return;
  }
  uint32_t slot = 0;
  if (c_call([&] { return ix->eng->append(point, 1, tree->dimension, tree->dimension, &slot); }) !=
      SVDB_OK) {
    fprintf(stderr, "kdtree_insert: %s\n", svdb::last_error());
    ix->by_label.erase(index);
    return;
  }
  ix->labels.push_back(index);  // capacity reserved above
}

void kdtree_free(KDTree* tree) {
  if (!tree) return;
  TreeIndex* ix = static_cast<TreeIndex*>(tree->svdb_index);
  if (ix) {
    if (ix->owned) delete ix->eng;
    delete ix;
  }
  free(tree);
}

size_t kdtree_nearest(KDTree* tree, const double* point) {
  if (!tree || !tree->svdb_index || !point) return (size_t)-1;
  TreeIndex* ix = static_cast<TreeIndex*>(tree->svdb_index);
  Engine* e = ix->eng;
  if (e->rows() == 0) return (size_t)-1;
  const size_t D = tree->dimension;
  if (D == 0 || D > e->dim()) {
    fprintf(stderr, "kdtree_nearest: tree dimension %zu outside [1, %zu]\n", D, e->dim());
    return (size_t)-1;
  }
  size_t k = 1, id = (size_t)-1;
  int metric = SVDB_METRIC_L2;
  double score;
  // One query: the low-precision pass carries ~100 µs of fixed work (query
  // prep, pre-pass, shortlist, rerank, certification); the FP64 scan reads
  // the compared columns of every row once, 512 MB at 2^26 elements, about
  // as long.  Below that the FP64 scan answers (the same ids either way).
  const bool small = (double)e->phys_rows() * (double)D <= (double)(1u << 26);
  const int prec = D == e->dim() && !small ? SVDB_PREC_AUTO : SVDB_PREC_FP64;
  // small stores: one launch answers (k_nearest1; the same id as the scan)
  bool handled = false;
  if (small && c_call([&] { return e->nearest1(point, D, &id, &handled); }) != SVDB_OK) {
    fprintf(stderr, "kdtree_nearest: %s\n", svdb::last_error());
    return (size_t)-1;
  }
  if (!handled &&
      c_call([&] { return e->search_host(point, 1, D, D, &k, &metric, prec, 1, &id, &score); }) !=
          SVDB_OK) {
    fprintf(stderr, "kdtree_nearest: %s\n", svdb::last_error());
    return (size_t)-1;
  }
  if (ix->owned && id != (size_t)-1) return id < ix->labels.size() ? ix->labels[id] : (size_t)-1;
  return id;
}

// ======================================================================
// vector_database.h
// ======================================================================

VectorDatabase* vector_db_init(size_t initial_capacity, size_t dimension) {
  VectorDatabase* db = (VectorDatabase*)calloc(1, sizeof(VectorDatabase));
  if (!db) {
    fprintf(stderr, "Failed to allocate memory for database\n");
    return nullptr;
  }
  db->size = 0;
  db->capacity = initial_capacity > 0 ? initial_capacity : 10;
  if (db->capacity > SIZE_MAX / sizeof(Vector)) {
    fprintf(stderr, "Capacity overflow detected, unable to allocate memory for vectors\n");
    free(db);
    return nullptr;
  }
  db->vectors = (Vector*)malloc(db->capacity * sizeof(Vector));
  if (!db->vectors) {
    fprintf(stderr, "Failed to allocate memory for vectors\n");
    free(db);
    return nullptr;
  }
  Engine* e = new (std::nothrow) Engine(db->capacity);
  if (!e || !e->ok()) {
    fprintf(stderr, "vector_db_init: %s\n", svdb::last_error());
    delete e;
    free(db->vectors);
    free(db);
    return nullptr;
  }
  db->svdb_engine = e;
  db->svdb_uuid = new (std::nothrow) UuidIndex();
  if (!db->svdb_uuid) {
    fprintf(stderr, "Failed to allocate the uuid index\n");
    delete e;
    free(db->vectors);
    free(db);
    return nullptr;
  }
  db->kdtree = make_tree(dimension, e);
  if (!db->kdtree) {
    fprintf(stderr, "Failed to create KDTree\n");
    delete uuids_of(db);
    delete e;
    free(db->vectors);
    free(db);
    return nullptr;
  }
  if (pthread_mutex_init(&db->mutex, nullptr) != 0) {
    fprintf(stderr, "Failed to initialize mutex\n");
    kdtree_free(db->kdtree);
    delete uuids_of(db);
    delete e;
    free(db->vectors);
    free(db);
    return nullptr;
  }
  return db;
}

void vector_db_free(VectorDatabase* db) {
  if (!db) return;
  for (size_t i = 0; i < db->size; ++i) free(db->vectors[i].data);
  kdtree_free(db->kdtree);
  delete eng_of(db);
  delete uuids_of(db);
  free(db->vectors);
  pthread_mutex_destroy(&db->mutex);
  free(db);
}

static int grow_mirror(VectorDatabase* db, size_t need) {
  if (need <= db->capacity) return 0;
  size_t nc = db->capacity;
  while (nc < need) {
    if (nc > SIZE_MAX / 2 / sizeof(Vector)) {
      fprintf(stderr, "Capacity overflow detected, unable to allocate more memory for vectors\n");
      return -1;
    }
    nc *= 2;
  }
  Vector* nv = (Vector*)realloc(db->vectors, nc * sizeof(Vector));
  if (!nv) {
    fprintf(stderr, "Failed to allocate more memory for vectors\n");
    return -1;
  }
  db->vectors = nv;
  db->capacity = nc;
  return 0;
}

size_t vector_db_insert(VectorDatabase* db, Vector vec) {
  if (!db) return (size_t)-1;
  pthread_mutex_lock(&db->mutex);
  if (grow_mirror(db, db->size + 1) != 0) {
    pthread_mutex_unlock(&db->mutex);
    return (size_t)-1;
  }
  if (!vec.data || c_call([&] {
        return eng_of(db)->append(vec.data, 1, vec.dimension, vec.dimension, nullptr);
      }) != SVDB_OK) {
    fprintf(stderr, "vector_db_insert: %s\n", vec.data ? svdb::last_error() : "NULL data");
    pthread_mutex_unlock(&db->mutex);
    return (size_t)-1;
  }
  db->vectors[db->size] = vec;
  index_append(db, vec);
  if (db->kdtree) kdtree_insert(db->kdtree, vec.data, db->size);
  size_t index = db->size++;
  pthread_mutex_unlock(&db->mutex);
  return index;
}

Vector* vector_db_read(VectorDatabase* db, size_t index) {
  if (!db) return nullptr;
  pthread_mutex_lock(&db->mutex);
  Vector* v = index < db->size ? &db->vectors[index] : nullptr;
  pthread_mutex_unlock(&db->mutex);
  return v;
}

// The first row (lowest index) whose uuid equals `uuid` (compared as
// strncmp over UUID_SIZE), as vector_database.c:145-158 finds by scanning;
// here an index lookup, exact as long as uuids change only through these
// entry points.  An empty uuid, or an index caught disagreeing with the
// mirror (its row no longer holds the uuid, its count is off, or it is
// stale after an allocation failure), falls back to the scan, and the
// latter rebuilds the index.
Vector* vector_db_read_by_uuid(VectorDatabase* db, const char* uuid) {
  if (!db || !uuid) return nullptr;
  pthread_mutex_lock(&db->mutex);
  Vector* v = nullptr;
  UuidIndex* ix = uuids_of(db);
  const size_t pos = uuid[0] ? ix->find(uuid, UUID_SIZE) : UuidIndex::kNone;
  if (pos != UuidIndex::kNone && pos < db->size &&
      strncmp(db->vectors[pos].uuid, uuid, UUID_SIZE) == 0 && ix->size() == db->size) {
    v = &db->vectors[pos];
  } else {
    for (size_t i = 0; i < db->size; ++i)
      if (strncmp(db->vectors[i].uuid, uuid, UUID_SIZE) == 0) {
        v = &db->vectors[i];
        break;
      }
    // the index missed a hit, held a stale entry or lost count: rebuild it
    // from the mirror
    const bool agreed = pos == UuidIndex::kNone && v == nullptr && !ix->stale() &&
                        ix->size() == db->size;
    if (uuid[0] && !agreed) {
      ix->clear();
      try {
        std::vector<const char*> us(db->size);
        for (size_t i = 0; i < db->size; ++i) us[i] = db->vectors[i].uuid;
        ix->append(db->size, us.data(), UUID_SIZE);
      } catch (...) {
        ix->mark_stale();  // rebuilt by the next lookup
      }
    }
  }
  pthread_mutex_unlock(&db->mutex);
  return v;
}

void vector_db_update(VectorDatabase* db, size_t index, Vector vec) {
  if (!db) return;
  pthread_mutex_lock(&db->mutex);
  if (index < db->size && vec.data &&
      c_call([&] { return eng_of(db)->overwrite(index, vec.data, vec.dimension); }) == SVDB_OK) {
    uuids_of(db)->rename(index, db->vectors[index].uuid, vec.uuid, UUID_SIZE);
    free(db->vectors[index].data);
    db->vectors[index] = vec;
    if (db->kdtree) kdtree_insert(db->kdtree, vec.data, index);
  } else {
    if (index < db->size) fprintf(stderr, "vector_db_update: %s\n", svdb::last_error());
    free(vec.data);
  }
  pthread_mutex_unlock(&db->mutex);
}

}  // extern "C"

// ---- host-mirror halves of the mutations, for the serving loop ----------
// The serving loop applies a mutation to the device store at once (searches
// queued behind it must see it) and to the host mirror (db->vectors) later,
// in the same order, on a thread of its own: a delete shifts the mirror,
// which for a large store takes far longer than the device update and is
// not needed to answer searches.
namespace svdb {

int mirror_reserve(VectorDatabase* db, size_t rows) {
  pthread_mutex_lock(&db->mutex);
  const int rc = grow_mirror(db, rows);
  pthread_mutex_unlock(&db->mutex);
  return rc ? SVDB_ENOMEM : SVDB_OK;
}

void mirror_insert(VectorDatabase* db, Vector vec) {
  pthread_mutex_lock(&db->mutex);
  if (grow_mirror(db, db->size + 1) == 0) {  // reserved by mirror_reserve
    index_append(db, vec);
    db->vectors[db->size++] = vec;
  } else
    free(vec.data);
  pthread_mutex_unlock(&db->mutex);
}

void mirror_update(VectorDatabase* db, size_t index, double* data, size_t dim) {
  pthread_mutex_lock(&db->mutex);
  if (index < db->size) {
    free(db->vectors[index].data);
    db->vectors[index].data = data;
    db->vectors[index].dimension = dim;
  } else {
    free(data);
  }
  pthread_mutex_unlock(&db->mutex);
}

void mirror_delete(VectorDatabase* db, size_t index) {
  pthread_mutex_lock(&db->mutex);
  if (index < db->size) {
    uuids_of(db)->erase(index, db->vectors[index].uuid, UUID_SIZE);
    free(db->vectors[index].data);
    svdb::shift_down_one(db->vectors, index, db->size);
    db->size--;
  }
  pthread_mutex_unlock(&db->mutex);
}

}  // namespace svdb

extern "C" {

void vector_db_delete(VectorDatabase* db, size_t index) {
  if (!db) return;
  pthread_mutex_lock(&db->mutex);
  if (index < db->size) {
    if (c_call([&] { return eng_of(db)->erase(index); }) != SVDB_OK) {
      fprintf(stderr, "vector_db_delete: %s\n", svdb::last_error());
    } else {
      uuids_of(db)->erase(index, db->vectors[index].uuid, UUID_SIZE);
      free(db->vectors[index].data);
      svdb::shift_down_one(db->vectors, index, db->size);
      db->size--;
    }
  }
  pthread_mutex_unlock(&db->mutex);
}

void vector_db_save(VectorDatabase* db, const char* filename) {
  if (!db || !filename) return;
  pthread_mutex_lock(&db->mutex);
  FILE* f = fopen(filename, "wb");
  if (!f) {
    perror("Failed to open file for writing");
    pthread_mutex_unlock(&db->mutex);
    return;
  }
  size_t n = 0;
  for (size_t i = 0; i < db->size; ++i)
    if (db->vectors[i].dimension != 0 && db->vectors[i].data) n++;
  fwrite(&n, sizeof(size_t), 1, f);
  for (size_t i = 0; i < db->size; ++i) {
    const Vector& v = db->vectors[i];
    if (v.dimension == 0 || !v.data) {
      fprintf(stderr, "Invalid vector at index %zu, skipping\n", i);
      continue;
    }
    fwrite(v.uuid, sizeof(char), UUID_SIZE, f);
    fwrite(&v.dimension, sizeof(size_t), 1, f);
    fwrite(v.data, sizeof(double), v.dimension, f);
  }
  fclose(f);
  pthread_mutex_unlock(&db->mutex);
}

VectorDatabase* vector_db_load(const char* filename, size_t dimension) {
  FILE* f = filename ? fopen(filename, "rb") : nullptr;
  if (!f) {
    perror("Failed to open file for reading");
    return nullptr;
  }
  size_t n = 0;
  if (fread(&n, sizeof(size_t), 1, f) != 1) {
    fprintf(stderr, "vector_db_load: truncated header\n");
    fclose(f);
    return nullptr;
  }
  // The header's count sizes the store only as far as the file can hold
  // that many records (each at least uuid[37] + size_t): a damaged or
  // crafted count must not become a huge allocation.
  size_t cap = n;
  long here = ftell(f);
  if (here >= 0 && fseek(f, 0, SEEK_END) == 0) {
    const long end = ftell(f);
    if (end >= here) cap = std::min(cap, (size_t)(end - here) / (UUID_SIZE + sizeof(size_t)));
    if (fseek(f, here, SEEK_SET) != 0) {
      perror("Failed to read file");
      fclose(f);
      return nullptr;
    }
  }
  VectorDatabase* db = vector_db_init(cap > 0 ? cap : 10, dimension);
  if (!db) {
    fclose(f);
    return nullptr;
  }
  // Read every record into the host mirror, then upload in bulk: one
  // device pass per run of equal-dimension rows.
  std::vector<double> batch;
  size_t batch_dim = 0, batch_first = 0;
  bool fail = false;
  auto flush = [&]() {
    size_t rows = batch_dim ? batch.size() / batch_dim : 0;
    if (rows && c_call([&] {
          return eng_of(db)->append(batch.data(), rows, batch_dim, batch_dim, nullptr);
        }) != SVDB_OK) {
      fprintf(stderr, "vector_db_load: %s\n", svdb::last_error());
      fail = true;
    }
    batch.clear();
    (void)batch_first;
  };
  for (size_t i = 0; i < n && !fail; ++i) {
    Vector v;
    memset(&v, 0, sizeof(v));
    if (fread(v.uuid, 1, UUID_SIZE, f) != UUID_SIZE || fread(&v.dimension, sizeof(size_t), 1, f) != 1) {
      fprintf(stderr, "vector_db_load: truncated record %zu\n", i);
      fail = true;
      break;
    }
    v.data = (double*)malloc((v.dimension ? v.dimension : 1) * sizeof(double));
    if (!v.data || fread(v.data, sizeof(double), v.dimension, f) != v.dimension) {
      fprintf(stderr, "vector_db_load: truncated record %zu\n", i);
      free(v.data);
      fail = true;
      break;
    }
    if (batch_dim != v.dimension && !batch.empty()) flush();
    batch_dim = v.dimension;
    try {
      batch.insert(batch.end(), v.data, v.data + v.dimension);
    } catch (...) {
      fprintf(stderr, "vector_db_load: out of host memory at record %zu\n", i);
      free(v.data);
      fail = true;
      break;
    }
    if (grow_mirror(db, db->size + 1) != 0) {
      free(v.data);
      fail = true;
      break;
    }
    index_append(db, v);
    db->vectors[db->size++] = v;
    if (batch.size() >= (size_t)1 << 24) flush();
  }
  if (!fail) flush();
  fclose(f);
  if (fail) {
    vector_db_free(db);
    return nullptr;
  }
  return db;
}

static float legacy_metric(const Vector& a, const Vector& b, int metric) {
  if (a.dimension != b.dimension) {
    fprintf(stderr, "Vectors have different dimensions\n");
    return -1.0f;
  }
  float out = NAN;
  if (svdb::pair_ref32_host(a.data, b.data, a.dimension, metric, &out) != SVDB_OK) {
    fprintf(stderr, "similarity metric: %s\n", svdb::last_error());
    return NAN;
  }
  return out;
}

float cosine_similarity(Vector vec1, Vector vec2) { return legacy_metric(vec1, vec2, SVDB_METRIC_COS); }
float euclidean_distance(Vector vec1, Vector vec2) { return legacy_metric(vec1, vec2, SVDB_METRIC_L2); }
float dot_product(Vector vec1, Vector vec2) { return legacy_metric(vec1, vec2, SVDB_METRIC_DOT); }

// ======================================================================
// svdb.h
// ======================================================================

const char* svdb_last_error(void) { return svdb::last_error(); }

int64_t svdb_debug_fail_host_alloc(int64_t n) {
  if (n >= 0) svdb::arm_fault(n);
  return svdb::fault_armed();
}

void* svdb_device_alloc(size_t bytes) {
  if (svdb::device_init() < 0) return nullptr;
  void* p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
    svdb::set_error(SVDB_ENOMEM, "hipMalloc failed");
    return nullptr;
  }
  return p;
}

void svdb_device_free(void* ptr) {
  if (ptr && svdb::device_init() >= 0) (void)hipFree(ptr);
}

int svdb_device_copy(void* dst, const void* src, size_t bytes, int kind) {
  if (svdb::device_init() < 0) return SVDB_EDEVICE;
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                        : kind == 1 ? hipMemcpyDeviceToHost
                        : kind == 2 ? hipMemcpyDeviceToDevice
                                    : hipMemcpyDefault;
  if (kind < 0 || kind > 2) {
    svdb::set_error(SVDB_EINVAL, "copy kind must be 0, 1 or 2");
    return SVDB_EINVAL;
  }
  if (kind == 2) {  // device → device with a plain copy kernel
    if (svdb::launch_copy(src, dst, bytes, nullptr) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
      svdb::set_error(SVDB_EDEVICE, "device copy failed");
      return SVDB_EDEVICE;
    }
    return SVDB_OK;
  }
  if (hipMemcpy(dst, src, bytes, k) != hipSuccess) {
    svdb::set_error(SVDB_EDEVICE, "hipMemcpy failed");
    return SVDB_EDEVICE;
  }
  return SVDB_OK;
}

int svdb_mfma_sustained(int kind, double* tflops) {
  if (!tflops || kind < SVDB_MFMA_FP64 || kind > SVDB_MFMA_FP8) {
    svdb::set_error(SVDB_EINVAL, "svdb_mfma_sustained: kind or output");
    return SVDB_EINVAL;
  }
  if (svdb::device_init() < 0) return SVDB_EDEVICE;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0) {
    svdb::set_error(SVDB_EDEVICE, "svdb_mfma_sustained: device attributes");
    return SVDB_EDEVICE;
  }
  // a few milliseconds a launch at the rates these loops reach (shorter
  // launches read low while the clocks ramp)
  const int iters = kind == SVDB_MFMA_FP64 ? 4096 : 16384;
  hipStream_t st = nullptr;
  float* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const uint32_t max_blocks = (uint32_t)cus * 4;
  int rc = SVDB_OK;
  double best = 0.0;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&out, (size_t)max_blocks * 256 * sizeof(float)) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    svdb::set_error(SVDB_EDEVICE, "svdb_mfma_sustained: setup");
    rc = SVDB_EDEVICE;
  }
  for (uint32_t waves = 2; rc == SVDB_OK && waves <= 4; waves += 2) {
    const uint32_t blocks = (uint32_t)cus * waves;  // 4 wavefronts a block, one per SIMD
    if (svdb::launch_mfma_loop(kind, blocks, iters, out, st) != hipSuccess) rc = SVDB_EDEVICE;
    for (int rep = 0; rc == SVDB_OK && rep < 3; rep++) {
      float ms = 0.f;
      if (hipEventRecord(e0, st) != hipSuccess ||
          svdb::launch_mfma_loop(kind, blocks, iters, out, st) != hipSuccess ||
          hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.f) {
        rc = SVDB_EDEVICE;
        break;
      }
      best = std::max(best, svdb::mfma_loop_flop(kind, blocks, iters) / (ms * 1e-3) / 1e12);
    }
  }
  if (rc != SVDB_OK && svdb::last_error()[0] == 0)
    svdb::set_error(SVDB_EDEVICE, "svdb_mfma_sustained: launch");
  if (st) (void)hipStreamSynchronize(st);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (out) (void)hipFree(out);
  if (st) (void)hipStreamDestroy(st);
  *tflops = best;
  return rc;
}

int svdb_device_sync(void) {
  if (svdb::device_init() < 0) return SVDB_EDEVICE;
  return hipDeviceSynchronize() == hipSuccess ? SVDB_OK : SVDB_EDEVICE;
}

int svdb_search(VectorDatabase* db, const double* queries, size_t nq, size_t dim, const size_t* k,
                const int* metric, int precision, size_t out_stride, size_t* ids_out,
                double* scores_out) {
  Engine* e = eng_of(db);
  if (!e || (!queries && nq) || !k || !metric || !ids_out || !scores_out) {
    svdb::set_error(SVDB_EINVAL, "NULL argument");
    return SVDB_EINVAL;
  }
  for (size_t i = 0; i < nq; i++)
    if (!check_metric_int(metric[i])) {
      svdb::set_error(SVDB_EINVAL, "unknown metric");
      return SVDB_EINVAL;
    }
  // dim below the store's: Euclidean over the first dim coordinates (the
  // kd-tree's projection); the engine refuses it for cosine and dot
  if (e->dim() != 0 && (dim == 0 || dim > e->dim())) {
    svdb::set_error(SVDB_EINVAL, "query dimension exceeds the store's");
    return SVDB_EINVAL;
  }
  return c_call([&] {
    return e->search_host(queries, nq, dim, dim, k, metric, precision, out_stride, ids_out,
                          scores_out);
  });
}

int svdb_search_device(VectorDatabase* db, const double* d_queries, size_t nq, size_t ldq, size_t k,
                       int metric, int precision, uint64_t* d_ids, double* d_scores) {
  Engine* e = eng_of(db);
  if (!e || !d_queries || !d_ids || !d_scores) {
    svdb::set_error(SVDB_EINVAL, "NULL argument");
    return SVDB_EINVAL;
  }
  return c_call([&] {
    return e->search_device(d_queries, nq, ldq, e->dim(), k, metric, precision, d_ids, d_scores);
  });
}

int svdb_pair_metrics(VectorDatabase* db, const size_t* index_a, const size_t* index_b,
                      size_t npairs, int metric, double* out) {
  Engine* e = eng_of(db);
  if (!e || ((!index_a || !index_b || !out) && npairs)) {
    svdb::set_error(SVDB_EINVAL, "NULL argument");
    return SVDB_EINVAL;
  }
  return c_call([&] { return e->pairs_f64(index_a, index_b, npairs, metric, out); });
}

int svdb_bulk_insert(VectorDatabase* db, const double* rows, size_t n, size_t dim,
                     const char* uuids, size_t* first_index) {
  Engine* e = eng_of(db);
  if (!e || (!rows && n)) {
    svdb::set_error(SVDB_EINVAL, "NULL argument");
    return SVDB_EINVAL;
  }
  pthread_mutex_lock(&db->mutex);
  if (grow_mirror(db, db->size + n) != 0) {
    pthread_mutex_unlock(&db->mutex);
    svdb::set_error(SVDB_ENOMEM, "host mirror allocation failed");
    return SVDB_ENOMEM;
  }
  // Host copies first, so a failure leaves the store unchanged.
  std::vector<double*> copies;
  std::vector<const char*> us;
  try {
    copies.assign(n, nullptr);
    us.assign(uuids ? n : 0, nullptr);
  } catch (...) {
    pthread_mutex_unlock(&db->mutex);
    svdb::set_error(SVDB_ENOMEM, "host allocation failed");
    return SVDB_ENOMEM;
  }
  for (size_t i = 0; i < n; i++) {
    copies[i] = (double*)malloc(dim * sizeof(double));
    if (!copies[i]) {
      for (size_t j = 0; j < i; j++) free(copies[j]);
      pthread_mutex_unlock(&db->mutex);
      svdb::set_error(SVDB_ENOMEM, "host mirror allocation failed");
      return SVDB_ENOMEM;
    }
    memcpy(copies[i], rows + i * dim, dim * sizeof(double));
  }
  int rc = c_call([&] { return e->append(rows, n, dim, dim, nullptr); });
  if (rc != SVDB_OK) {
    for (double* p : copies) free(p);
    pthread_mutex_unlock(&db->mutex);
    return rc;
  }
  if (first_index) *first_index = db->size;
  for (size_t i = 0; i < n; i++) {
    Vector& v = db->vectors[db->size + i];
    memset(v.uuid, 0, UUID_SIZE);
    if (uuids) memcpy(v.uuid, uuids + i * UUID_SIZE, UUID_SIZE);
    v.dimension = dim;
    v.data = copies[i];
    if (uuids) us[i] = v.uuid;
  }
  uuids_of(db)->append(n, uuids ? us.data() : nullptr, UUID_SIZE);
  db->size += n;
  pthread_mutex_unlock(&db->mutex);
  return SVDB_OK;
}

int svdb_set_shortlist(VectorDatabase* db, size_t width) {
  Engine* e = eng_of(db);
  if (!e) return SVDB_EINVAL;
  return e->set_shortlist(width);
}

int svdb_set_budget_bytes(VectorDatabase* db, size_t bytes) {
  Engine* e = eng_of(db);
  if (!e) return SVDB_EINVAL;
  return e->set_budget(bytes);
}

int svdb_enable_shadow(VectorDatabase* db, int precision) {
  Engine* e = eng_of(db);
  if (!e) return SVDB_EINVAL;
  return c_call([&] { return e->enable_shadow(precision); });
}

int svdb_get_stats(VectorDatabase* db, svdb_stats* out) {
  Engine* e = eng_of(db);
  if (!e || !out) return SVDB_EINVAL;
  svdb::EngineStats s = e->stats();
  out->rows = e->rows();
  out->dim = e->dim();
  out->phys_rows = e->phys_rows();
  out->device_bytes = e->device_bytes();
  out->budget_bytes = e->budget();
  out->searches = s.searches;
  out->queries = s.queries;
  out->certified = s.certified;
  out->fallbacks = s.fallbacks;
  out->last_max_rel_err = s.last_max_rel_err;
  out->widened = s.widened;
  out->second_pass = s.second_pass;
  out->compactions = s.compactions;
  out->graph_launches = s.graph_launches;
  out->graph_captures = s.graph_captures;
  return SVDB_OK;
}

int svdb_set_profiling(VectorDatabase* db, int on) {
  Engine* e = eng_of(db);
  if (!e) return SVDB_EINVAL;
  return c_call([&] { return e->set_profiling(on != 0); });
}

int svdb_get_kernel_times(VectorDatabase* db, svdb_kernel_time* out) {
  Engine* e = eng_of(db);
  if (!e || !out) return SVDB_EINVAL;
  svdb::KernelTime kt[SVDB_K_COUNT];
  e->kernel_times(kt);
  for (int i = 0; i < SVDB_K_COUNT; i++) {
    out[i].launches = kt[i].launches;
    out[i].total_ms = kt[i].total_ms;
  }
  return SVDB_OK;
}

}  // extern "C"
