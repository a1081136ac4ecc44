// host_common.h — host-side helpers shared by the engine's C++ files.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/boosters_amd.h"
#include "engine.h"

namespace bst {

struct Error {
  int code;
  std::string msg;
};

[[noreturn]] inline void raise(int code, const std::string& msg) { throw Error{code, msg}; }

// Thread-local message returned by bst_last_error() (defined in api.cpp).
std::string& last_error_ref();

// Runs one C-ABI entry point: exceptions become status codes and the message
// of bst_last_error(); nothing throws across the boundary.
template <class Fn>
int guard(Fn&& fn) {
  try {
    return fn();
  } catch (const Error& e) {
    last_error_ref() = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    last_error_ref() = "host allocation failed";
    return BST_E_OOM;
  } catch (const std::exception& e) {
    last_error_ref() = e.what();
    return BST_E_INVALID;
  } catch (...) {
    last_error_ref() = "unknown error";
    return BST_E_INVALID;
  }
}

inline void need(bool ok, int code, const char* msg) {
  if (!ok) raise(code, msg);
}

#define BST_HIP(x)                                                                   \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess)                                                            \
      ::bst::raise(BST_E_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_));    \
  } while (0)

// Cache of one context's device and pinned host blocks, reused by size
// between training runs: a run of the same shape then allocates and frees
// nothing (hipFree and hipHostFree synchronise and cost tens of microseconds
// each).  Blocks are only handed out inside a PoolScope on the context's
// thread, and all training work runs on the context's one stream, so a
// reused block's earlier work is ordered before its new work.  A scope ends
// by freeing the blocks the run did not reuse, so the cache holds at most
// one run's working set.
struct BufPool {
  std::multimap<size_t, void*> dev, host;              // free, returned by this run
  std::multimap<size_t, void*> dev_prev, host_prev;    // free, left by the previous run
  static void* take(std::multimap<size_t, void*>& a, std::multimap<size_t, void*>& b,
                    size_t bytes) {
    for (auto* m : {&a, &b}) {
      auto it = m->find(bytes);
      if (it != m->end()) {
        void* q = it->second;
        m->erase(it);
        return q;
      }
    }
    return nullptr;
  }
  void* take_dev(size_t bytes) { return take(dev, dev_prev, bytes); }
  void* take_host(size_t bytes) { return take(host, host_prev, bytes); }
  void put_dev(void* q, size_t bytes) { dev.emplace(bytes, q); }
  void put_host(void* q, size_t bytes) { host.emplace(bytes, q); }
  static void free_all(std::multimap<size_t, void*>& m, bool pinned) {
    for (auto& kv : m) (void)(pinned ? hipHostFree(kv.second) : hipFree(kv.second));
    m.clear();
  }
  // Start of a run: what is free now is the previous run's.
  void begin_run() {
    dev_prev.merge(dev);
    host_prev.merge(host);
  }
  // End of a run: free the previous run's blocks this run did not take.
  void end_run() {
    free_all(dev_prev, false);
    free_all(host_prev, true);
  }
  void release() {
    end_run();
    free_all(dev, false);
    free_all(host, true);
  }
  ~BufPool() { release(); }
};

// The pool that device buffers allocated on this thread come from (null:
// plain hipMalloc / hipFree).
inline thread_local BufPool* tls_pool = nullptr;

struct PoolScope {
  BufPool* prev;
  explicit PoolScope(BufPool* p) : prev(tls_pool) {
    p->begin_run();
    tls_pool = p;
  }
  ~PoolScope() {
    tls_pool->end_run();
    tls_pool = prev;
  }
  PoolScope(const PoolScope&) = delete;
  PoolScope& operator=(const PoolScope&) = delete;
};

// Device buffer owned by the engine.  Allocated inside a PoolScope, its block
// comes from (and returns to) that pool.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  BufPool* pool = nullptr;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), pool(o.pool) {
    o.p = nullptr;
    o.n = 0;
    o.pool = nullptr;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      reset();
      p = o.p;
      n = o.n;
      pool = o.pool;
      o.p = nullptr;
      o.n = 0;
      o.pool = nullptr;
    }
    return *this;
  }
  ~DevBuf() { reset(); }
  void alloc(size_t count) {
    reset();
    if (count == 0) return;
    pool = tls_pool;
    if (pool) {
      if (void* q = pool->take_dev(count * sizeof(T))) {
        p = (T*)q;
        n = count;
        return;
      }
    }
    hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      pool = nullptr;
      raise(BST_E_OOM, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                           " bytes failed: " + hipGetErrorString(e));
    }
    n = count;
  }
  // Grow-only reallocation (contents are not preserved).
  void ensure(size_t count) {
    if (count > n) alloc(count);
  }
  void reset() {
    if (p) {
      if (pool)
        pool->put_dev(p, n * sizeof(T));
      else
        (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
    pool = nullptr;
  }
  size_t bytes() const { return n * sizeof(T); }
};

template <class T>
void h2d(hipStream_t s, T* dst, const T* src, size_t count) {
  if (count) BST_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, s));
}
template <class T>
void d2h(hipStream_t s, T* dst, const T* src, size_t count) {
  if (count) BST_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost, s));
}

// A numeric or categorical bin mapper (BinMapper, data/binned/bin_mapper.rs:55-90).
struct Mapper {
  bool categorical = false;
  bool trivial = false;
  std::vector<double> bounds;  // numeric upper bounds incl. f64::MAX sentinels
  std::vector<int32_t> cats;   // categorical: sorted category values
  uint32_t n_bins = 0;
  int missing = 0;  // 0 None, 2 NaN
  uint32_t default_bin = 0;
  int64_t n_nonzero = 0;  // values != 0 and not NaN (the density of sparsity detection)
};

// One Exclusive Feature Bundle: sparse categorical features sharing one
// column whose code is offsets[j] + bin for a row where feature j (the first
// in bundle order that is non-zero there) is set, 0 where none is.
struct Bundle {
  std::vector<uint32_t> features;  // original feature indices, bundle order
  std::vector<uint32_t> offsets;   // code of bin 0 of each feature (the first is 1)
  uint32_t total_bins = 0;         // 1 + the features' bins
};

// The bundles the reference builds for a [F][n] host matrix
// (create_bundle_plan): empty when fewer than two sparse categorical
// features, more than 1000, or too many conflicting pairs.
std::vector<Bundle> plan_bundles(const float* X, int64_t n, int32_t F,
                                 const std::vector<Mapper>& maps, const bst_binning_config& cfg);
// A bundle's codes for every row (total_bins <= 256).
void encode_bundle(const Bundle& b, const float* X, int64_t n, const std::vector<Mapper>& maps,
                   uint8_t* out);
// BinMapper::value_to_bin of a categorical feature for a non-NaN value.
uint32_t cat_bin(const Mapper& m, float v);

// Build bin mappers for every feature of a [F][n] host matrix
// (BinnedDataset::from_dataset, data/binned/dataset.rs:207-309).
std::vector<Mapper> build_mappers(const float* X, int64_t n, int32_t F, const uint8_t* is_cat,
                                  const bst_binning_config& cfg);

float next_up_f32(float x);
float round_down_f32(double b);

}  // namespace bst

namespace bst {
// Histogram launch shape of one training run.
struct HistTuning {
  int fg = 32;                 // features per histogram block: 32 (one block per CU) or 16 (two)
  uint32_t tile_min = 1024;    // fewest rows per block in a step (power of two, 512..8192)
  uint32_t blocks_per_cu = 2;  // blocks a step's rows are spread over, per CU
  int priv = 1;                // sub-histograms: 1 in LDS, 2 in global memory (device atomics)
  int tuned = 0;               // 1 = measured on this data shape
  double tune_ms = 0;          // time spent measuring (0 when fixed or cached)
};
}  // namespace bst

struct bst_ctx {
  int device = 0;
  int n_cu = 256;  // compute units (sizes grids of resident, grid-stride kernels)
  size_t lds_per_block = 65536;  // LDS a workgroup may use (hipDeviceProp sharedMemPerBlock)
  // Device address of this device's static predict scratch (the
  // allocation-free entry point's; looked up at creation).
  float* noalloc_scratch = nullptr;
  hipStream_t stream = nullptr;
  double last_ms_total = 0, last_ms_per_tree = 0;
  int64_t last_hist_rows = 0;
  std::vector<double> last_metric;
  // Training predictions [K][n] after the last round of the last train, kept
  // only when keep_preds is set (bst_ctx_set_keep_predictions).
  bool keep_preds = false;
  std::vector<float> last_preds;
  // Kernel profiling: events bracketing launches on `stream`.
  struct Pending {
    int kind;
    hipEvent_t a, b;
    double bytes;
  };
  bool profile = false;
  std::vector<Pending> pending;
  // Grow-only pinned host buffer for device -> host staging.
  void* staging = nullptr;
  size_t staging_bytes = 0;
  void* pinned_staging(size_t bytes) {
    if (bytes > staging_bytes) {
      if (staging) {
        BST_HIP(hipStreamSynchronize(stream));
        (void)hipHostFree(staging);
        staging = nullptr;
        staging_bytes = 0;
      }
      BST_HIP(hipHostMalloc(&staging, bytes, hipHostMallocDefault));
      staging_bytes = bytes;
    }
    return staging;
  }
  // Per-tree leaf values of two-phase predictions (grow-only, freed with the ctx).
  bst::DevBuf<float> pred_leaf;
  // Autotuner results per data shape, and the shape the last train used.
  std::map<std::vector<int64_t>, bst::HistTuning> tune_cache;
  // Device and pinned blocks of the last training run, reused by the next.
  bst::BufPool pool;
  bst::HistTuning last_tuning;
  int64_t k_launches[3] = {0, 0, 0};
  double k_ms[3] = {0, 0, 0}, k_bytes[3] = {0, 0, 0};
  // Record the start of a launch of `kind`; returns an index for prof_end.
  int prof_begin(int kind, double bytes) {
    if (!profile) return -1;
    Pending p{kind, nullptr, nullptr, bytes};
    BST_HIP(hipEventCreate(&p.a));
    BST_HIP(hipEventCreate(&p.b));
    BST_HIP(hipEventRecord(p.a, stream));
    pending.push_back(p);
    return (int)pending.size() - 1;
  }
  void prof_end(int i) {
    if (i >= 0) BST_HIP(hipEventRecord(pending[i].b, stream));
  }
  // After a stream synchronisation: fold the finished launches in.
  void prof_drain() {
    for (auto& p : pending) {
      float ms = 0.0f;
      BST_HIP(hipEventElapsedTime(&ms, p.a, p.b));
      k_launches[p.kind]++;
      k_ms[p.kind] += ms;
      k_bytes[p.kind] += p.bytes;
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    pending.clear();
  }
};

struct bst_dataset {
  bst_ctx* ctx = nullptr;
  std::vector<bst::Mapper> mappers;      // per original feature
  std::vector<uint32_t> eff_to_orig;     // effective column -> original feature (a
                                         // bundle's: its first feature)
  std::vector<bst::Bundle> bundles;      // effective columns 0 .. bundles.size() - 1
  std::vector<bst::FeatMeta> fmeta;      // per effective column (host copy)
  std::vector<float> thr_up;             // [E][256]: next_up((f32)bound[bin])
  bst::DatasetDev d;
  bst::DevBuf<uint8_t> bins_rm, bins_cm;
  bst::DevBuf<uint32_t> hi_cm;
  bst::DevBuf<bst::FeatMeta> meta;
  bst::DevBuf<float> raw;
  double ms_binning = 0;
};

struct bst_model {
  bst_model_info info{};
  std::vector<float> base_score;
  std::vector<int64_t> tree_offset;
  std::vector<uint32_t> tree_group, split_index, left, right, cat_start, cat_size, cat_words;
  std::vector<float> split_threshold, leaf_value, gain, cover;
  std::vector<uint8_t> default_left, is_leaf, split_type;
  // Model metadata carried through .bstr files (persist/schema.rs:27-44):
  // meta.num_groups (-1: same as info.n_groups), optional feature names,
  // feature types (0 numeric, 1 categorical) and objective name.
  int64_t meta_num_groups = -1;
  bool has_feature_names = false, has_feature_types = false, has_objective_name = false;
  std::vector<std::string> feature_names;
  std::vector<uint8_t> feature_types;
  std::string objective_name;
  // Device copy for traversal (built on first predict).
  int device_for = -1;
  bst::DevBuf<bst::DevNode> d_nodes;
  bst::DevBuf<int64_t> d_tree_off;
  bst::DevBuf<uint32_t> d_tree_group, d_cat_start, d_cat_size, d_cat_words;
  bst::DevBuf<float> d_base;
  // Predictor layout (engine.h): compact nodes of the LDS chunks' trees, their
  // offsets, levels and original ids, the categorical references and the
  // chunk table.
  bst::DevBuf<uint2> d_cnodes;
  bst::DevBuf<uint32_t> d_ctree_off, d_cdepth, d_corig, d_cat_ref;
  bst::DevBuf<bst::PredChunk> d_chunks;
  int n_chunks = 0;
  int pred_has_cat = 0;
  void upload(hipStream_t s);
  // The device copy as the predictor's view (after upload).
  bst::PredModel view() const {
    bst::PredModel v{};
    v.nodes = d_nodes.p;
    v.tree_off = d_tree_off.p;
    v.tree_group = d_tree_group.p;
    v.cat_start = d_cat_start.p;
    v.cat_size = d_cat_size.p;
    v.cat_words = d_cat_words.p;
    v.cnodes = d_cnodes.p;
    v.ctree_off = d_ctree_off.p;
    v.cdepth = d_cdepth.p;
    v.corig = d_corig.p;
    v.cat_ref = d_cat_ref.p;
    v.chunks = d_chunks.p;
    v.n_chunks = n_chunks;
    v.n_features = (int)info.n_features;
    v.has_cat = pred_has_cat;
    v.n_trees = info.n_trees;
    v.base = d_base.p;
    v.K = (int)info.n_groups;
    v.g0 = 0;
    v.Kw = v.K < bst::kPredMaxGroups ? v.K : bst::kPredMaxGroups;
    return v;
  }
};

namespace bst {
// The histogram launch shape for training `ds` with `cfg`/`ext` and T trees
// grown together: fixed fields of `ext` are kept; with ext.autotune the free
// ones are measured on this data (cached in ctx per shape).
HistTuning choose_hist_tuning(bst_ctx* ctx, const DatasetDev& d, const bst_gbdt_config& cfg,
                              const bst_ext_config& ext, int T);

// Structural check of a forest before it is traversed on the device: array
// lengths, groups, and per tree an acyclic walk from the root in which every
// split node has in-range, distinct, unvisited children (the rules of
// repr/gbdt/tree_view.rs:216-290), split features < n_features and < 2^24,
// and category bitsets inside cat_words.  Raises BST_E_FORMAT.
void check_forest(const bst_model& m);
}  // namespace bst
