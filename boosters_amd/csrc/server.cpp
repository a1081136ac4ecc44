// server.cpp — device-resident prediction serving: continuous batching into
// one persistent traversal launch.
//
// No counterpart in the reference (its predict path is GBDTModel::predict,
// model/gbdt/model.rs:217-258, one call per batch).  Requests of any row count
// are queued by bst_server_submit (thread-safe).  A worker thread coalesces
// queued requests into batches, one per free slot of a ring of kServeRing
// slots, packs each batch's rows side by side into the slot's pinned input
// arena ([F][rows]), and publishes it to the serving kernel (k_serve): one
// block per CU stays resident and takes the published batches' work units
// in order, so batches that arrive while others are on the GPU are traversed
// alongside them.  A unit is a 256-row tile traversed through the whole
// forest; a batch too small to occupy the device is instead cut into leaf
// units (its tiles times ranges of the forest, leaf values to the slot's
// leaf buffer in HBM) and one sum unit per tile, which adds each row's
// leaves in forest order once the leaf units are done, so that a small
// request is spread over many CUs.  The kernel reads the rows and writes the
// raw margins directly in pinned coherent host memory; the block that
// finishes a batch's last unit publishes the batch's sequence number, and
// the worker copies the outputs to the requests and frees the slot.
//
// Admission: the slot arenas and leaf buffers are sized at creation from the
// caller's budget, capped by the device's free HBM (hipMemGetInfo), and
// allocated then; a request is admitted when a slot is free and its rows fit
// the slot beside the requests already packed there.  The kernel leaves after 50 ms without
// work (or on destroy) and the worker relaunches it when a batch is published
// and no launch is live.  Every row is traversed by the prediction kernel's
// own tile code with the same per-row summation order, so a result is
// bit-identical whether its request ran alone or packed with others, and
// equal to bst_predict_device's.
#include <hip/hip_runtime.h>
#include <sys/prctl.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "host_common.h"

using namespace bst;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t ticket = 0;
  const float* X = nullptr;
  int64_t n = 0;
  int32_t F = 0;
  float* out = nullptr;
  Clock::time_point t0;
  bool done = false;
  bool claimed = false;  // a bst_server_wait holds this ticket
  int status = BST_OK;
  int64_t latency_ns = 0;
};

struct Slot {
  std::vector<Request*> reqs;
  int64_t rows = 0;
  uint32_t seq = 0;  // of the batch in flight
  bool busy = false;
};

// Run fn(i) for i in [0, n) on up to `threads` threads (the caller's included).
template <class Fn>
void parallel_for(int n, int threads, Fn&& fn) {
  threads = std::max(1, std::min(threads, n));
  if (threads == 1) {
    for (int i = 0; i < n; i++) fn(i);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve((size_t)threads - 1);
  auto part = [&](int t) {
    for (int i = t; i < n; i += threads) fn(i);
  };
  for (int t = 1; t < threads; t++) pool.emplace_back(part, t);
  part(0);
  for (auto& th : pool) th.join();
}

// Threads for packing a batch: packs below 8 MB are copied by the worker alone.
int pack_threads(size_t bytes) {
  if (bytes < ((size_t)8 << 20)) return 1;
  const unsigned hw = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(8u, hw / 2));
}

constexpr int64_t kMaxSlotRows = 1 << 20;
constexpr double kIdleMs = 50.0;
// Host-link bytes that the leaf units of one batch may read in all.
constexpr int64_t kSplitHostBytes = (int64_t)2 << 20;
// Batches of at least this many input bytes are packed on a thread of their
// own, so that smaller batches admitted meanwhile are not held behind the copy.
constexpr size_t kAsyncPackBytes = (size_t)4 << 20;

}  // namespace

struct bst_server {
  bst_ctx* ctx = nullptr;
  const bst_model* model = nullptr;
  int64_t budget = 0;      // admission budget after the free-HBM cap
  int32_t F = 0, K = 1;
  int tile_rows = 256;
  int64_t slot_rows = 0;   // rows one batch holds
  // Batches of at most split_rows rows whose tiles leave most CUs idle are
  // cut into leaf and sum units; leaf holds [n_trees][split_rows] per slot.
  int64_t split_rows = 0;
  int n_chunks = 0, n_cu = 1;
  int64_t n_trees = 0;
  DevBuf<float> leaf;
  std::vector<Slot> slots;
  // Pinned coherent host memory: state, input and output arenas, and their
  // device addresses.
  ServeHostState* hs = nullptr;
  ServeHostState* hs_dev = nullptr;
  float* in_arena = nullptr;
  float* in_dev = nullptr;
  float* out_arena = nullptr;
  float* out_dev = nullptr;
  DevBuf<ServeDevState> dstate;
  hipStream_t stream = nullptr;
  hipEvent_t kernel_done = nullptr;
  bool launched = false;
  uint64_t idle_ticks = 0;
  uint64_t next_unit[2] = {0, 0};        // units published so far per lane
  uint64_t cum_done[kServeRing] = {};    // tiles_done[slot] once its batches finish
  uint32_t next_seq = 1;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::unordered_map<uint64_t, Request*> reqs;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread worker;
  std::string worker_error;
  // Background packing of large batches: the slot's thread, and its state
  // (0 idle, 1 packing, 2 packed, 3 failed).
  std::thread packer[kServeRing];
  std::atomic<int> pack_state[kServeRing] = {};
};

namespace {

// The persistent kernel is live, or (re)launched now.
void ensure_kernel(bst_server* srv) {
  if (srv->launched) {
    const hipError_t q = hipEventQuery(srv->kernel_done);
    if (q == hipErrorNotReady) return;
    if (q != hipSuccess) raise(BST_E_DEVICE, std::string("serving kernel: ") + hipGetErrorString(q));
  }
  const bst_model* m = srv->model;
  launch_serve(srv->stream, m->view(), srv->hs_dev, srv->dstate.p, srv->in_dev, srv->out_dev,
               srv->leaf.p, srv->idle_ticks, srv->ctx->lds_per_block, srv->ctx->n_cu);
  BST_HIP(hipGetLastError());
  BST_HIP(hipEventRecord(srv->kernel_done, srv->stream));
  srv->launched = true;
}

// Pack a slot's requests side by side into its input arena ([F][rows]).
void pack(bst_server* srv, int si) {
  const Slot& s = srv->slots[si];
  const int32_t F = srv->F;
  const int64_t rows = s.rows;
  float* base = srv->in_arena + (size_t)si * srv->slot_rows * F;
  std::vector<int64_t> offs(s.reqs.size());
  int64_t off = 0;
  for (size_t i = 0; i < s.reqs.size(); i++) {
    offs[i] = off;
    off += s.reqs[i]->n;
  }
  parallel_for(F, pack_threads((size_t)rows * F * 4), [&](int f) {
    for (size_t i = 0; i < s.reqs.size(); i++) {
      const Request* r = s.reqs[i];
      std::memcpy(base + (size_t)f * rows + offs[i], r->X + (size_t)f * r->n, (size_t)r->n * 4);
    }
  });
}

// Describe a packed slot's batch and publish its units.
void publish(bst_server* srv, int si) {
  Slot& s = srv->slots[si];
  const int32_t F = srv->F;
  const int64_t rows = s.rows;
  const uint32_t n_tiles = (uint32_t)((rows + srv->tile_rows - 1) / srv->tile_rows);
  // Leaf units per tile: enough to spread the batch over the CUs, when its
  // tiles alone would leave more than half of them idle.  Each leaf unit
  // reads its tile's rows across the host link, so the units of a batch
  // read at most kSplitHostBytes in all (a batch of more rows takes fewer
  // units per tile, or none).
  uint32_t split = 0;
  const int64_t in_bytes = rows * F * 4;
  const int64_t by_bytes = kSplitHostBytes / std::max<int64_t>(in_bytes, 1);
  if (rows <= srv->split_rows && srv->n_chunks > 1 && 2 * (int64_t)n_tiles <= srv->n_cu &&
      by_bytes >= 2) {
    split = (uint32_t)std::min<int64_t>(
        {(int64_t)kServeMaxSplit, (int64_t)srv->n_chunks, by_bytes,
         std::max<int64_t>(2, srv->n_cu / (int64_t)n_tiles)});
    // Ranges of `per` chunks: only as many as cover the forest, so that no
    // range starts past its last chunk.
    const uint32_t per = ((uint32_t)srv->n_chunks + split - 1) / split;
    split = ((uint32_t)srv->n_chunks + per - 1) / per;
  }
  const uint32_t n_units = split ? n_tiles * (split + 1) : n_tiles;
  ServeDesc& d = srv->hs->desc[si];
  d.leaf_target = srv->cum_done[si] + (uint64_t)n_tiles * split;
  srv->cum_done[si] += n_units;
  s.seq = srv->next_seq++;
  d.in_off = (uint64_t)si * srv->slot_rows * F;
  d.out_off = (uint64_t)si * srv->slot_rows * srv->K;
  d.leaf_off = (uint64_t)si * (uint64_t)srv->n_trees * (uint64_t)srv->split_rows;
  d.done_target = srv->cum_done[si];
  d.rows = (uint32_t)rows;
  d.seq = s.seq;
  d.split = split;
  d.tiles = n_tiles;
  // Batches cut into units go to lane 0, which blocks look at first.
  const uint64_t lane = split ? 0u : 1u;
  __atomic_store_n(&d.range, (lane << 63) | (srv->next_unit[lane] << 16) | n_units,
                   __ATOMIC_RELAXED);
  srv->next_unit[lane] += n_units;
  // The rows and the descriptor are written before the units are published.
  __atomic_store_n(&srv->hs->published[lane], srv->next_unit[lane], __ATOMIC_RELEASE);
}

// Copy a finished slot's outputs to its requests (outside the lock).
void unpack(bst_server* srv, int si) {
  const Slot& s = srv->slots[si];
  const int32_t K = srv->K;
  const float* base = srv->out_arena + (size_t)si * srv->slot_rows * K;
  int64_t off = 0;
  for (Request* r : s.reqs) {
    for (int32_t k = 0; k < K; k++)
      std::memcpy(r->out + (size_t)k * r->n, base + (size_t)k * s.rows + off, (size_t)r->n * 4);
    off += r->n;
  }
}

// The worker stopped on an error: every queued and admitted request fails
// with its code, and later submits are refused with its message.
void fail_pending(bst_server* srv, int code, const std::string& msg) {
  std::lock_guard<std::mutex> lk(srv->mu);
  srv->worker_error = msg;
  for (auto* r : srv->queue) {
    r->status = code;
    r->done = true;
  }
  srv->queue.clear();
  for (auto& s : srv->slots)
    for (auto* r : s.reqs) {
      r->status = code;
      r->done = true;
    }
  srv->cv_done.notify_all();
}

// Ask the kernel to leave and wait for it (bounded: it polls the flag).
void stop_kernel(bst_server* srv) {
  __atomic_store_n(&srv->hs->stop, 1u, __ATOMIC_RELEASE);
  if (srv->launched) (void)hipEventSynchronize(srv->kernel_done);
  srv->launched = false;
}

void worker_loop(bst_server* srv) {
  // The loop sleeps 10 us between polls while batches are in flight; with
  // Linux's default 50 us timer slack each sleep would last about 60 us,
  // which every request would wait out once more at its completion.
  (void)prctl(PR_SET_TIMERSLACK, 1000UL, 0UL, 0UL, 0UL);
  try {
    BST_HIP(hipSetDevice(srv->ctx->device));
    std::vector<int> finished, admitted;
    for (;;) {
      finished.clear();
      admitted.clear();
      // Batches whose last unit is done (their seq published by the kernel);
      // a slot still being packed holds its previous batch's seq.
      for (int si = 0; si < (int)srv->slots.size(); si++) {
        const Slot& s = srv->slots[si];
        if (s.busy && srv->pack_state[si].load(std::memory_order_acquire) == 0 &&
            __atomic_load_n(&srv->hs->done_seq[si], __ATOMIC_ACQUIRE) == s.seq)
          finished.push_back(si);
      }
      if (__atomic_load_n(&srv->hs->error, __ATOMIC_ACQUIRE))
        raise(BST_E_DEVICE,
              "serving kernel found a unit without a batch, or a batch's leaf units did not "
              "finish within the wait bound");
      for (int si : finished) unpack(srv, si);
      bool any_busy = false;
      {
        std::unique_lock<std::mutex> lk(srv->mu);
        const auto now = Clock::now();
        for (int si : finished) {
          Slot& s = srv->slots[si];
          for (Request* r : s.reqs) {
            r->latency_ns = std::chrono::duration_cast<std::chrono::nanoseconds>(now - r->t0).count();
            r->done = true;
          }
          s.reqs.clear();
          s.rows = 0;
          s.busy = false;
        }
        // Admit queued requests into free slots, coalescing while they fit.
        for (int si = 0; si < (int)srv->slots.size() && !srv->queue.empty(); si++) {
          Slot& s = srv->slots[si];
          if (s.busy) continue;
          while (!srv->queue.empty() && s.rows + srv->queue.front()->n <= srv->slot_rows) {
            Request* r = srv->queue.front();
            srv->queue.pop_front();
            s.reqs.push_back(r);
            s.rows += r->n;
          }
          if (!s.reqs.empty()) {
            s.busy = true;
            admitted.push_back(si);
          }
        }
        for (auto& s : srv->slots) any_busy = any_busy || s.busy;
        if (srv->stop && srv->queue.empty() && !any_busy) {
          lk.unlock();
          if (!finished.empty()) srv->cv_done.notify_all();
          stop_kernel(srv);
          return;
        }
        if (finished.empty() && admitted.empty() && !any_busy)
          srv->cv_work.wait_for(lk, std::chrono::milliseconds(20),
                                [&] { return srv->stop || !srv->queue.empty(); });
      }
      if (!finished.empty()) srv->cv_done.notify_all();
      // Packing and publishing outside the lock: only this thread (and the
      // slot's packer) touches an admitted slot, and its requests' buffers
      // belong to it until done.  Large batches are packed in the background
      // and published once packed; the others at once.
      for (int si : admitted) {
        const Slot& s = srv->slots[si];
        if ((size_t)s.rows * srv->F * 4 < kAsyncPackBytes) {
          pack(srv, si);
          publish(srv, si);
        } else {
          srv->pack_state[si].store(1, std::memory_order_relaxed);
          srv->packer[si] = std::thread([srv, si] {
            int st = 2;
            try {
              pack(srv, si);
            } catch (...) {
              st = 3;
            }
            srv->pack_state[si].store(st, std::memory_order_release);
          });
        }
      }
      // Packed in the background: publish (which gives the slot its new
      // seq), then mark the slot idle for the finished check.
      for (int si = 0; si < (int)srv->slots.size(); si++) {
        const int st = srv->pack_state[si].load(std::memory_order_acquire);
        if (st < 2) continue;
        srv->packer[si].join();
        if (st == 3) raise(BST_E_OOM, "serving worker: packing a batch failed");
        publish(srv, si);
        srv->pack_state[si].store(0, std::memory_order_release);
        admitted.push_back(si);  // the loop does not idle after a publish
      }
      if (any_busy || !admitted.empty()) ensure_kernel(srv);
      if (finished.empty() && admitted.empty() && any_busy)
        std::this_thread::sleep_for(std::chrono::microseconds(10));
    }
  } catch (const Error& e) {
    fail_pending(srv, e.code, e.msg);
  } catch (const std::bad_alloc&) {
    fail_pending(srv, BST_E_OOM, "serving worker: host allocation failed");
  } catch (const std::exception& e) {
    fail_pending(srv, BST_E_DEVICE, std::string("serving worker: ") + e.what());
  } catch (...) {
    fail_pending(srv, BST_E_DEVICE, "serving worker: unknown failure");
  }
  for (auto& th : srv->packer)
    if (th.joinable()) th.join();
  stop_kernel(srv);
}

template <class Fn>
int sguard(Fn&& fn) {
  try {
    return fn();
  } catch (const Error& e) {
    ::bst::last_error_ref() = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    ::bst::last_error_ref() = e.what();
    return BST_E_INVALID;
  }
}

void release(bst_server* srv) {
  if (srv->stream) (void)hipStreamSynchronize(srv->stream);
  if (srv->kernel_done) (void)hipEventDestroy(srv->kernel_done);
  if (srv->stream) (void)hipStreamDestroy(srv->stream);
  if (srv->hs) (void)hipHostFree(srv->hs);
  if (srv->in_arena) (void)hipHostFree(srv->in_arena);
  if (srv->out_arena) (void)hipHostFree(srv->out_arena);
  for (auto& kv : srv->reqs) delete kv.second;
  delete srv;
}

template <class T>
T* pinned(size_t bytes, T** dev) {
  void* p = nullptr;
  BST_HIP(hipHostMalloc(&p, bytes, hipHostMallocCoherent | hipHostMallocMapped));
  void* d = nullptr;
  const hipError_t e = hipHostGetDevicePointer(&d, p, 0);
  if (e != hipSuccess) {
    (void)hipHostFree(p);
    raise(BST_E_DEVICE, std::string("hipHostGetDevicePointer: ") + hipGetErrorString(e));
  }
  *dev = static_cast<T*>(d);
  return static_cast<T*>(p);
}

}  // namespace

extern "C" {

int bst_server_create(bst_ctx* ctx, const bst_model* model, int64_t budget, int32_t max_inflight,
                      bst_server** out) {
  return sguard([&] {
    if (!ctx || !model || !out) raise(BST_E_INVALID, "null argument");
    if (budget <= 0) raise(BST_E_CONFIG, "hbm budget must be positive");
    if (model->info.n_groups > kPredMaxGroups) raise(BST_E_CONFIG, "too many output groups");
    BST_HIP(hipSetDevice(ctx->device));
    auto* m = const_cast<bst_model*>(model);
    if (m->device_for != ctx->device) {
      m->upload(ctx->stream);
      m->device_for = ctx->device;
    }
    const int32_t F = (int32_t)model->info.n_features, K = (int32_t)model->info.n_groups;
    const int R = serve_tile_rows(F, ctx->lds_per_block);
    if (R == 0) raise(BST_E_CONFIG, "the model tests too many features for the serving kernel's LDS tile");
    auto* srv = new bst_server();
    std::unique_ptr<bst_server, void (*)(bst_server*)> guard_srv(srv, release);
    srv->ctx = ctx;
    srv->model = model;
    srv->F = F;
    srv->K = K;
    srv->tile_rows = R;
    // Budget: the caller's, capped by the device's free HBM now.
    size_t free_b = 0, total_b = 0;
    BST_HIP(hipMemGetInfo(&free_b, &total_b));
    srv->budget = std::min<int64_t>(budget, (int64_t)(free_b / 10 * 9));
    const int n_slots = std::max(1, std::min<int>(max_inflight, kServeRing));
    const int64_t row_bytes = (int64_t)(std::max(F, 1) + K) * 4;
    const bst::PredModel pm = m->view();
    srv->n_chunks = pm.n_chunks;
    srv->n_trees = pm.n_trees;
    srv->n_cu = std::max(1, ctx->n_cu);
    // Leaf buffers (HBM) for batches of up to half the CUs' worth of tiles,
    // within a quarter of the budget; none if that holds less than a tile.
    const int64_t leaf_row_bytes = (int64_t)n_slots * std::max<int64_t>(srv->n_trees, 1) * 4;
    srv->split_rows = std::min<int64_t>((int64_t)(srv->n_cu / 2) * R, srv->budget / 4 / leaf_row_bytes);
    srv->split_rows = srv->split_rows / R * R;
    if (srv->split_rows < R || srv->n_chunks < 2) srv->split_rows = 0;
    const int64_t leaf_bytes = srv->split_rows * leaf_row_bytes;
    srv->slot_rows =
        std::min<int64_t>(kMaxSlotRows, (srv->budget - leaf_bytes) / (n_slots * row_bytes));
    srv->slot_rows = std::min<int64_t>(srv->slot_rows / R * R, (int64_t)kServeMaxTiles * R);
    if (srv->slot_rows < R) raise(BST_E_OOM, "serving budget holds less than one tile per slot");
    srv->split_rows = std::min(srv->split_rows, srv->slot_rows);
    if (srv->split_rows > 0)
      srv->leaf.alloc((size_t)n_slots * (size_t)srv->n_trees * (size_t)srv->split_rows);
    srv->slots.resize(n_slots);
    srv->hs = pinned<ServeHostState>(sizeof(ServeHostState), &srv->hs_dev);
    std::memset(srv->hs, 0, sizeof(ServeHostState));
    srv->in_arena = pinned<float>((size_t)n_slots * srv->slot_rows * std::max(F, 1) * 4, &srv->in_dev);
    srv->out_arena = pinned<float>((size_t)n_slots * srv->slot_rows * K * 4, &srv->out_dev);
    srv->dstate.alloc(1);
    BST_HIP(hipStreamCreateWithFlags(&srv->stream, hipStreamNonBlocking));
    BST_HIP(hipEventCreateWithFlags(&srv->kernel_done, hipEventDisableTiming));
    BST_HIP(hipMemsetAsync(srv->dstate.p, 0, sizeof(ServeDevState), srv->stream));
    BST_HIP(hipStreamSynchronize(srv->stream));
    int rate_khz = 0;
    BST_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device));
    srv->idle_ticks = (uint64_t)(std::max(rate_khz, 1) * kIdleMs);
    srv->worker = std::thread(worker_loop, srv);
    *out = guard_srv.release();
    return BST_OK;
  });
}

int bst_server_submit(bst_server* srv, const float* X, int64_t n, int32_t F, float* out,
                      uint64_t* ticket) {
  return sguard([&] {
    if (!srv || !X || !out || !ticket) raise(BST_E_INVALID, "null argument");
    if (F != srv->F) raise(BST_E_SHAPE, "request feature count differs from the model's");
    if (n <= 0) raise(BST_E_SHAPE, "empty request");
    if (n > srv->slot_rows)
      raise(BST_E_OOM, "request has more rows than a serving batch holds (" +
                           std::to_string(srv->slot_rows) + " under this budget)");
    auto* r = new Request();
    r->X = X;
    r->n = n;
    r->F = F;
    r->out = out;
    r->t0 = Clock::now();
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      if (!srv->worker_error.empty()) {
        delete r;
        raise(BST_E_DEVICE, srv->worker_error);
      }
      r->ticket = srv->next_ticket++;
      srv->reqs[r->ticket] = r;
      srv->queue.push_back(r);
    }
    srv->cv_work.notify_one();
    *ticket = r->ticket;
    return BST_OK;
  });
}

int bst_server_wait(bst_server* srv, uint64_t ticket, int64_t* latency_ns) {
  return sguard([&] {
    if (!srv) raise(BST_E_INVALID, "null server");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(ticket);
    if (it == srv->reqs.end()) raise(BST_E_INVALID, "unknown ticket");
    Request* r = it->second;
    // One waiter per ticket: a second concurrent wait on the same ticket is
    // refused instead of racing the first one's erase and delete.
    if (r->claimed) raise(BST_E_INVALID, "ticket is already being waited on");
    r->claimed = true;
    srv->cv_done.wait(lk, [&] { return r->done; });
    const int st = r->status;
    if (latency_ns) *latency_ns = r->latency_ns;
    srv->reqs.erase(ticket);
    delete r;
    if (st != BST_OK) raise(st, srv->worker_error.empty() ? "request failed" : srv->worker_error);
    return BST_OK;
  });
}

int bst_server_destroy(bst_server* srv) {
  return sguard([&] {
    if (!srv) return BST_OK;
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      srv->stop = true;
    }
    srv->cv_work.notify_all();
    if (srv->worker.joinable()) srv->worker.join();
    (void)hipSetDevice(srv->ctx->device);
    release(srv);
    return BST_OK;
  });
}

}  // extern "C"
