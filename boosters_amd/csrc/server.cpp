// server.cpp — device-resident prediction serving with continuous batching.
//
// No counterpart in the reference (its predict path is GBDTModel::predict,
// model/gbdt/model.rs:217-258, one call per batch).  Requests of any row count
// are queued by bst_server_submit (thread-safe).  A worker thread admits
// queued requests while their device footprint fits the free part of the HBM
// budget, packs them side by side into one [F][rows] batch, and runs one
// traversal launch for the whole batch.  Up to max_inflight batches are on the
// device at once, each on its own stream, so the copies of one batch overlap
// the traversal of another and new requests are admitted as soon as a batch
// completes and frees its budget.  The worker only decides under the server's
// lock (retire, admit, budget); packing requests into the pinned batch (split
// over a few threads when large), the launches and unpacking results run
// outside it, so submit and wait never queue behind a copy.  Every row is
// traversed by the same kernel with the same per-row summation order, so a
// result is bit-identical whether its request ran alone or packed with others.
#include <hip/hip_runtime.h>

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
  int64_t bytes = 0;
  Clock::time_point t0;
  bool done = false;
  bool claimed = false;  // a bst_server_wait holds this ticket
  int status = BST_OK;
  int64_t latency_ns = 0;
};

struct Batch {
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
  float* h_in = nullptr;   // pinned [F][cap]
  float* h_out = nullptr;  // pinned [K][cap]
  int64_t cap = 0;
  DevBuf<float> d_in, d_out;
  std::vector<Request*> reqs;
  int64_t rows = 0, bytes = 0;
  bool busy = false;      // admitted, until its requests are done
  bool retiring = false;  // its results are being unpacked
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

}  // namespace


struct bst_server {
  bst_ctx* ctx = nullptr;
  const bst_model* model = nullptr;
  int64_t budget = 0;
  int64_t used = 0;
  int32_t F = 0, K = 1;
  std::vector<Batch> batches;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::unordered_map<uint64_t, Request*> reqs;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread worker;
  std::string worker_error;
};

namespace {

void ensure_batch(Batch& b, int64_t rows, int32_t F, int32_t K) {
  if (rows <= b.cap) return;
  int64_t cap = std::max<int64_t>(rows, b.cap * 2);
  if (b.h_in) (void)hipHostFree(b.h_in);
  if (b.h_out) (void)hipHostFree(b.h_out);
  b.h_in = b.h_out = nullptr;
  BST_HIP(hipHostMalloc((void**)&b.h_in, (size_t)cap * F * sizeof(float), hipHostMallocDefault));
  BST_HIP(hipHostMalloc((void**)&b.h_out, (size_t)cap * K * sizeof(float), hipHostMallocDefault));
  b.d_in.alloc((size_t)cap * F);
  b.d_out.alloc((size_t)cap * K);
  b.cap = cap;
}

void launch_batch(bst_server* srv, Batch& b) {
  const int32_t F = srv->F, K = srv->K;
  ensure_batch(b, b.rows, F, K);
  // Pack requests side by side in the row dimension: [F][rows], one task per
  // feature row.
  std::vector<int64_t> offs(b.reqs.size());
  int64_t off = 0;
  for (size_t i = 0; i < b.reqs.size(); i++) {
    offs[i] = off;
    off += b.reqs[i]->n;
  }
  parallel_for(F, pack_threads((size_t)b.rows * F * 4), [&](int f) {
    for (size_t i = 0; i < b.reqs.size(); i++) {
      const Request* r = b.reqs[i];
      std::memcpy(b.h_in + (size_t)f * b.rows + offs[i], r->X + (size_t)f * r->n,
                  (size_t)r->n * 4);
    }
  });
  BST_HIP(hipMemcpyAsync(b.d_in.p, b.h_in, (size_t)b.rows * F * 4, hipMemcpyHostToDevice, b.stream));
  const PredModel pm = srv->model->view();
  launch_predict(b.stream, pm, b.d_in.p, b.rows, b.rows, F, 0, b.d_out.p, b.rows, nullptr,
                 nullptr, srv->ctx->lds_per_block, srv->ctx->n_cu);
  BST_HIP(hipGetLastError());
  BST_HIP(hipMemcpyAsync(b.h_out, b.d_out.p, (size_t)b.rows * K * 4, hipMemcpyDeviceToHost, b.stream));
  BST_HIP(hipEventRecord(b.ev, b.stream));
  b.busy = true;
}

// Copy a finished batch's results to its requests (outside the lock).
void unpack_batch(bst_server* srv, Batch& b) {
  const int32_t K = srv->K;
  int64_t off = 0;
  for (Request* r : b.reqs) {
    for (int32_t k = 0; k < K; k++)
      std::memcpy(r->out + (size_t)k * r->n, b.h_out + (size_t)k * b.rows + off, (size_t)r->n * 4);
    off += r->n;
  }
}

// Mark a batch's requests done and free its budget (under the lock).
void retire_batch(bst_server* srv, Batch& b) {
  const auto now = Clock::now();
  for (Request* r : b.reqs) {
    r->latency_ns = std::chrono::duration_cast<std::chrono::nanoseconds>(now - r->t0).count();
    r->done = true;
  }
  b.reqs.clear();
  b.busy = false;
  b.retiring = false;
  srv->used -= b.bytes;
  b.bytes = 0;
  b.rows = 0;
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
  for (auto& b : srv->batches)
    for (auto* r : b.reqs) {
      r->status = code;
      r->done = true;
    }
  srv->cv_done.notify_all();
}

void worker_loop(bst_server* srv) {
  try {
    BST_HIP(hipSetDevice(srv->ctx->device));
    const int64_t max_batch_rows = 1 << 20;
    std::vector<Batch*> finished, admitted;
    for (;;) {
      finished.clear();
      admitted.clear();
      {
        std::unique_lock<std::mutex> lk(srv->mu);
        // Batches whose device work is done (results unpacked below).
        for (auto& b : srv->batches) {
          if (b.busy && !b.retiring && hipEventQuery(b.ev) == hipSuccess) {
            b.retiring = true;
            finished.push_back(&b);
          }
        }
        // Admit queued requests into a free batch while the budget allows.
        for (auto& b : srv->batches) {
          if (b.busy || srv->queue.empty()) continue;
          while (!srv->queue.empty()) {
            Request* r = srv->queue.front();
            if (srv->used + r->bytes > srv->budget) break;
            if (b.rows > 0 && b.rows + r->n > max_batch_rows) break;
            srv->queue.pop_front();
            b.reqs.push_back(r);
            b.rows += r->n;
            b.bytes += r->bytes;
            srv->used += r->bytes;
          }
          if (!b.reqs.empty()) {
            b.busy = true;
            admitted.push_back(&b);
          }
        }
        bool any_busy = false;
        for (auto& b : srv->batches) any_busy = any_busy || b.busy;
        if (srv->stop && srv->queue.empty() && !any_busy) return;
        if (finished.empty() && admitted.empty() && !any_busy)
          srv->cv_work.wait(lk, [&] { return srv->stop || !srv->queue.empty(); });
      }
      // Copies and launches without the lock: only this thread touches an
      // admitted or finished batch, and its requests' buffers belong to it
      // until they are marked done.
      for (Batch* b : admitted) launch_batch(srv, *b);
      for (Batch* b : finished) unpack_batch(srv, *b);
      if (!finished.empty()) {
        {
          std::lock_guard<std::mutex> lk(srv->mu);
          for (Batch* b : finished) retire_batch(srv, *b);
        }
        srv->cv_done.notify_all();
      }
      if (finished.empty() && admitted.empty())
        std::this_thread::sleep_for(std::chrono::microseconds(20));
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
    auto* srv = new bst_server();
    srv->ctx = ctx;
    srv->model = model;
    srv->budget = budget;
    srv->F = (int32_t)model->info.n_features;
    srv->K = (int32_t)model->info.n_groups;
    srv->batches.resize(std::max(1, std::min(max_inflight, 8)));
    for (auto& b : srv->batches) {
      BST_HIP(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
      BST_HIP(hipEventCreateWithFlags(&b.ev, hipEventDisableTiming));
    }
    srv->worker = std::thread(worker_loop, srv);
    *out = srv;
    return BST_OK;
  });
}

int bst_server_submit(bst_server* srv, const float* X, int64_t n, int32_t F, float* out,
                      uint64_t* ticket) {
  return sguard([&] {
    if (!srv || !X || !out || !ticket) raise(BST_E_INVALID, "null argument");
    if (F != srv->F) raise(BST_E_SHAPE, "request feature count differs from the model's");
    if (n <= 0) raise(BST_E_SHAPE, "empty request");
    auto* r = new Request();
    r->X = X;
    r->n = n;
    r->F = F;
    r->out = out;
    r->bytes = n * (int64_t)(F + srv->K) * 4;
    if (r->bytes > srv->budget) {
      delete r;
      raise(BST_E_OOM, "request exceeds the serving HBM budget");
    }
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
    for (auto& b : srv->batches) {
      if (b.stream) (void)hipStreamSynchronize(b.stream);
      if (b.h_in) (void)hipHostFree(b.h_in);
      if (b.h_out) (void)hipHostFree(b.h_out);
      if (b.ev) (void)hipEventDestroy(b.ev);
      if (b.stream) (void)hipStreamDestroy(b.stream);
    }
    for (auto& kv : srv->reqs) delete kv.second;
    delete srv;
    return BST_OK;
  });
}

}  // extern "C"
