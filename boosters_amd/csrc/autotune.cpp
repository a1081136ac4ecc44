// autotune.cpp — measured choice of the histogram launch shape.
//
// No counterpart in the reference (its histograms are built on the CPU).
// Four knobs decide how a step's histogram work is cut into blocks and where
// each block keeps its sub-histogram:
//   fg            features per block: 32 (128 KB LDS sub-histogram, one block
//                 per CU, 4 waves per SIMD) or 16 (64 KB, two blocks per CU,
//                 8 waves per SIMD, each row read by twice as many blocks);
//   tile_min      fewest rows per block; a step's tile (hist_step_tile) is
//                 the 512-row multiple that spreads its rows over at most
//   blocks_per_cu x n_cu blocks (more blocks balance better, fewer make fewer
//                 partial histograms for the finish pass to add);
//   priv          sub-histograms in LDS (zeroed, filled with LDS atomics and
//                 flushed per item) or in global memory (the item's own
//                 section, filled with device atomics at L2: no LDS zeroing
//                 or flush, but an L2 atomic per (row, feature) and g/h).
// Histogram sums are exact integers, so every shape gives the same
// histograms: the choice changes time only, never the trees.
//
// The tuner times the histogram + finish launches of the steps a tree of this
// configuration will run, on this dataset's own bins: for depth-wise growth
// the root (rows in place when the round is unsampled with one tree) and
// every level's smaller children (n/2 x 0.75^(L-1) rows at level L, node
// sizes spread over 0.25x .. 1.75x of the mean), for leaf-wise the root and single
// nodes of n/2 .. n/256 rows.  Gathered rows follow a scattered permutation,
// as after partitioning.  Three rounds over all candidates (the first warms
// clocks and caches and is discarded) keep each (candidate, step) pair's
// fastest time; the cheapest weighted total wins.  Results are cached in the
// context per data shape.  BST_TRACE_TUNE=1 prints every candidate's total.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "host_common.h"

namespace bst {

namespace {

constexpr uint32_t kTileMax = 65536;         // the device planner's largest tile
constexpr uint64_t kMaxStepRows = 4u << 20;  // rows of the largest measured step
constexpr uint32_t kSlots = 64;              // histogram slots the measurement cycles over

struct Shape {
  uint32_t nodes;      // built nodes in the step
  uint32_t rows;       // mean rows per node
  double weight;       // how many such steps a tree runs
  bool gather = true;  // rows read through the row-id list
  bool spread = false; // node sizes spread over 0.25x .. 1.75x of the mean
};

uint32_t step_tile(uint32_t tile_min, uint32_t target, uint32_t NG, uint64_t rows) {
  return hist_step_tile(tile_min, kTileMax, target, NG, rows);
}

std::vector<Shape> tree_shapes(int64_t n, const bst_gbdt_config& cfg, int T, bool root_gather) {
  std::vector<Shape> sh;
  const uint32_t N = (uint32_t)std::min<int64_t>(n, 0xFFFFFFFFll);
  sh.push_back({(uint32_t)T, N, 1.0, root_gather, false});
  if (cfg.growth == BST_GROW_DEPTHWISE) {
    // Level L builds the smaller child of each of its 2^(L-1) parents, in
    // nodes of uneven sizes.  Uneven splits make the smaller children hold
    // less than half the rows, the more so the deeper the level: taken as
    // n/2 x 0.75^(L-1).
    for (int L = 1; L < cfg.max_depth; L++) {
      const uint32_t nodes = 1u << std::min(L - 1, 10);
      const double total = 0.5 * (double)n * std::pow(0.75, L - 1);
      const uint32_t rows = (uint32_t)std::max(1.0, total / nodes);
      if (rows < 64) break;
      sh.push_back({nodes * (uint32_t)T, rows, 1.0, true, nodes > 1});
    }
  } else {
    const double steps = std::max(1, (int)cfg.max_leaves - 1);
    const int ks[4] = {1, 3, 5, 8};  // n/2, n/8, n/32, n/256
    for (int k : ks) {
      const uint32_t rows = (uint32_t)std::max<int64_t>(1, n >> k);
      sh.push_back({(uint32_t)T, rows, steps / 4.0, true, false});
    }
  }
  // Keep each measured step within kMaxStepRows rows (wide class-batched
  // steps): fewer rows per node, the weight scaled to the rows it stands for.
  for (Shape& x : sh) {
    const uint64_t total = (uint64_t)x.nodes * x.rows;
    if (total > kMaxStepRows) {
      const uint32_t r = (uint32_t)std::max<uint64_t>(64, kMaxStepRows / x.nodes);
      x.weight *= (double)x.rows / (double)r;
      x.rows = r;
    }
  }
  return sh;
}

// One step's work list: (tile, group) items and finish entries.
struct Plan {
  std::vector<HistItem> items;
  std::vector<uint32_t> fin;
  uint32_t nodes = 0;
  bool gather = true;
};

Plan plan_step(const Shape& sh, uint32_t NG, uint32_t slots, int64_t n, uint32_t tile_min,
               uint32_t target) {
  Plan pl;
  pl.nodes = sh.nodes;
  pl.gather = sh.gather;
  const uint64_t total = (uint64_t)sh.nodes * sh.rows;
  const uint32_t tile = step_tile(tile_min, target, NG, total);
  uint64_t pos = 0;
  for (uint32_t v = 0; v < sh.nodes; v++) {
    // Spread sizes: 0.25x .. 1.75x of the mean, linear in the node index.
    const double f = sh.spread ? 0.25 + 1.5 * (double)v / (double)(sh.nodes - 1) : 1.0;
    const uint32_t rows =
        (uint32_t)std::min<double>((double)n, std::max(1.0, std::floor(f * (double)sh.rows)));
    const uint32_t nt = hist_node_tiles(rows, tile, kTileMax);
    const uint32_t ts = hist_node_tile_rows(rows, nt);
    const uint32_t slot = v % slots;
    const uint64_t span = (uint64_t)std::max<int64_t>(1, n - (int64_t)rows + 1);
    const uint32_t begin = (uint32_t)(pos % span);  // consecutive ranges, wrapped inside n
    pos += rows;
    pl.fin.insert(pl.fin.end(), {slot, 0xFFFFFFFFu, (uint32_t)pl.items.size(), nt});
    for (uint32_t k = 0; k < nt; k++) {
      const uint32_t b0 = k * ts, cnt = std::min(ts, rows - b0);
      for (uint32_t g = 0; g < NG; g++)
        pl.items.push_back({v, slot, begin + b0, cnt, g, nt == 1 ? 0x80000000u : 0u});
    }
  }
  return pl;
}

struct Bench {
  bst_ctx* ctx = nullptr;
  DatasetDev d{};
  DevBuf<uint32_t> idx;
  DevBuf<float2> ogh;
  DevBuf<int32_t> qexp;
  DevBuf<long long> hist, partial;
  DevBuf<HistItem> items;
  DevBuf<uint32_t> fin, count;
  hipEvent_t a = nullptr, b = nullptr;

  ~Bench() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }

  // Time (ms) of one step's histogram + finish launches: the fastest of
  // `runs` runs.
  float run(const Plan& pl, int fg, int priv, int runs) {
    DatasetDev dd = d;
    dd.FG = fg;
    dd.hist_priv = priv;
    dd.NG = (d.E + fg - 1) / fg;
    const int64_t slot_elems = (int64_t)dd.NG * 256 * fg * 2;
    h2d(ctx->stream, items.p, pl.items.data(), pl.items.size());
    h2d(ctx->stream, fin.p, pl.fin.data(), pl.fin.size());
    const uint32_t n_items = (uint32_t)pl.items.size();
    h2d(ctx->stream, count.p, &n_items, 1);
    // The training launch: a resident grid walking a device-side item count.
    const int grid = (hist_blocks_resident(fg) * ctx->n_cu + 7) / 8 * 8;
    float best = 1e30f;
    for (int r = 0; r < runs; r++) {
      BST_HIP(hipEventRecord(a, ctx->stream));
      launch_hist(ctx->stream, dd, items.p, grid, idx.p, ogh.p, 0, qexp.p, hist.p, slot_elems,
                  nullptr, pl.gather, 0, count.p, partial.p, nullptr, nullptr);
      launch_hist_finish(ctx->stream, fin.p, (int)pl.nodes, dd.NG, fg, partial.p, hist.p,
                         slot_elems, nullptr);
      BST_HIP(hipEventRecord(b, ctx->stream));
      BST_HIP(hipEventSynchronize(b));
      float ms = 0.0f;
      BST_HIP(hipEventElapsedTime(&ms, a, b));
      best = std::min(best, ms);
    }
    return best;
  }
};

}  // namespace

HistTuning choose_hist_tuning(bst_ctx* ctx, const DatasetDev& d, const bst_gbdt_config& cfg,
                              const bst_ext_config& ext, int T) {
  HistTuning t;
  const bool fix_fg = ext.hist_fg != 0, fix_tile = ext.hist_rows_min != 0,
             fix_bpc = ext.hist_blocks_per_cu != 0, fix_priv = ext.hist_priv != 0;
  if (fix_fg) t.fg = ext.hist_fg;
  if (fix_tile) t.tile_min = (uint32_t)ext.hist_rows_min;
  if (fix_bpc) t.blocks_per_cu = (uint32_t)ext.hist_blocks_per_cu;
  if (fix_priv) t.priv = ext.hist_priv;
  if (!ext.autotune || (fix_fg && fix_tile && fix_bpc && fix_priv) || d.n < 4096 || d.E <= 0)
    return t;

  const std::vector<int64_t> key = {d.n, d.E, d.S, d.hi_off >= 0, T, cfg.growth,
                                    cfg.growth == BST_GROW_DEPTHWISE ? cfg.max_depth : cfg.max_leaves,
                                    ext.hist_fg, ext.hist_rows_min, ext.hist_blocks_per_cu,
                                    ext.hist_priv, ctx->n_cu};
  auto hit = ctx->tune_cache.find(key);
  if (hit != ctx->tune_cache.end()) {
    HistTuning c = hit->second;
    c.tune_ms = 0;
    return c;
  }
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<int> fgs = fix_fg ? std::vector<int>{t.fg} : std::vector<int>{32, 16};
  std::vector<uint32_t> tiles =
      fix_tile ? std::vector<uint32_t>{t.tile_min} : std::vector<uint32_t>{512, 1024, 2048};
  std::vector<uint32_t> bpcs =
      fix_bpc ? std::vector<uint32_t>{t.blocks_per_cu} : std::vector<uint32_t>{1, 2, 4};
  std::vector<int> privs = fix_priv ? std::vector<int>{t.priv} : std::vector<int>{1, 2};
  // Unsampled single-tree rounds read the root's rows in place (no gather).
  const bool sampled = !(std::fabs(cfg.subsample - 1.0f) < 1e-6f) || ext.goss_top_rate > 0.0f;
  const std::vector<Shape> shapes = tree_shapes(d.n, cfg, T, T > 1 || sampled);

  // Work lists of every candidate first, to size the buffers once.
  struct Cand {
    int fg;
    uint32_t tile_min, bpc;
    int priv;
    std::vector<Plan> plans;
  };
  std::vector<Cand> cands;
  size_t max_items = 0, max_fin = 0, max_partial = 0;
  for (int priv : privs)
  for (int fg : fgs)
    for (uint32_t tm : tiles)
      for (uint32_t bpc : bpcs) {
        // Global-memory sub-histograms are measured at the default spread
        // (two blocks per CU): their cost is the atomics, not the block count.
        if (priv == 2 && !fix_bpc && bpc != 2) continue;
        Cand c{fg, tm, bpc, priv, {}};
        const uint32_t NG = (uint32_t)((d.E + fg - 1) / fg);
        const uint32_t target = (uint32_t)std::max(64, (int)bpc * ctx->n_cu);
        for (const Shape& sh : shapes) {
          c.plans.push_back(plan_step(sh, NG, kSlots, d.n, tm, target));
          const Plan& pl = c.plans.back();
          max_items = std::max(max_items, pl.items.size());
          max_fin = std::max(max_fin, pl.fin.size());
          max_partial = std::max(max_partial, pl.items.size() * (size_t)256 * fg * 2);
        }
        cands.push_back(std::move(c));
      }

  Bench B;
  B.ctx = ctx;
  B.d = d;
  const size_t n = (size_t)d.n;
  // A scattered permutation of the rows (multiplier coprime to n), like the
  // row order of a node after partitioning.
  auto gcd = [](uint64_t x, uint64_t y) {
    while (y) {
      const uint64_t r = x % y;
      x = y;
      y = r;
    }
    return x;
  };
  uint64_t mul = 2654435761ull % n;
  while (mul < 2 || gcd(mul, n) != 1) mul++;
  std::vector<uint32_t> perm(n);
  for (size_t i = 0; i < n; i++) perm[i] = (uint32_t)((i * mul) % n);
  B.idx.alloc(n);
  h2d(ctx->stream, B.idx.p, perm.data(), n);
  std::vector<float2> gh(n, make_float2(0.25f, 0.125f));
  B.ogh.alloc(n);
  h2d(ctx->stream, B.ogh.p, gh.data(), n);
  const int32_t qe[2] = {20, 20};
  B.qexp.alloc(2);
  h2d(ctx->stream, B.qexp.p, qe, 2);
  const int64_t slot_elems = std::max<int64_t>((int64_t)((d.E + 15) / 16) * 256 * 16 * 2,
                                               (int64_t)((d.E + 31) / 32) * 256 * 32 * 2);
  B.hist.alloc((size_t)kSlots * slot_elems);
  B.partial.alloc(std::max<size_t>(max_partial, 1));
  B.items.alloc(std::max<size_t>(max_items, 1));
  B.fin.alloc(std::max<size_t>(max_fin, 1));
  B.count.alloc(1);
  BST_HIP(hipEventCreate(&B.a));
  BST_HIP(hipEventCreate(&B.b));

  // Rounds over all candidates (the first one warms the clocks and caches),
  // keeping each (candidate, step) pair's fastest time: interleaving the
  // repeats spreads clock drift over every candidate alike.
  constexpr int kRounds = 3;
  std::vector<std::vector<float>> tmin(cands.size(), std::vector<float>(shapes.size(), 1e30f));
  // LDS candidates (global ones too when the knob is fixed to global) first,
  // over all rounds.
  const bool lds_first = privs.size() > 1;
  for (int r = 0; r < kRounds; r++)
    for (size_t ci = 0; ci < cands.size(); ci++) {
      if (lds_first && cands[ci].priv == 2) continue;
      for (size_t i = 0; i < shapes.size(); i++) {
        const float ms = B.run(cands[ci].plans[i], cands[ci].fg, cands[ci].priv, 1);
        if (r > 0) tmin[ci][i] = std::min(tmin[ci][i], ms);
      }
    }
  auto total_of = [&](size_t ci) {
    double total = 0;
    for (size_t i = 0; i < shapes.size(); i++) total += shapes[i].weight * tmin[ci][i];
    return total;
  };
  if (lds_first) {
    // Global-memory candidates after: the clocks and caches are warm, so one
    // run per step, the cheapest steps first, and a candidate is dropped as
    // soon as the steps measured so far already cost more than the best LDS
    // candidate's whole tree (an L2 atomic per (row, feature) and g/h is
    // tens of times the LDS cost on every shape measured, so this stops
    // after a step or two instead of timing whole trees of it).
    double best_lds = 1e300;
    for (size_t ci = 0; ci < cands.size(); ci++)
      if (cands[ci].priv != 2) best_lds = std::min(best_lds, total_of(ci));
    std::vector<size_t> order(shapes.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
      return (double)shapes[a].nodes * shapes[a].rows * shapes[a].weight <
             (double)shapes[b].nodes * shapes[b].rows * shapes[b].weight;
    });
    for (size_t ci = 0; ci < cands.size(); ci++) {
      if (cands[ci].priv != 2) continue;
      double part = 0;
      for (size_t i : order) {
        tmin[ci][i] = B.run(cands[ci].plans[i], cands[ci].fg, cands[ci].priv, 1);
        part += shapes[i].weight * tmin[ci][i];
        if (part > best_lds) break;  // the unmeasured steps stay at 1e30: never picked
      }
    }
  }
  const bool trace = std::getenv("BST_TRACE_TUNE") != nullptr;
  double best = 1e300;
  HistTuning pick = t;
  for (size_t ci = 0; ci < cands.size(); ci++) {
    const Cand& c = cands[ci];
    const double total = total_of(ci);
    if (trace) {
      // A pruned global candidate shows the cost of the steps it was timed on.
      double part = 0;
      int timed = 0;
      for (size_t i = 0; i < shapes.size(); i++)
        if (tmin[ci][i] < 1e29f) {
          part += shapes[i].weight * tmin[ci][i];
          timed++;
        }
      std::fprintf(stderr, "[tune] %s fg %d rows_min %u blocks/cu %u: %.3f ms%s\n",
                   c.priv == 2 ? "global" : "lds", c.fg, c.tile_min, c.bpc, part,
                   timed < (int)shapes.size() ? " (pruned: steps timed so far)" : "");
    }
    if (total < best) {
      best = total;
      pick.fg = c.fg;
      pick.tile_min = c.tile_min;
      pick.blocks_per_cu = c.bpc;
      pick.priv = c.priv;
    }
  }
  BST_HIP(hipStreamSynchronize(ctx->stream));
  pick.tuned = 1;
  pick.tune_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  ctx->tune_cache[key] = pick;
  return pick;
}

}  // namespace bst
