// engine.h — internal interfaces of the gfx950 GBDT engine (host <-> kernels).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>
#include <stdint.h>

#include "../../include/boosters_amd.h"

namespace bst {

constexpr int kMaxBinsDev = 256;  // u8 bins per feature; bin 256 lives in the hi plane
constexpr int kMaxCatWords = 8;   // categorical bitsets cover bins 0..255

// Per effective feature, for value_to_bin (BinMapper, bin_mapper.rs:210-258).
struct BinMeta {
  uint32_t orig;        // column in the raw [F][N] matrix
  uint32_t n_bins;
  uint32_t n_search;    // numeric: bins searched (NaN bin excluded)
  uint32_t table_off;   // offset into the f32 threshold / i32 category tables
  uint32_t default_bin;
  uint8_t is_cat;
  uint8_t missing_nan;
  uint8_t pad[2];
};

// Per effective feature, for split finding and partitioning.
struct FeatMeta {
  uint32_t n_bins;
  uint32_t default_bin;
  uint8_t is_cat;
  uint8_t has_missing;
  uint8_t pad[2];
};

// One block of histogram work: a tile of a node's rows x one feature group.
struct HistItem {
  uint32_t node;     // global node id (tree slot * max_nodes + partition id)
  uint32_t slot;     // histogram slot to accumulate into
  uint32_t begin;    // first position in idx/ogh of this tile
  uint32_t count;    // rows in this tile
  uint32_t group;    // feature group
  uint32_t tree;     // tree slot (selects the quantisation scale); bit 31: the
                     // tile is the whole node (store directly); bit 30: the rows
                     // live in the alternate idx/ogh buffer
};

// A split decision on device (SplitInfo, split/types.rs:34-55).
struct DevSplit {
  float gain;  // -inf = invalid
  uint32_t feature;
  uint32_t bin;
  uint32_t flags;  // bit0 default_left, bit1 categorical
  uint32_t cats[kMaxCatWords];
  long long lg, lh;  // exact fixed-point (g, h) sums of the rows going left
};

// One node to split: rows [begin, begin+count) of idx/ogh.
struct PartItem {
  uint32_t node;       // global id of the parent (keeps the left rows)
  uint32_t begin;
  uint32_t count;
  uint32_t tile0;      // first partition tile of this node
  uint32_t n_tiles;
  uint32_t right_node; // global id of the new right child
  uint32_t feature;    // effective column
  uint32_t pad;        // device-driven partition: bit 0 = rows are in the alternate buffer
  // Device-driven partition: the node's split decision, copied by the planner
  // so that the partition need not load the split and the feature meta:
  uint32_t split_bin;    // left: bin <= split_bin (numerical)
  uint32_t split_flags;  // DevSplit::flags (bit 0 default left, bit 1 categorical), bit 2 below
  uint32_t miss_bin;     // the feature's missing bin, or ~0 without one
  uint32_t pad2;
  // split_flags bit 2 (kItemLeaves): the children are leaves (depth-wise last
  // step) and the partition adds their values to the predictions instead of
  // moving the rows; values of the left (kept) and right child, learning rate
  // applied.
  float leaf_l, leaf_r;
  uint32_t pad3, pad4;
};
constexpr uint32_t kItemLeaves = 4u;
// PartTile::src bits: 1 = rows in the alternate buffer, 2 = the item applies
// leaf values (kItemLeaves), so the tile needs no (g, h), 4 = the node's rows
// are stored from its last position down (a right child of the one-pass
// partition), so the tile's rows are begin, begin - 1, ...
constexpr uint32_t kTileAlt = 1u, kTileLeaves = 2u, kTileRev = 4u;

struct PartTile {
  uint32_t item;   // index into PartItem list
  uint32_t begin;  // absolute position of the tile's first row (in row order)
  uint32_t count;
  uint32_t src;    // device planner: kTileAlt / kTileLeaves / kTileRev bits; host: 0
};

// Per-node segment for sums / leaf updates.
struct Segment {
  uint32_t node;
  uint32_t begin;
  uint32_t count;
  uint32_t tree;  // tree slot; bit 31: rows in the alternate idx buffer
};

// Packed tree node for traversal (Tree SoA, repr/gbdt/tree.rs:32-47).
struct DevNode {
  uint32_t meta;   // bits 0..23 feature, 24 leaf, 25 default_left, 26 categorical
  float value;     // threshold (numeric) or leaf value
  uint32_t left;   // tree-local ids
  uint32_t right;
};
constexpr uint32_t kNodeLeaf = 1u << 24, kNodeDefLeft = 1u << 25, kNodeCat = 1u << 26;

struct DatasetDev {
  int64_t n = 0;
  int32_t F = 0;        // original features
  int32_t E = 0;        // effective (non-trivial) features
  int32_t S = 0;        // row stride of bins_rm in bytes
  int32_t hi_off = -1;  // byte offset of the hi-bit mask in a row, -1 = none
  int32_t FG = 32;      // features per histogram group (lanes of one row)
  int32_t NG = 0;       // number of groups
  int32_t max_cat_bins = 0;  // most categories of a categorical feature (0: none)
  int32_t hist_priv = 1;     // histogram sub-histograms: 1 in LDS, 2 in global memory
  int64_t hi_words = 0; // u32 words per feature in hi_cm
  uint8_t* bins_rm = nullptr;   // [n][S]
  uint8_t* bins_cm = nullptr;   // [E][n]
  uint32_t* hi_cm = nullptr;    // [E][hi_words]
  FeatMeta* meta = nullptr;     // [E]
  float* raw = nullptr;         // [F][n] raw features (sampled prediction updates)
};

// ---- launchers (kernels.hip) ----------------------------------------------
void launch_value_to_bin(hipStream_t s, const float* X, int64_t n, int E, const BinMeta* meta,
                         const float* thr, const int32_t* cats, uint8_t* bins_cm, uint32_t* hi_cm,
                         int64_t hi_words);
void launch_to_rowmajor(hipStream_t s, const DatasetDev& d);

// One leaf of a renewal pass: its rows are positions begin .. begin + count - 1
// of the partition's row list; output picks the prediction row and alpha the
// quantile.
struct RenewLeaf {
  uint32_t begin;
  uint32_t count;
  uint32_t output;
  float alpha;
};
// Leaf renewal (quantile objectives): out[i] = the alpha-quantile of
// y[row] - preds[output][row] over leaf i's rows (NaN for an empty leaf).
size_t renew_temp_bytes(int64_t n_pos, int n_leaves);
void renew_leaves(hipStream_t s, const RenewLeaf* leaves, int n_leaves, int64_t n_pos,
                  const uint32_t* idx, const float* y, const float* preds, const float* w,
                  int64_t n, unsigned long long* keys, unsigned long long* keys_sorted,
                  uint32_t* seg_begin, uint32_t* seg_end, void* temp, size_t temp_bytes,
                  float* out);

// Metric payloads: accuracy threshold, Huber delta, one quantile alpha per
// output.
struct MetricParams {
  float accuracy_threshold = 0.5f;
  double huber_delta = 1.0;
  float alphas[16] = {};
  int n_alphas = 0;
};
void launch_metric(hipStream_t s, int metric, int transform, const MetricParams& mp,
                   const float* preds, int K, const float* y, const float* w, int64_t n,
                   double* partial, int n_blocks);

// Objective payloads: pseudo-Huber delta, one pinball alpha per output.
struct ObjParams {
  float huber_delta = 1.0f;
  float alphas[16] = {};
};
void launch_gradients(hipStream_t s, int objective, const ObjParams& op, const float* preds,
                      const float* y, const float* w, int64_t n, int K, float* g, float* h,
                      uint32_t* absmax);
// LambdaRank gradients of one output over ragged query groups
// (group_off[n_groups + 1] row offsets, each group <= 4096 rows); disc[r] =
// 1 / log2(2 + r) for r below the largest group; absmax as launch_gradients.
constexpr int kRankMaxGroupRows = 4096;
void launch_lambdarank(hipStream_t s, const float* preds, const float* y, const float* w,
                       const int64_t* group_off, int64_t n_groups, const double* disc, float* g,
                       float* h, uint32_t* absmax);
// Reads absmax[0..2K) and resets it to zero for the next round.
void launch_qscale(hipStream_t s, uint32_t* absmax, int K, int64_t n_rows, int32_t* qexp);
// Fixed-point exponent of one gradient or hessian column from its max |x|: the
// e with max|x| * 2^e * n_rows < 2^62, so any sum over the rows fits (0 for
// an all-zero or non-finite column).
__device__ inline int32_t qscale_exponent(float m, int64_t n_rows) {
  int lg = 0;
  while ((1LL << lg) < n_rows && lg < 62) lg++;
  int e = 0;
  if (m > 0.0f && isfinite(m)) {
    int x;
    frexpf(m, &x);  // m < 2^x
    e = 62 - lg - x;
    if (e > 1000) e = 1000;
  }
  return e;
}

void launch_init_rows(hipStream_t s, uint32_t* idx, float2* ogh, const float* g, const float* h,
                      const uint32_t* sampled, int64_t m);

void launch_hist(hipStream_t s, const DatasetDev& d, const HistItem* items, int n_items,
                 const uint32_t* idx, const float2* ogh, int64_t tree_stride_rows,
                 const int32_t* qexp, long long* hist, int64_t slot_elems, long long* node_sum,
                 bool gather, int rows_per_block_hint, const uint32_t* dyn = nullptr,
                 long long* partial = nullptr, const uint32_t* idx_alt = nullptr,
                 const float2* ogh_alt = nullptr);
// node_sum may be null (no totals); partial non-null: multi-tile nodes store
// one partial histogram per item there instead of global atomics, and
// launch_hist_finish sums them (and applies the subtraction trick).
// Histogram tile (rows per (tile, group) item) of a step over `rows` rows:
// as many tiles per group as fit target items (target / NG), so that the
// step's items fill the resident grid once and no block takes two, rounded
// up to whole 512-row chunks, within [tile_min, tile_max].  Nodes round up
// to whole tiles, so a step of several nodes gets up to one extra item per
// node and group.
__host__ __device__ inline uint32_t hist_step_tile(uint32_t tile_min, uint32_t tile_max,
                                                   uint32_t target, uint32_t NG, uint64_t rows) {
  const uint64_t per_group = target / NG > 0 ? target / NG : 1;
  uint64_t tile = (rows + per_group - 1) / per_group;
  tile = (tile + 511) / 512 * 512;
  if (tile < tile_min) tile = tile_min;
  if (tile > tile_max) tile = tile_max;
  return (uint32_t)tile;
}

// Tiles of one node of `rows` rows at a step's tile size: the nearest whole
// number of tiles (one at least for a non-empty node; tiles of at most
// tile_max rows), so that no node ends in a sliver tile that costs a block,
// a partial histogram and its finish for a few rows.  Never more than
// ceil(rows / tile), so work-list capacities sized for that hold.
__host__ __device__ inline uint32_t hist_node_tiles(uint32_t rows, uint32_t tile,
                                                    uint32_t tile_max) {
  if (rows == 0) return 0;
  uint32_t nt = (uint32_t)(((uint64_t)rows + tile / 2) / tile);
  const uint32_t lo = (uint32_t)(((uint64_t)rows + tile_max - 1) / tile_max);
  if (nt < lo) nt = lo;
  return nt < 1 ? 1 : nt;
}
// Rows per tile of a node split into nt tiles (the last may be shorter).
__host__ __device__ inline uint32_t hist_node_tile_rows(uint32_t rows, uint32_t nt) {
  return nt ? (uint32_t)(((uint64_t)rows + nt - 1) / nt) : 0;
}

// Histogram blocks one CU holds at once for a feature-group width (LDS).
inline int hist_blocks_resident(int FG) { return FG == 16 ? 2 : 1; }
// Histogram lanes of padding features (index E .. FG * NG - 1) read their
// own bin byte and NaN bit: the row must hold them (as zeros).
inline bool hist_layout_padded(const DatasetDev& d, int FG) {
  const int ng = (d.E + FG - 1) / FG;
  return d.E > 0 && FG * ng <= d.S && (d.hi_off < 0 || d.hi_off + FG / 8 * ng <= d.S);
}

void launch_hist_finish(hipStream_t s, const uint32_t* fin, int n, int NG, int FG,
                        const long long* partial, long long* hist, int64_t slot_elems,
                        const uint32_t* dyn);
void launch_init_rows_sums(hipStream_t s, uint32_t* idx, float2* ogh, const float* g,
                           const float* h, int64_t n, int T, const int32_t* qexp,
                           uint32_t max_nodes, long long* node_sum);
void launch_node_sums(hipStream_t s, const Segment* segs, int n_segs, const float2* ogh,
                      int64_t tree_stride_rows, const int32_t* qexp, long long* node_sum,
                      const uint32_t* dyn = nullptr);
void launch_subtract(hipStream_t s, const uint32_t* quads, int n, long long* hist,
                     int64_t slot_elems, long long* node_sum, const uint32_t* dyn = nullptr);
void launch_zero_slots(hipStream_t s, const uint32_t* slots, int n, long long* hist,
                       int64_t slot_elems, const uint32_t* dyn = nullptr);

struct SplitParams {
  double lambda;
  double min_gain;
  double min_child_weight;
  uint32_t min_samples_leaf;
  uint32_t max_onehot_cats;
  int32_t rightmost;  // rayon tie rule (rightmost max) vs sequential (first max)
};
void launch_split(hipStream_t s, const DatasetDev& d, const uint32_t* nodes, const uint32_t* slots,
                  int n_nodes, const uint32_t* feats, int n_feats, const long long* hist,
                  int64_t slot_elems, const long long* node_sum, const uint32_t* node_count,
                  const int32_t* qexp, uint32_t max_nodes, SplitParams p,
                  unsigned long long* node_key, DevSplit* cand, DevSplit* node_split,
                  const uint32_t* dyn = nullptr, bool finalize = true, int feat_stride = 0,
                  const uint32_t* minus = nullptr);
// feats: n_feats effective columns, shared by all trees (feat_stride 0) or
// one list per tree slot at feats + slot * feat_stride; null = columns
// 0 .. n_feats-1.  minus (optional, per node entry): a slot whose histogram
// the split subtracts from the entry's slot before scanning, writing the
// difference back into the entry's slot (for its candidate features), or
// ~0 for none; with kMinusNoWriteback set the difference is only scanned
// (no later step reads the node's histogram).  With dyn, n_nodes (the grid)
// must not exceed the entry lists: each block reads its first entry ahead
// of the count.
constexpr uint32_t kMinusNoWriteback = 0x80000000u;

// Stable partition of every listed node (partition.rs:73-145): count, scan,
// scatter, copy back, then children counts/sums.
void launch_partition(hipStream_t s, const DatasetDev& d, const PartItem* items, int n_items,
                      const PartTile* tiles, int n_tiles, const DevSplit* node_split,
                      uint32_t* idx, float2* ogh, uint32_t* idx_tmp, float2* ogh_tmp,
                      uint8_t* flag, uint32_t* tile_left, uint32_t* tile_off,
                      uint32_t* item_left, uint32_t* node_count, long long* node_sum,
                      const uint32_t* dyn_items = nullptr, const uint32_t* dyn_tiles = nullptr);

// Device-driven partition, one pass, stable: each node's left rows fill its
// range in the other buffer of the pair from the front in row order, its
// right rows fill the range from the back in row order (so the right child
// is read from its last position down: kTileRev).  A node's tiles are
// consecutive in the tile list and are claimed in order through *ticket
// (zeroed by the caller); tstat[k] (zeroed by the caller) carries tile k's
// left count and then its inclusive prefix among its node's tiles, and the
// last tile of item i writes the node's left count to cursor[2 i].  A
// look-back that does not finish sets *err.  Counts read from dyn_tiles;
// tiles[k].src names the source buffer and direction.  Items flagged
// kItemLeaves instead add leaf_l / leaf_r to
// preds[tree_output[node / max_nodes] * n_rows + row] for each of their rows
// (rows stay in place, cursors untouched).
constexpr uint32_t kFusedTileRows = 4096;  // rows per tile of launch_partition_fused (at most)
// Tile of the depth-wise steps whose partition applies leaf values (random
// prediction updates: more, shorter tiles keep more of them in flight).
constexpr uint32_t kFusedTileRowsSmall = 2048;
void launch_partition_fused(hipStream_t s, const DatasetDev& d, const PartItem* items,
                            const PartTile* tiles, int grid_tiles, const DevSplit* node_split,
                            uint32_t* idx0, float2* ogh0, uint32_t* idx1, float2* ogh1,
                            uint32_t* cursor, unsigned long long* tstat, uint32_t* ticket,
                            uint32_t* err, const uint32_t* dyn_tiles,
                            const uint32_t* tree_output, float* preds, int64_t n_rows,
                            uint32_t max_nodes);

void launch_leaf_update(hipStream_t s, const Segment* segs, int n_segs, const float* leaf_val,
                        const uint32_t* idx, int64_t tree_stride_rows, float* preds,
                        const uint32_t* tree_output, int64_t n_rows,
                        const uint32_t* dyn = nullptr, const uint32_t* idx_alt = nullptr);

// ---- prediction (inference/gbdt/predictor.rs:197-302) ----------------------
// The forest is cut into chunks of consecutive whole trees.  A tree of at
// most kPredChunkNodes nodes whose features fit 16 bits is stored a second
// time in a compact 8-byte layout (PredCNode) and belongs to an LDS chunk of
// at most kPredChunkNodes nodes: a block stages one chunk at a time in LDS
// (the next chunk's loads in flight while the current one is traversed) and
// its lanes traverse from there.  Any other tree forms a chunk of its own
// that is traversed from global memory in the 16-byte layout (n_nodes = 0).
constexpr int kPredChunkNodes = 3584;  // 28 KiB of 8-byte nodes per LDS buffer
constexpr int kPredMaxRows = 256;      // rows per block at most, one per lane
// Trees a lane traverses together (independent chains).  3584 = 7 x 512, so
// a chunk holds 7 complete depth-8 trees (511 nodes), 14 of depth 7, ...;
// a chunk that fits a few more than a multiple of 7 takes only that
// multiple (a group of fewer trees leaves chains idle).
constexpr int kPredILP = 7;
// A block stages the rows' feature values as order keys (pred_order_key):
// staged row 0 is all zeros and feature f is staged row f + 1.
// Compact node, x: split: staged row of its feature (f + 1, bits 0..15) |
// default_left (17) | categorical (18) | the left child's index in the LDS
// chunk (bits 19..31; the right child is the next index); leaf: staged row 0
// (bits 0..15 zero) | its own index in the chunk (bits 19..31).  y: numeric
// split: the threshold's order key + default_left; categorical split: its
// index in PredModel::cat_ref + 1; leaf: the bitwise NOT of its value's bits
// (a NaN value canonical, 0x7FC00000).  A traversal step goes to the child
// (x >> 19) + [key(x) + default_left >= y] (unsigned): NaN (key 0xFFFFFFFF)
// wraps to 0 with default_left and goes left, else right, and a leaf (key 0,
// y never 0) stays where it is.  Every LDS chunk ends with a sentinel leaf
// that chains past the chunk's last tree sit on.
constexpr uint32_t kCDefLeft = 1u << 17, kCCat = 1u << 18;
constexpr int kCShift = 19;
// Order key of an f32 for the compact traversal: unsigned order of keys is
// the order of the values (-0 and +0 equal), NaN is 0xFFFFFFFF, and every
// other key is at most 0xFF800000 (+inf).
__host__ __device__ inline uint32_t pred_order_key(uint32_t u) {
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) return 0xFFFFFFFFu;  // NaN
  if (u == 0x80000000u) u = 0u;                                // -0 -> +0
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
// The value of an order key (-0 comes back as +0, NaN as 0x7FFFFFFF).
__host__ __device__ inline uint32_t pred_key_bits(uint32_t k) {
  return (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
}
struct PredChunk {
  uint32_t first_tree, end_tree;  // trees [first_tree, end_tree)
  uint32_t node_begin;            // LDS chunk: first compact node (even); else unused
  uint32_t n_nodes;               // compact nodes staged; 0: one tree read from global memory
};
struct PredModel {
  const DevNode* nodes;        // 16-byte nodes, original ids (global-memory trees)
  const int64_t* tree_off;     // [n_trees + 1] into nodes
  const uint2* cnodes;         // compact nodes of the LDS chunks' trees
  const uint32_t* ctree_off;   // [n_trees] first compact node of each compact tree
  const uint32_t* cdepth;      // [n_trees] levels of each compact tree (its deepest leaf)
  const uint32_t* corig;       // [compact nodes] the node's original tree-local id
  const uint32_t* cat_ref;     // [2 x categorical compact nodes]: cat_start, cat_size
  const uint32_t* tree_group;
  const uint32_t *cat_start, *cat_size, *cat_words;
  const PredChunk* chunks;
  int n_chunks;
  int n_features;  // features the trees test (rows staged per block)
  int has_cat;     // some compact node is a categorical split
  int64_t n_trees;
  const float* base;
  int K;   // outputs of the model
  // The outputs one launch accumulates, g0 .. g0 + Kw - 1 (Kw <= kPredMaxK =
  // 16 registers per row): a model of more outputs is predicted in windows,
  // each launch adding only its window's trees.
  int g0, Kw;
};
// ---- serving: one persistent traversal launch --------------------------------
// The serving kernel keeps one block per CU resident and takes 256-row tiles
// (fewer rows per tile when the model's features need more LDS) of the
// batches the host publishes, in publication order, until the host asks it
// to stop or no tile has come for its idle time.  Batches live in pinned,
// coherent host memory (inputs [F][rows], outputs [K][rows]) that the kernel
// reads and writes directly; the descriptors and counters below are the
// whole protocol.
constexpr int kServeRing = 16;  // batches in flight at most
// A batch is a run of work units (claimed in order).  A batch of `split` = 0
// has one unit per tile of rows, each traversing the whole forest.  A batch
// too small to occupy the device has split > 0: tiles x split leaf units
// (tile u / split, the u % split-th of split ranges of the forest's chunks),
// each writing its trees' leaf values to the slot's leaf buffer, then one
// sum unit per tile, which waits until every leaf unit of the batch is done
// and adds the leaves of each row in forest order.
// Units are claimed in two lanes, each in publication order: lane 0 holds
// the batches cut into leaf and sum units (small ones), lane 1 the others,
// and a block looks at lane 0 first, so a small request is not queued behind
// a large batch's tiles.
struct ServeDesc {
  // The batch's units, one 64-bit word so that a reader never sees half of a
  // rewrite: lane << 63 | first unit number (units are numbered per lane)
  // << 16 | count.
  uint64_t range;
  uint64_t in_off, out_off;  // float offsets of its rows and outputs in the arenas
  uint64_t done_target;  // ServeDevState::tiles_done[slot] once its last unit is done
  uint64_t leaf_target;  // split > 0: tiles_done[slot] once every leaf unit is done
  uint64_t leaf_off;     // split > 0: float offset of the slot's leaf values [trees][rows]
  uint32_t rows, seq;
  uint32_t split, tiles;
};
constexpr uint32_t kServeMaxTiles = 65535;  // units of one batch (16 bits of range)
constexpr uint64_t kServeUnitMask = (1ull << 47) - 1;  // first-unit field of range
constexpr uint32_t kServeMaxSplit = 8;      // leaf units per tile at most (each reads its tile's rows)
struct ServeHostState {  // pinned coherent host memory
  uint64_t published[2];  // units published so far per lane (host; release stores)
  uint32_t stop;         // host: exit now
  uint32_t error;        // kernel: a claimed unit matched no descriptor, or a sum unit's
                         // wait for its batch's leaf units ran out (never expected)
  ServeDesc desc[kServeRing];
  uint32_t done_seq[kServeRing];  // kernel: seq of the slot's finished batch (a release store)
};
struct ServeDevState {   // device memory, zeroed before the first launch
  uint64_t next_unit[2];  // the next unit to claim per lane
  uint64_t seen_pub[2];   // published[] as last read from host memory by any block
  uint64_t tiles_done[kServeRing];  // units done per slot, over all its batches
  uint64_t last_claim;   // constant-rate clock of the latest claim by any block
};
// Tile rows of the serving kernel for a model testing F features (0 when the
// rows cannot be staged in LDS).
int serve_tile_rows(int F, size_t lds_limit);
// One persistent launch of n_cu blocks; `h` is the device address of the
// pinned state; the kernel exits when h->stop is set or once no block has
// claimed a unit for idle_ticks of the device's constant-rate clock (the
// blocks leave together, so a launch never lingers with a few of them).
void launch_serve(hipStream_t s, const PredModel& m, ServeHostState* h, ServeDevState* d,
                  const float* in_base, float* out_base, float* leaf_base, uint64_t idle_ticks,
                  size_t lds_limit, int n_cu);

// Rows per block and LDS bytes of a predict launch for a model testing F
// features: rows are staged feature-major when F x rows x 4 B fits beside the
// two node buffers (256, 128 or 64 rows); otherwise *stage_rows = false and
// the kernel reads features from global memory (256 rows per block).
size_t pred_lds_bytes(int F, size_t lds_limit, bool* stage_rows, int* rows_per_block);
// out[k][row] = transform(base[k] + leaves of the trees of group k, in forest
// order, f32).  leaf_idx nullable: [n_trees][n] tree-local leaf ids.
// leaf_scratch non-null: two-phase launch for batches too small to fill the
// device -- (row tile, chunk range) blocks write every tree's leaf value to
// leaf_scratch [n_trees][n], then one pass per row adds them in forest order
// (the same f32 sums, bit for bit, as the one-phase launch).
void launch_predict(hipStream_t s, const PredModel& m, const float* X, int64_t n,
                    int64_t x_stride, int F, int transform, float* out, int64_t out_stride,
                    uint32_t* leaf_idx, float* leaf_scratch, size_t lds_limit, int n_cu);
// The same over rows [0, n) of a larger call: X, out and leaf_idx rows are
// addressed with their own strides (x_stride, out_stride, leaf_stride);
// leaf_scratch holds [n_trees][n] for this launch.
void launch_predict_rows(hipStream_t s, const PredModel& m, const float* X, int64_t n,
                         int64_t x_stride, int F, int transform, float* out, int64_t out_stride,
                         uint32_t* leaf_idx, int64_t leaf_stride, float* leaf_scratch,
                         size_t lds_limit, int n_cu);
// out[k][row] += the listed trees' leaves, in order (training-time updates).
void launch_predict_accum(hipStream_t s, const PredModel& m, const float* X, int64_t n,
                          int64_t x_stride, int F, float* out, int64_t out_stride,
                          size_t lds_limit);
// Whether launch_predict takes the two-phase path for n rows (and the scratch
// it then needs, in floats).
bool pred_two_phase(const PredModel& m, int64_t n, int n_cu, size_t lds_limit,
                    int64_t* scratch_floats);
void launch_copy16(hipStream_t s, void* dst, const void* src, int64_t bytes);
constexpr int kPredMaxGroups = 16;   // outputs one prediction launch (and the serving loop) holds
constexpr int kMaxOutputs = 1024;    // outputs of a model (softmax classes)

}  // namespace bst
