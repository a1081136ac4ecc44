// grow_device.h — device-side tree growth control (the graph schedule).
//
// The host grower (trainer.cpp) decides which nodes to expand, allocates
// partition / tree node ids and builds the per-step work lists on the CPU,
// with two stream synchronisations per step.  The device grower moves exactly
// that control onto the GPU: single-block "controller" kernels read the
// split results, make the same decisions in the same order (depth-wise level
// order, leaf-wise max (gain, partition id) with n_leaves++ on pop), assign
// the same ids, and write the work lists and their counts to device memory.
// The heavy kernels are launched with worst-case grids and walk the
// device-side counts, so one boosting round is a fixed launch sequence that is
// captured once in a hipGraph and replayed every round.
//
// One step (per tree, one depth level or one leaf-wise pop):
//   plan      finalise the previous step's splits, (depth-wise) advance to the
//             next level, choose / decide / allocate, write partition items
//             and tiles
//   partition one pass into the other buffer of a pair (+ per-tile exact
//             totals); each partition id records which buffer holds it
//   children  counts, totals, histogram slots, histogram items, finish list,
//             split list
//   histogram per-item tiles (single-tile nodes direct, others as partials)
//   finish    partial sums + the subtraction trick
//   split     per (node, feature) candidates; the next plan picks the best
#pragma once
#include "engine.h"

namespace bst {

// Trees the controllers grow together (class batching) at most.
constexpr uint32_t kCtlMaxTrees = 16;

struct CtlParams {
  uint32_t T;           // trees grown together
  uint32_t MN;          // partition ids per tree
  uint32_t SPT;         // histogram slots per tree
  uint32_t MNT;         // tree nodes per tree
  uint32_t W;           // max candidates expanded per step per tree
  uint32_t WO;          // open-set capacity per tree
  uint32_t growth;      // 0 depth-wise, 1 leaf-wise
  uint32_t max_depth, max_leaves;
  uint32_t NG, E, FG;
  uint32_t hist_tile, part_tile, seg_tile;  // hist_tile: the largest histogram tile
  uint32_t part_tile_small;                  // partition tile of leaf-applying steps
  uint32_t hist_tile_min, hist_target;       // smallest tile; items wanted per step
  uint32_t rightmost;   // split tie rule across features (rayon: rightmost max)
  int64_t stride;       // positions per tree slot in idx/ogh
  float lr;
  double lambda, alpha, min_gain;
  int32_t out0;         // output of tree slot 0 (qexp / gradient row)
  // Depth-wise rounds over every row: the last step's partition adds the
  // children's leaf values to the predictions (no leaf-range update for them).
  uint32_t leaf_apply;
};

struct TreeCtl {
  uint32_t n_part, n_tnodes, n_leaves, n_open, n_next, n_free, depth, rows;
};

// Count slots in DeviceBufs::counts.
// C_HROWS / C_HNODES accumulate over a whole round (rows and nodes
// histogrammed), for the algorithmic-byte accounting.
// C_TICKET is the one-pass partition's tile ticket (reset by each plan);
// C_ERR is set when a partition look-back did not finish (reset per round).
enum : int {
  C_ITEMS = 0, C_TILES, C_HIST, C_FIN, C_SPLIT, C_LSEGS, C_HROWS, C_HNODES, C_TICKET, C_ERR,
  C_NCOUNTS
};

// Per-round outputs, laid out back to back in one device buffer so that one
// copy brings a round's trees to the host.
struct TreeOut {
  TreeCtl* ctl;                                   // [T]
  uint32_t* counts;                               // [C_NCOUNTS]
  uint32_t *tr_split, *tr_left, *tr_right;        // [T][MNT]
  uint32_t* tr_cats;                              // [T][MNT][8]
  float *tr_thr, *tr_leaf, *tr_gain, *tr_cover;   // [T][MNT]
  uint8_t *tr_dleft, *tr_isleaf, *tr_stype;       // [T][MNT]
};
size_t tree_out_layout(uint32_t T, uint32_t MNT, char* base, TreeOut* out);

struct DeviceBufs {
  TreeCtl* ctl;
  uint32_t *open_node, *open_tnode, *open_depth;  // [T][WO]
  uint32_t *next_node, *next_tnode, *next_depth;  // [T][2W]
  uint32_t *ch_node, *ch_tnode, *ch_depth;        // [T][W]
  uint32_t* pbegin;                               // [T*MN] global positions
  uint32_t* pslot;                                // [T*MN] tree-local slot, ~0 = none
  uint8_t* pbuf;                                  // [T*MN] bit 0: rows in the alternate buffer,
                                                  // bit 1: rows stored from the last position down
  uint32_t* free_slot;                            // [T][SPT] stack
  float* leaf_val;                                // [T*MN]
  // Tree arrays [T][MNT] (cats: [T][MNT][8]).
  uint32_t *tr_split, *tr_left, *tr_right, *tr_cats;
  float *tr_thr, *tr_leaf, *tr_gain, *tr_cover;
  uint8_t *tr_dleft, *tr_isleaf, *tr_stype;
  // Step work lists.
  PartItem* items;
  PartTile* tiles;
  unsigned long long* tstat;  // [tiles] look-back status of the step's partition tiles
  uint32_t* pend;  // [items][5]: tree, left tree node, right tree node, rank, child depth
  HistItem* hitems;
  uint32_t* fin;   // [n][4]: small slot, large slot or ~0, first item, tiles
  uint32_t *split_nodes, *split_slots;
  // Per split entry: the slot holding the smaller sibling's histogram, which
  // the split subtracts from the entry's slot (the parent's histogram, kept
  // by the larger child) and writes back; kNone for a histogram that is
  // complete.
  uint32_t* split_minus;
  Segment* leaf_segs;
  uint32_t* counts;  // [C_NCOUNTS]
  // Engine buffers shared with the trainer.
  long long* node_sum;
  uint32_t* node_count;
  DevSplit* node_split;
  unsigned long long* node_key;  // packed best (gain, feature) per node
  const DevSplit* cand;          // [node][E] per-feature best
  const float* thr_up;           // [E][256]
  const uint32_t* eff_to_orig;   // [E]
  const FeatMeta* fmeta;         // [E]
  const int32_t* qexp;           // [2K]
  uint32_t* cursor;              // [items][2] partition cursors; [2 i] ends as the left count
};

// Also the round's quantisation scales (what launch_qscale does: qexp from
// absmax[0..n_q), absmax reset).  rows_per_tree null: every tree has all rows.
void launch_ctl_init(hipStream_t s, const CtlParams& p, const DeviceBufs& b,
                     const uint32_t* rows_per_tree, uint32_t* absmax, int n_q, int64_t n_rows,
                     int32_t* qexp);
// advance: depth-wise, the previous step's children form the level to plan.
void launch_ctl_plan(hipStream_t s, const CtlParams& p, const DeviceBufs& b, bool advance);
void launch_ctl_children(hipStream_t s, const CtlParams& p, const DeviceBufs& b);
// Also advances the boosting-round counter read by the sampling kernels.
// advance: 0 the open set remains; 1 depth-wise, the last step's children
// (next level) remain; 2 the last step's partition applied its children's
// values (leaf_apply) and its children step was skipped: the items'
// children become leaves here.
void launch_ctl_finalize(hipStream_t s, const CtlParams& p, const DeviceBufs& b, int advance,
                         uint32_t* round);
void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, int64_t n);
// Element copy by a kernel; dst may be mapped pinned host memory (device
// pointer from hipHostGetDevicePointer).
void launch_copy_u32(hipStream_t s, uint32_t* dst, const uint32_t* src, int64_t n);

}  // namespace bst
