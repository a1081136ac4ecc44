// grow_device.hip — controller kernels of the device-side grower.
//
// These single-block kernels restate, on the GPU, the host control of
// trainer.cpp (itself restating training/gbdt/{grower,expansion}.rs):
//   plan      = GrowthState::pop_next + should_expand + apply_split_to_builder
//               + RowPartitioner::split id allocation (grower.rs:346-394,
//               530-650; expansion.rs:150-231), preceded by the best-split
//               selection of the previous step's nodes (find.rs per-feature
//               bests reduced across features)
//   children  = the subtraction-trick bookkeeping of process_candidate
//               (grower.rs:396-480): smaller child built, larger = parent - smaller
//   finalize  = finalize_remaining + apply_learning_rate (grower.rs:484-514,
//               mutable_tree.rs:267-273) and the partition-range prediction
//               update list (grower.rs:208-219)
// Decisions, orders and ids are the same as the host grower's, so the
// forests are identical (tested).
#include <math.h>

#include "grow_device.h"

namespace bst {

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kDirectFlag = 0x80000000u;  // HistItem.tree: tile = whole node
constexpr uint32_t kAltFlag = 0x40000000u;     // HistItem.tree: alternate buffer
constexpr int kCtlThreads = 1024;
constexpr uint32_t kPlanLdsItems = 1024;  // items whose partition tiles the plan fills block-wide
static_assert(kCtlThreads <= kPlanLdsItems, "a paired level records every item in LDS");

// Exclusive scan over the block (kCtlThreads threads); *total receives the sum.
__device__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_tmp, uint32_t* total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(x, d, 64);
    if (lane >= d) x += o;
  }
  if (lane == 63) s_tmp[wv] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int i = 0; i < kCtlThreads / 64; i++) {
      const uint32_t t = s_tmp[i];
      s_tmp[i] = s;
      s += t;
    }
    s_tmp[kCtlThreads / 64] = s;
  }
  __syncthreads();
  const uint32_t r = x - v + s_tmp[wv];
  *total = s_tmp[kCtlThreads / 64];
  __syncthreads();
  return r;
}

// Gain as an unsigned key with the float order (-inf lowest).
__device__ __forceinline__ uint32_t ord_gain(float v) {
  const uint32_t u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// PartTile::src of a node's tiles from its pbuf bits (bit 0 alternate
// buffer, bit 1 rows stored from the last position down).
__device__ __forceinline__ uint32_t tile_src(uint32_t pbuf, bool leaves) {
  return ((pbuf & 1u) ? kTileAlt : 0u) | ((pbuf & 2u) ? kTileRev : 0u) |
         (leaves ? kTileLeaves : 0u);
}

// Position of the node's row b0 (in row order) for a tile of `src`.
__device__ __forceinline__ uint32_t tile_begin(uint32_t begin, uint32_t count, uint32_t b0,
                                               uint32_t src) {
  return (src & kTileRev) ? begin + count - 1u - b0 : begin + b0;
}

__device__ __forceinline__ bool expandable(const DevSplit& s, double min_gain) {
  return s.gain > -INFINITY && (double)s.gain > min_gain;
}

__device__ void init_tree_node(const CtlParams& P, const DeviceBufs& B, uint32_t t, uint32_t tn) {
  const uint64_t k = (uint64_t)t * P.MNT + tn;
  B.tr_split[k] = 0;
  B.tr_left[k] = 0;
  B.tr_right[k] = 0;
  B.tr_thr[k] = 0.0f;
  B.tr_leaf[k] = 0.0f;
  B.tr_gain[k] = 0.0f;
  B.tr_cover[k] = 0.0f;
  B.tr_dleft[k] = 0;
  B.tr_isleaf[k] = 0;
  B.tr_stype[k] = 0;
  for (int w = 0; w < kMaxCatWords; w++) B.tr_cats[k * kMaxCatWords + w] = 0;
}

// compute_leaf_weight (split/gain.rs:151-168) from exact fixed-point sums,
// with the learning rate applied (apply_learning_rate scales every leaf by
// the same float product once the tree is done).
__device__ float leaf_weight(const CtlParams& P, long long sg, long long sh, int qg, int qh) {
  const double G = ldexp((double)sg, -qg);
  const double H = ldexp((double)sh, -qh);
  float w;
  if (P.alpha == 0.0) {
    w = (float)(-G / (H + P.lambda));
  } else {
    const double ag = fabs(G);
    w = ag <= P.alpha ? 0.0f : (float)((G > 0.0 ? -1.0 : 1.0) * (ag - P.alpha) / (H + P.lambda));
  }
  return w * P.lr;
}

// The node's sums (sg, sh), the round's scale exponents (qg, qh) and its
// histogram slot are passed in already loaded.
__device__ void make_leaf_from(const CtlParams& P, const DeviceBufs& B, uint32_t t, uint32_t gid,
                               uint32_t tn, long long sg, long long sh, int qg, int qh,
                               uint32_t slot) {
  const float wl = leaf_weight(P, sg, sh, qg, qh);
  const double H = ldexp((double)sh, -qh);
  const uint64_t k = (uint64_t)t * P.MNT + tn;
  B.tr_isleaf[k] = 1;
  B.tr_leaf[k] = wl;
  B.tr_cover[k] = (float)H;
  B.leaf_val[gid] = wl;
  if (slot != kNone) {
    const uint32_t pos = atomicAdd(&B.ctl[t].n_free, 1u);
    B.free_slot[(uint64_t)t * P.SPT + pos] = slot;
    B.pslot[gid] = kNone;
  }
}

__device__ void make_leaf(const CtlParams& P, const DeviceBufs& B, uint32_t t, uint32_t p,
                          uint32_t tn) {
  const uint32_t gid = t * P.MN + p;
  const int out = P.out0 + (int)t;
  make_leaf_from(P, B, t, gid, tn, B.node_sum[2 * gid], B.node_sum[2 * gid + 1],
                 B.qexp[2 * out], B.qexp[2 * out + 1], B.pslot[gid]);
}

// Depth-wise: the previous step's children become the level to plan.
__device__ void advance_level(const CtlParams& P, const DeviceBufs& B) {
  for (uint32_t t = 0; t < P.T; t++) {
    const uint32_t nn = B.ctl[t].n_next;
    for (uint32_t i = threadIdx.x; i < nn; i += blockDim.x) {
      const uint64_t s = (uint64_t)t * 2 * P.W + i, d = (uint64_t)t * P.WO + i;
      B.open_node[d] = B.next_node[s];
      B.open_tnode[d] = B.next_tnode[s];
      B.open_depth[d] = B.next_depth[s];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      TreeCtl& c = B.ctl[t];
      c.n_open = nn;
      c.n_next = 0;
      c.depth += 1;
    }
    __syncthreads();
  }
}

// Histogram tile of a step: about hist_target (tile, group) items for `rows`
// rows, so small steps still spread over the CUs and large ones do not make
// needless partials.
__device__ uint32_t step_tile(const CtlParams& P, uint64_t rows) {
  return hist_step_tile(P.hist_tile_min, P.hist_tile, P.hist_target, P.NG, rows);
}

__device__ unsigned long long block_sum_u64(unsigned long long v, unsigned long long* s_w) {
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  unsigned long long r = 0ull;
  for (int i = 0; i < kCtlThreads / 64; i++) r += s_w[i];
  __syncthreads();
  return r;
}

}  // namespace

// Root state of every tree, root histogram work, its finish entry and the
// root split list.  Root totals are added by the row initialisation.
__global__ __launch_bounds__(kCtlThreads) void k_ctl_init(CtlParams P, DeviceBufs B,
                                                          const uint32_t* __restrict__ rows,
                                                          uint32_t* __restrict__ absmax, int n_q,
                                                          int64_t n_rows, int32_t* __restrict__ qexp) {
  __shared__ uint32_t s_hist;
  if (threadIdx.x == 0) s_hist = 0;
  // The round's quantisation scales; absmax is then ready for the next round.
  if ((int)threadIdx.x < n_q) {
    qexp[threadIdx.x] = qscale_exponent(__uint_as_float(absmax[threadIdx.x]), n_rows);
    absmax[threadIdx.x] = 0u;
  }
  auto rows_of = [&](uint32_t t) { return rows ? rows[t] : (uint32_t)P.stride; };
  uint64_t all_rows = 0;
  for (uint32_t t = 0; t < P.T; t++) all_rows += rows_of(t);
  const uint32_t tile = step_tile(P, all_rows);
  __syncthreads();
  for (uint32_t t = 0; t < P.T; t++) {
    for (uint32_t p = threadIdx.x; p < P.MN; p += blockDim.x) {
      B.pslot[t * P.MN + p] = kNone;
      B.leaf_val[t * P.MN + p] = NAN;
    }
    for (uint32_t i = threadIdx.x; i + 1 < P.SPT; i += blockDim.x)
      B.free_slot[(uint64_t)t * P.SPT + i] = P.SPT - 1 - i;  // slot 0 goes to the root
    __syncthreads();
    const uint32_t m = rows_of(t);
    const uint32_t nt = hist_node_tiles(m, tile, P.hist_tile);
    const uint32_t ts = hist_node_tile_rows(m, nt);
    if (threadIdx.x == 0) {
      TreeCtl c;
      c.n_part = 1;
      c.n_tnodes = 1;
      c.n_leaves = 1;
      c.n_open = 1;
      c.n_next = 0;
      c.n_free = P.SPT - 1;
      c.depth = 0;
      c.rows = m;
      B.ctl[t] = c;
      const uint32_t gid = t * P.MN;
      B.pslot[gid] = 0;
      B.pbuf[gid] = 0;
      B.pbegin[gid] = (uint32_t)(t * P.stride);
      B.node_count[gid] = m;
      B.node_sum[2 * gid] = 0;
      B.node_sum[2 * gid + 1] = 0;
      init_tree_node(P, B, t, 0);
      B.open_node[(uint64_t)t * P.WO] = 0;
      B.open_tnode[(uint64_t)t * P.WO] = 0;
      B.open_depth[(uint64_t)t * P.WO] = 0;
      B.split_nodes[t] = gid;
      B.split_slots[t] = t * P.SPT;
      B.split_minus[t] = kNone;
      B.fin[4 * t + 0] = t * P.SPT;
      B.fin[4 * t + 1] = kNone;
      B.fin[4 * t + 2] = s_hist;
      B.fin[4 * t + 3] = nt;
    }
    __syncthreads();
    const uint32_t base = s_hist;
    for (uint32_t k = threadIdx.x; k < nt * P.NG; k += blockDim.x) {
      const uint32_t q = k / P.NG, g = k % P.NG;
      const uint32_t b0 = q * ts;
      HistItem it;
      it.node = t * P.MN;
      it.slot = t * P.SPT;
      it.begin = (uint32_t)(t * P.stride) + b0;
      it.count = min(ts, m - b0);
      it.group = g;
      it.tree = t | (nt == 1 ? kDirectFlag : 0u);
      B.hitems[base + k] = it;
    }
    __syncthreads();
    if (threadIdx.x == 0) s_hist += nt * P.NG;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int i = 0; i < C_NCOUNTS; i++) B.counts[i] = 0;
    B.counts[C_HIST] = s_hist;
    B.counts[C_FIN] = P.T;
    B.counts[C_SPLIT] = P.T;
    uint32_t r = 0;
    for (uint32_t t = 0; t < P.T; t++) r += rows_of(t);
    B.counts[C_HROWS] = r;
    B.counts[C_HNODES] = P.T;
  }
}

// An expanded node (global id gid, tree node tn at depth dep): its tree
// record, both children's tree nodes a and a + 1, partition item g (tiles
// tile0 .. tile0 + ntile - 1, right child's partition id `right`) and the
// record the children step reads back (tree, children's tree nodes, e, depth).
// The split feature's original index, threshold (numerical splits) and
// missing-value bin (kNone without missing values) are passed in loaded.
// With `leaves` (the children are leaves and the partition applies them),
// the children's values come from the split's left sums and the parent's
// totals (sg, sh), as the leaves made at the end of the tree would compute
// them from the children's sums.
__device__ void emit_expand(const CtlParams& P, const DeviceBufs& B, uint32_t t, uint32_t gid,
                            uint32_t tn, uint32_t dep, const DevSplit& sp, double H,
                            uint32_t count, uint32_t begin, uint32_t pbuf, uint32_t right,
                            uint32_t a, uint32_t e, uint32_t g, uint32_t tile0, uint32_t ntile,
                            uint32_t orig, float thr, uint32_t miss_bin, bool leaves,
                            long long sg, long long sh, int qg, int qh) {
  const uint32_t b = a + 1;
  const uint64_t k = (uint64_t)t * P.MNT + tn;
  B.tr_gain[k] = sp.gain;
  B.tr_cover[k] = (float)H;
  B.tr_split[k] = orig;
  B.tr_left[k] = a;
  B.tr_right[k] = b;
  uint32_t lt, rt;
  if (!(sp.flags & 2u)) {
    B.tr_thr[k] = thr;
    B.tr_dleft[k] = (uint8_t)(sp.flags & 1u);
    B.tr_stype[k] = 0;
    lt = a;
    rt = b;
  } else {
    B.tr_thr[k] = 0.0f;
    B.tr_dleft[k] = (sp.flags & 1u) ? 0 : 1;
    B.tr_stype[k] = 1;
    for (int w = 0; w < kMaxCatWords; w++) B.tr_cats[k * kMaxCatWords + w] = sp.cats[w];
    lt = b;  // training-left rows take the inference-right child
    rt = a;
  }
  init_tree_node(P, B, t, a);
  init_tree_node(P, B, t, b);
  PartItem it;
  it.node = gid;
  it.begin = begin;
  it.count = count;
  it.tile0 = tile0;
  it.n_tiles = ntile;
  it.right_node = t * P.MN + right;
  it.feature = sp.feature;
  it.pad = pbuf;
  it.split_bin = sp.bin;
  it.split_flags = sp.flags | (leaves ? kItemLeaves : 0u);
  it.miss_bin = miss_bin;
  it.pad2 = 0;
  it.leaf_l = leaves ? leaf_weight(P, sp.lg, sp.lh, qg, qh) : 0.0f;
  it.leaf_r = leaves ? leaf_weight(P, sg - sp.lg, sh - sp.lh, qg, qh) : 0.0f;
  it.pad3 = 0;
  it.pad4 = 0;
  B.items[g] = it;
  B.cursor[2 * g] = 0;
  B.cursor[2 * g + 1] = 0;
  B.pend[5 * g + 0] = t;
  B.pend[5 * g + 1] = lt;
  B.pend[5 * g + 2] = rt;
  B.pend[5 * g + 3] = e;
  B.pend[5 * g + 4] = dep + 1;
}

// Original index, threshold (numerical bins only; any bin reads in bounds)
// and missing-value bin of the split feature.
__device__ __forceinline__ void split_feature_info(const DeviceBufs& B, const DevSplit& sp,
                                                   uint32_t* orig, float* thr,
                                                   uint32_t* miss_bin) {
  const uint32_t f = sp.feature;
  *orig = B.eff_to_orig[f];
  *thr = B.thr_up[(uint64_t)f * 256 + (sp.bin & 255u)];
  const FeatMeta fm = B.fmeta[f];
  *miss_bin = fm.has_missing ? fm.default_bin : kNone;
}

// The best split of a listed node from its packed key (best gain and winning
// feature under the tie rule); no candidate (key 0) gives an invalid split.
__device__ __forceinline__ DevSplit split_from_key(const CtlParams& P, const DeviceBufs& B,
                                                   uint32_t node, unsigned long long key) {
  DevSplit sp;
  if (key == 0) {
    sp.gain = -INFINITY;
    sp.feature = 0;
    sp.bin = 0;
    sp.flags = 0;
    for (int w = 0; w < kMaxCatWords; w++) sp.cats[w] = 0;
    sp.lg = 0;
    sp.lh = 0;
  } else {
    const uint32_t lo = (uint32_t)(key & 0xFFFFFFFFu);
    const uint32_t f = P.rightmost ? lo : (0xFFFFFFFFu - lo);
    sp = B.cand[(uint64_t)node * P.E + f];
  }
  return sp;
}

// Depth-wise growth of one tree.  The children step lists a level's nodes in
// the same order in the split list and in the next-level list (entries 2i,
// 2i + 1 are the children of item i in both), so thread i carries node i from
// its split candidates to its tree nodes and partition item in registers,
// with every load that does not depend on the split issued together.
// Returns false, having written nothing, when the two lists do not pair up
// or do not fit one block; the general path then runs.
__device__ bool plan_level_paired(const CtlParams& P, const DeviceBufs& B, int advance,
                                  uint32_t n_sp, uint32_t* s_tmp, uint32_t* s_any,
                                  uint32_t* s_tfirst, uint32_t* s_tbegin, uint32_t* s_tcount,
                                  uint8_t* s_tsrc,
                                  uint32_t* s_items, uint32_t* s_tiles, uint32_t* s_ptile) {
  const uint32_t i = threadIdx.x;
  const TreeCtl c0 = B.ctl[0];
  const int qg = B.qexp[2 * P.out0], qh = B.qexp[2 * P.out0 + 1];
  const uint32_t nn = advance ? c0.n_next : c0.n_open;
  if (nn > (uint32_t)kCtlThreads || n_sp > (uint32_t)kCtlThreads) return false;
  // Every thread loads (entry 0 past the lists; each list holds at least
  // one entry) so that the loads issue together.
  const bool in = i < nn, ins = i < n_sp;
  const uint32_t li = in ? i : 0u;
  const uint32_t p0 = (advance ? B.next_node : B.open_node)[li];
  const uint32_t tn = (advance ? B.next_tnode : B.open_tnode)[li];
  const uint32_t dep = (advance ? B.next_depth : B.open_depth)[li];
  const uint32_t sn = B.split_nodes[ins ? i : 0u];
  // One tree: partition ids are global ids.
  const unsigned long long wb = __ballot(in != ins || (in && sn != p0));
  if ((i & 63) == 0) s_any[i >> 6] = wb != 0ull ? 1u : 0u;
  __syncthreads();
  uint32_t any = 0;
#pragma unroll
  for (int w = 0; w < kCtlThreads / 64; w++) any |= s_any[w];
  if (any) return false;
  const uint32_t p = in ? p0 : 0u;
  const unsigned long long key = B.node_key[p];
  const uint32_t count = B.node_count[p], begin = B.pbegin[p], pbuf = B.pbuf[p];
  const uint32_t slot = B.pslot[p];
  const long long sg = B.node_sum[2 * p], sh = B.node_sum[2 * p + 1];
  const DevSplit sp = split_from_key(P, B, p, key);
  // Split-dependent loads ahead of the stores and the scans' barriers.
  uint32_t orig, miss;
  float thr;
  split_feature_info(B, sp, &orig, &thr, &miss);
  if (in) {
    B.node_key[p] = 0;
    B.node_split[p] = sp;
  }
  const uint32_t depth = c0.depth + (advance ? 1u : 0u);
  const bool act = in && depth < P.max_depth;
  const bool ex = act && expandable(sp, P.min_gain);
  // A step whose partition applies the children's leaf values reads and
  // writes every row's prediction at random: shorter tiles put more blocks
  // on those accesses.  Steps that move rows take full tiles (fewer tiles,
  // fewer look-back rounds).
  const bool lv_step = P.leaf_apply && depth + 1u >= P.max_depth;
  const uint32_t ptile = lv_step ? P.part_tile_small : P.part_tile;
  const uint32_t ntile = ex ? max(1u, (count + ptile - 1) / ptile) : 0u;
  uint32_t tot_e, tot_t;
  const uint32_t e = block_excl_scan(ex ? 1u : 0u, s_tmp, &tot_e);
  const uint32_t toff = block_excl_scan(ntile, s_tmp, &tot_t);
  if (act && !ex) make_leaf_from(P, B, 0, p, tn, sg, sh, qg, qh, slot);
  if (ex) {
    const bool lv = P.leaf_apply && dep + 1u >= P.max_depth;
    emit_expand(P, B, 0, p, tn, dep, sp, ldexp((double)sh, -qh), count, begin, pbuf,
                c0.n_part + e, c0.n_tnodes + 2 * e, e, e, toff, ntile, orig, thr, miss, lv, sg,
                sh, qg, qh);
    s_tfirst[e] = toff;  // e < kCtlThreads <= kPlanLdsItems
    s_tbegin[e] = begin;
    s_tcount[e] = count;
    s_tsrc[e] = (uint8_t)tile_src(pbuf, lv);
  }
  if (i == 0) {
    // The level is consumed (advance_level + the step's updates); n_free is
    // left to the leaves' atomics.
    TreeCtl& c = B.ctl[0];
    c.depth = depth;
    c.n_open = 0;
    c.n_next = 0;
    c.n_part = c0.n_part + tot_e;
    c.n_tnodes = c0.n_tnodes + 2 * tot_e;
    *s_items = tot_e;
    *s_tiles = tot_t;
    *s_ptile = ptile;
  }
  __syncthreads();
  return true;
}

// Pick each listed node's best split from its per-feature candidates (the
// packed key holds the best gain and the winning feature under the tie
// rule), then choose the nodes to expand this step, decide leaf / split,
// allocate ids, write the tree nodes and the partition work list.
// Leaf-wise step: every tree pops its open node of max (gain, partition id)
// (BinaryHeap order of NodeCandidate) and, when that node splits, yields one
// item.  Trees are independent, so wave t handles tree t and the T pops run
// side by side; only the item and tile offsets are a prefix over the trees,
// taken in tree order, so items, ids and tiles are those of a tree-by-tree
// walk.  The popped node's state is loaded by the lane that found it, right
// after the argmax, to keep the dependent global loads per tree few.
__device__ void plan_leafwise(const CtlParams& P, const DeviceBufs& B, uint32_t* s_ex,
                              uint32_t* s_nt, uint32_t* s_tfirst, uint32_t* s_tbegin,
                              uint32_t* s_tcount, uint8_t* s_tsrc, uint32_t* s_items,
                              uint32_t* s_tiles) {
  const uint32_t lane = threadIdx.x & 63u, t = threadIdx.x >> 6;
  const bool tree = t < P.T;
  const uint64_t ob = (uint64_t)t * P.WO;
  bool pop = false;
  uint32_t p = 0, tn = 0, dep = 0, n_part = 0, n_tnodes = 0, count = 0, ntile = 0;
  DevSplit sp;
  if (tree) {
    const TreeCtl c = B.ctl[t];
    unsigned long long best = 0ull;
    uint32_t bi = 0;
    for (uint32_t i = lane; i < c.n_open; i += 64) {
      const uint32_t q = B.open_node[ob + i];
      const float g = B.node_split[t * P.MN + q].gain + 0.0f;  // -0 ties with +0
      const unsigned long long key = ((unsigned long long)ord_gain(g) << 32) | q;
      if (key > best) {
        best = key;
        bi = i;
      }
    }
    unsigned long long wbest = best;
    for (int m = 32; m >= 1; m >>= 1) {
      const unsigned long long o = __shfl_xor(wbest, m, 64);
      wbest = o > wbest ? o : wbest;
    }
    // Partition ids are unique in the open set, so one lane holds the max.
    const unsigned long long hit = __ballot(c.n_open > 0 && best == wbest);
    const int src = hit ? __ffsll((long long)hit) - 1 : 0;
    bi = __shfl(bi, src, 64);
    pop = c.n_leaves < P.max_leaves && c.n_open > 0;
    if (pop && lane == 0) {
      p = (uint32_t)(wbest & 0xFFFFFFFFull);
      const uint32_t last = c.n_open - 1;
      tn = B.open_tnode[ob + bi];
      dep = B.open_depth[ob + bi];
      const uint32_t ln = B.open_node[ob + last], lt = B.open_tnode[ob + last],
                     ld = B.open_depth[ob + last];
      sp = B.node_split[t * P.MN + p];
      count = B.node_count[t * P.MN + p];
      B.open_node[ob + bi] = ln;
      B.open_tnode[ob + bi] = lt;
      B.open_depth[ob + bi] = ld;
      B.ctl[t].n_open = last;
      B.ctl[t].n_leaves = c.n_leaves + 1;
      n_part = c.n_part;
      n_tnodes = c.n_tnodes;
      const bool ex = expandable(sp, P.min_gain);
      ntile = ex ? max(1u, (count + P.part_tile - 1) / P.part_tile) : 0u;
      s_ex[t] = ex ? 1u : 0u;
      s_nt[t] = ntile;
    } else if (lane == 0) {
      s_ex[t] = 0u;
      s_nt[t] = 0u;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t e = 0, k = 0;
    for (uint32_t u = 0; u < P.T; u++) {
      const uint32_t x = s_ex[u], y = s_nt[u];
      s_ex[u] = e;
      s_nt[u] = k;
      e += x;
      k += y;
    }
    *s_items = e;
    *s_tiles = k;
  }
  __syncthreads();
  if (tree && pop && lane == 0) {
    if (ntile == 0) {
      make_leaf(P, B, t, p, tn);
    } else {
      const uint32_t gid = t * P.MN + p;
      const int out = P.out0 + (int)t;
      const long long sg = B.node_sum[2 * gid], sh = B.node_sum[2 * gid + 1];
      const int qg = B.qexp[2 * out], qh = B.qexp[2 * out + 1];
      const uint32_t begin = B.pbegin[gid], pb = B.pbuf[gid];
      uint32_t orig, miss;
      float thr;
      split_feature_info(B, sp, &orig, &thr, &miss);
      const uint32_t g = s_ex[t], tile0 = s_nt[t];
      emit_expand(P, B, t, gid, tn, dep, sp, ldexp((double)sh, -qh), count, begin, pb, n_part,
                  n_tnodes, 0u, g, tile0, ntile, orig, thr, miss, false, sg, sh, qg, qh);
      s_tfirst[g] = tile0;
      s_tbegin[g] = begin;
      s_tcount[g] = count;
      s_tsrc[g] = (uint8_t)tile_src(pb, false);
      B.ctl[t].n_part = n_part + 1;
      B.ctl[t].n_tnodes = n_tnodes + 2;
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(kCtlThreads) void k_ctl_plan(CtlParams P, DeviceBufs B,
                                                          int advance) {
  __shared__ uint32_t s_tmp[kCtlThreads / 64 + 1];
  __shared__ uint32_t s_items, s_tiles, s_nch;
  __shared__ uint32_t s_part, s_tnodes;
  // Partition tiles of the first kPlanLdsItems items, written by the whole
  // block at the end: first tile, row range of each item.
  __shared__ uint32_t s_tfirst[kPlanLdsItems], s_tbegin[kPlanLdsItems], s_tcount[kPlanLdsItems];
  __shared__ uint8_t s_tsrc[kPlanLdsItems];  // PartTile::src bits of the item's tiles
  __shared__ uint32_t s_tlds_end;
  __shared__ uint32_t s_any[kCtlThreads / 64];
  __shared__ uint32_t s_ptile;  // the step's partition tile
  __shared__ uint32_t s_lw_ex[kCtlMaxTrees], s_lw_nt[kCtlMaxTrees];
  if (threadIdx.x == 0) {
    s_tlds_end = kNone;
    s_ptile = P.part_tile;
  }
  const uint32_t n_sp = B.counts[C_SPLIT];
  const bool paired = P.growth == 0 && P.T == 1 &&
                      plan_level_paired(P, B, advance, n_sp, s_tmp, s_any, s_tfirst,
                                        s_tbegin, s_tcount, s_tsrc, &s_items, &s_tiles,
                                        &s_ptile);
  if (!paired) {
    for (uint32_t i = threadIdx.x; i < n_sp; i += blockDim.x) {
      const uint32_t node = B.split_nodes[i];
      const unsigned long long key = B.node_key[node];
      B.node_key[node] = 0;
      B.node_split[node] = split_from_key(P, B, node, key);
    }
    __syncthreads();
    if (advance && P.growth == 0) advance_level(P, B);
    if (threadIdx.x == 0) {
      s_items = 0;
      s_tiles = 0;
    }
    __syncthreads();
    if (P.growth == 1)
      plan_leafwise(P, B, s_lw_ex, s_lw_nt, s_tfirst, s_tbegin, s_tcount, s_tsrc, &s_items,
                    &s_tiles);
    // Depth-wise: every tree's open level, tree by tree.
    for (uint32_t t = 0; t < (P.growth == 1 ? 0u : P.T); t++) {
      const uint64_t ob = (uint64_t)t * P.WO, cb = (uint64_t)t * P.W;
      if (threadIdx.x == 0) {
        const TreeCtl& c = B.ctl[t];
        s_nch = c.depth < P.max_depth ? c.n_open : 0u;
        s_part = c.n_part;
        s_tnodes = c.n_tnodes;
      }
      __syncthreads();
      const uint32_t nch = s_nch;
      for (uint32_t i = threadIdx.x; i < nch; i += blockDim.x) {
        B.ch_node[cb + i] = B.open_node[ob + i];
        B.ch_tnode[cb + i] = B.open_tnode[ob + i];
        B.ch_depth[cb + i] = B.open_depth[ob + i];
      }
      __syncthreads();
      uint32_t carry_e = 0, carry_t = 0;
      for (uint32_t base = 0; base < nch; base += blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const bool in = i < nch;
        uint32_t p = 0, tn = 0, dep = 0, count = 0;
        DevSplit sp;
        bool ex = false;
        if (in) {
          p = B.ch_node[cb + i];
          tn = B.ch_tnode[cb + i];
          dep = B.ch_depth[cb + i];
          sp = B.node_split[t * P.MN + p];
          ex = expandable(sp, P.min_gain);
          count = B.node_count[t * P.MN + p];
        }
        const uint32_t ntile = ex ? max(1u, (count + P.part_tile - 1) / P.part_tile) : 0u;
        uint32_t tot_e, tot_t;
        const uint32_t e = block_excl_scan(ex ? 1u : 0u, s_tmp, &tot_e) + carry_e;
        const uint32_t toff = block_excl_scan(ntile, s_tmp, &tot_t) + carry_t;
        if (in && !ex) make_leaf(P, B, t, p, tn);
        if (in && ex) {
          const uint32_t gid = t * P.MN + p;
          const int out = P.out0 + (int)t;
          const long long sg = B.node_sum[2 * gid], sh = B.node_sum[2 * gid + 1];
          const int qg = B.qexp[2 * out], qh = B.qexp[2 * out + 1];
          const double H = ldexp((double)sh, -qh);
          const uint32_t g = s_items + e;
          const uint32_t begin = B.pbegin[gid], pb = B.pbuf[gid];
          const bool lv = P.leaf_apply && P.growth == 0 && dep + 1u >= P.max_depth;
          const uint32_t src = tile_src(pb, lv);
          uint32_t orig, miss;
          float thr;
          split_feature_info(B, sp, &orig, &thr, &miss);
          emit_expand(P, B, t, gid, tn, dep, sp, H, count, begin, pb, s_part + e,
                      s_tnodes + 2 * e, e, g, s_tiles + toff, ntile, orig, thr, miss, lv, sg, sh,
                      qg, qh);
          if (g < kPlanLdsItems) {
            s_tfirst[g] = s_tiles + toff;
            s_tbegin[g] = begin;
            s_tcount[g] = count;
            s_tsrc[g] = (uint8_t)src;
          } else {
            if (g == kPlanLdsItems) s_tlds_end = s_tiles + toff;
            for (uint32_t q = 0; q < ntile; q++) {
              const uint32_t b0 = q * P.part_tile;
              PartTile pt;
              pt.item = g;
              pt.begin = tile_begin(begin, count, b0, src);
              pt.count = count > b0 ? min(P.part_tile, count - b0) : 0u;
              pt.src = src;
              B.tiles[s_tiles + toff + q] = pt;
              B.tstat[s_tiles + toff + q] = 0ull;
            }
          }
        }
        carry_e += tot_e;
        carry_t += tot_t;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        TreeCtl& c = B.ctl[t];
        c.n_part += carry_e;
        c.n_tnodes += 2 * carry_e;
        if (P.growth == 0) {
          c.n_open = 0;
          c.n_next = 0;
        }
        s_items += carry_e;
        s_tiles += carry_t;
      }
      __syncthreads();
    }
  }
  // Tiles of the items recorded in LDS: tile k belongs to the last item g
  // with s_tfirst[g] <= k (every item has at least one tile).
  const uint32_t n_lds_items = min(s_items, kPlanLdsItems);
  const uint32_t end_k = min(s_tiles, s_tlds_end);
  for (uint32_t k = threadIdx.x; k < end_k; k += blockDim.x) {
    uint32_t lo = 0, hi = n_lds_items - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (s_tfirst[mid] <= k) lo = mid; else hi = mid - 1;
    }
    const uint32_t ptile = s_ptile;
    const uint32_t b0 = (k - s_tfirst[lo]) * ptile, count = s_tcount[lo];
    PartTile pt;
    pt.item = lo;
    pt.src = s_tsrc[lo];
    pt.begin = tile_begin(s_tbegin[lo], count, b0, pt.src);
    pt.count = count > b0 ? min(ptile, count - b0) : 0u;
    B.tiles[k] = pt;
    B.tstat[k] = 0ull;
  }
  if (threadIdx.x == 0) {
    B.counts[C_ITEMS] = s_items;
    B.counts[C_TILES] = s_tiles;
    B.counts[C_TICKET] = 0u;
  }
}

// One item of a children step and its parent's state: pending record (tree,
// children's tree nodes, index in the tree, children's depth), the item, its
// left count, and the parent's left sums, totals, first row and slot.
struct ChildIn {
  PartItem it;
  uint32_t t, lt, rt, e, cdepth, L, R, pb, ps;
  long long lg, lh, pg, ph;
};

// Nothing is loaded for an invalid item (all fields zero).  Loading the
// parent's state early is safe: an item's writes touch only its own parent
// and right child, and right children are fresh ids.
__device__ __forceinline__ ChildIn child_load(const DeviceBufs& B, uint32_t i, bool valid) {
  ChildIn c{};
  if (valid) {
    c.t = B.pend[5 * i];
    c.lt = B.pend[5 * i + 1];
    c.rt = B.pend[5 * i + 2];
    c.e = B.pend[5 * i + 3];
    c.cdepth = B.pend[5 * i + 4];
    c.it = B.items[i];
    c.L = B.cursor[2 * i];
    c.R = c.it.count - c.L;
    const uint32_t parent = c.it.node;
    c.lg = B.node_split[parent].lg;
    c.lh = B.node_split[parent].lh;
    c.pg = B.node_sum[2 * parent];
    c.ph = B.node_sum[2 * parent + 1];
    c.pb = B.pbegin[parent];
    c.ps = B.pslot[parent];
  }
  return c;
}

// After the partition: child counts, slots, the smaller child's histogram
// items, the finish list (partials + subtraction), the split list, and the
// children pushed to the open set (leaf-wise) or the next level (depth-wise).
constexpr uint32_t kChildLdsItems = 1024;  // items whose histogram items are filled block-wide

__global__ __launch_bounds__(kCtlThreads) void k_ctl_children(CtlParams P, DeviceBufs B) {
  __shared__ unsigned long long s_red64[kCtlThreads / 64];
  __shared__ uint32_t s_tmp[kCtlThreads / 64 + 1];
  // Per tree: items this step, free slots and open-set size at entry.
  __shared__ uint32_t s_cnt[kCtlMaxTrees], s_free[kCtlMaxTrees], s_open[kCtlMaxTrees];
  // Small child of item i, for the block-wide fill of the histogram items.
  __shared__ uint32_t s_h0[kChildLdsItems], s_hnode[kChildLdsItems], s_hslot[kChildLdsItems],
      s_hbegin[kChildLdsItems], s_hcount[kChildLdsItems], s_hts[kChildLdsItems],
      s_htree[kChildLdsItems];
  const uint32_t n_items = B.counts[C_ITEMS];
  const bool lds_items = n_items <= kChildLdsItems;
  // One pass covers every item (the usual case): each thread loads its
  // item's state once, first.
  const bool one = n_items <= (uint32_t)kCtlThreads;
  ChildIn pre{};
  if (one) pre = child_load(B, threadIdx.x, threadIdx.x < n_items);
  if (threadIdx.x < P.T) {
    s_cnt[threadIdx.x] = 0;
    s_free[threadIdx.x] = B.ctl[threadIdx.x].n_free;
    s_open[threadIdx.x] = B.ctl[threadIdx.x].n_open;
  }
  // Every item of a step is at the same depth (depth-wise levels advance
  // together; leaf-wise always splits), so whether the children are split
  // next is one decision for the step.
  const bool need = n_items > 0 && (P.growth == 1 || B.pend[4] < P.max_depth);
  // Rows this step histograms (smaller children that will be split), which
  // sets the step's tile size.
  unsigned long long step_rows = 0ull;
  if (need) {
    if (one) {
      if (threadIdx.x < n_items) step_rows = min(pre.L, pre.R);
    } else {
      for (uint32_t i = threadIdx.x; i < n_items; i += blockDim.x) {
        const uint32_t L = B.cursor[2 * i], R = B.items[i].count - L;
        step_rows += L <= R ? L : R;
      }
    }
  }
  const uint32_t tile = step_tile(P, block_sum_u64(step_rows, s_red64));
  // Items per tree: the plan numbers a tree's items e = 0, 1, ... in order.
  if (one) {
    if (threadIdx.x < n_items) atomicMax(&s_cnt[pre.t], pre.e + 1u);
  } else {
    for (uint32_t i = threadIdx.x; i < n_items; i += blockDim.x)
      atomicMax(&s_cnt[B.pend[5 * i]], B.pend[5 * i + 3] + 1u);
  }
  __syncthreads();
  // Work lists are laid out by item order: histogram items by an exclusive
  // scan of each child's tile count, finish entries by a scan over the items
  // that need one, split entries 2i, 2i + 1 for item i; the e-th item of tree
  // t takes that tree's e-th free slot from the top of its free list.
  uint32_t carry_h = 0, carry_f = 0;
  unsigned long long hrows = 0ull;
  for (uint32_t base = 0; base < n_items; base += blockDim.x) {
    const uint32_t i = base + threadIdx.x;
    const bool in = i < n_items;
    const ChildIn c = one ? pre : child_load(B, i, in);
    const uint32_t t = c.t, lt = c.lt, rt = c.rt, e = c.e, cdepth = c.cdepth, L = c.L, R = c.R;
    const PartItem& it = c.it;
    const uint32_t scount = L <= R ? L : R;
    const uint32_t nt = in && need ? hist_node_tiles(scount, tile, P.hist_tile) : 0u;
    const uint32_t ts = hist_node_tile_rows(scount, nt);
    // The small child's histogram slot, loaded ahead of the scans.
    const uint32_t ss =
        in && need ? B.free_slot[(uint64_t)t * P.SPT + (s_free[t] - 1u - e)] : kNone;
    uint32_t tot;
    const uint32_t h0 = block_excl_scan(nt * P.NG, s_tmp, &tot) + carry_h;
    carry_h += tot;
    // Finish entries only for small children whose histogram is not already
    // whole in its slot: several tiles (partials to add) or none (zeros).
    const uint32_t fneed = (in && need && nt != 1) ? 1u : 0u;
    const uint32_t fi = block_excl_scan(fneed, s_tmp, &tot) + carry_f;
    carry_f += tot;
    if (in) {
      const uint32_t parent = it.node, right = it.right_node;
      // Children whose values the partition applied (the rows did not move)
      // take no rows, so the end of the tree lists no leaf range for them.
      const bool applied = (it.split_flags & kItemLeaves) != 0u;
      B.node_count[parent] = applied ? 0u : L;
      B.node_count[right] = applied ? 0u : R;
      // Children's exact totals: the split's left sums, and the rest.
      B.node_sum[2 * parent] = c.lg;
      B.node_sum[2 * parent + 1] = c.lh;
      B.node_sum[2 * right] = c.pg - c.lg;
      B.node_sum[2 * right + 1] = c.ph - c.lh;
      // The partition wrote the other buffer, the right child's rows from the
      // end of the parent's range down.
      const uint8_t cbuf = (uint8_t)((it.pad & 1u) ^ 1u);
      B.pbuf[parent] = cbuf;
      B.pbuf[right] = (uint8_t)(cbuf | 2u);
      const uint32_t pb = c.pb;
      B.pbegin[right] = pb + L;
      const bool left_small = L <= R;
      const uint32_t small = left_small ? parent : right, large = left_small ? right : parent;
      const uint32_t sbegin = left_small ? pb : pb + L;
      const uint32_t ps = c.ps;
      B.pslot[large] = ps;
      if (need) {
        B.pslot[small] = ss;
        const uint32_t gss = t * P.SPT + ss;
        hrows += scount;
        const uint32_t hflags = t | (nt == 1 ? kDirectFlag : 0u) | (cbuf ? kAltFlag : 0u);
        if (lds_items) {  // written by the whole block below
          s_h0[i] = h0;
          s_hnode[i] = small;
          s_hslot[i] = gss;
          s_hbegin[i] = sbegin;
          s_hcount[i] = scount;
          s_hts[i] = ts;
          s_htree[i] = hflags;
        } else {
          for (uint32_t q = 0; q < nt; q++)
            for (uint32_t g = 0; g < P.NG; g++) {
              HistItem h;
              h.node = small;
              h.slot = gss;
              h.begin = sbegin + q * ts;
              h.count = min(ts, scount - q * ts);
              h.group = g;
              h.tree = hflags;
              B.hitems[h0 + q * P.NG + g] = h;
            }
        }
        // The finish only adds the small child's partials; the split of the
        // large child subtracts the small one from the parent's histogram,
        // which the large child keeps in the parent's slot (ps), and writes
        // the result back there.
        if (fneed) {
          B.fin[4 * fi + 0] = gss;
          B.fin[4 * fi + 1] = kNone;
          B.fin[4 * fi + 2] = h0;
          B.fin[4 * fi + 3] = nt;
        }
        // Depth-wise, the children's children at the depth limit get no
        // histograms, so the large child's histogram is used only by its split.
        const uint32_t msub =
            gss | (P.growth == 0 && cdepth + 1u >= P.max_depth ? kMinusNoWriteback : 0u);
        B.split_nodes[2 * i] = parent;
        B.split_slots[2 * i] = t * P.SPT + (left_small ? ss : ps);
        B.split_minus[2 * i] = left_small ? kNone : msub;
        B.split_nodes[2 * i + 1] = right;
        B.split_slots[2 * i + 1] = t * P.SPT + (left_small ? ps : ss);
        B.split_minus[2 * i + 1] = left_small ? msub : kNone;
      } else {
        B.pslot[small] = kNone;
      }
      const uint32_t lp = parent - t * P.MN, rp = right - t * P.MN;
      if (P.growth == 0) {
        const uint64_t nb = (uint64_t)t * 2 * P.W + 2 * e;
        B.next_node[nb] = lp;
        B.next_tnode[nb] = lt;
        B.next_depth[nb] = cdepth;
        B.next_node[nb + 1] = rp;
        B.next_tnode[nb + 1] = rt;
        B.next_depth[nb + 1] = cdepth;
      } else {
        const uint64_t ob = (uint64_t)t * P.WO + s_open[t] + 2 * e;
        B.open_node[ob] = lp;
        B.open_tnode[ob] = lt;
        B.open_depth[ob] = cdepth;
        B.open_node[ob + 1] = rp;
        B.open_tnode[ob + 1] = rt;
        B.open_depth[ob + 1] = cdepth;
      }
    }
  }
  const unsigned long long rows_total = block_sum_u64(hrows, s_red64);  // also syncs s_h*
  if (need && lds_items) {
    // Histogram item k belongs to the last item i with s_h0[i] <= k (every
    // item has at least one tile, so s_h0 is strictly increasing).
    for (uint32_t k = threadIdx.x; k < carry_h; k += blockDim.x) {
      uint32_t lo = 0, hi = n_items - 1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (s_h0[mid] <= k) lo = mid; else hi = mid - 1;
      }
      const uint32_t r = k - s_h0[lo], q = r / P.NG, g = r % P.NG;
      const uint32_t ts = s_hts[lo];
      HistItem h;
      h.node = s_hnode[lo];
      h.slot = s_hslot[lo];
      h.begin = s_hbegin[lo] + q * ts;
      h.count = min(ts, s_hcount[lo] - q * ts);
      h.group = g;
      h.tree = s_htree[lo];
      B.hitems[k] = h;
    }
  }
  if (threadIdx.x == 0) {
    B.counts[C_HIST] = carry_h;
    B.counts[C_FIN] = carry_f;
    B.counts[C_SPLIT] = need ? 2u * n_items : 0u;
    B.counts[C_HROWS] += (uint32_t)rows_total;
    B.counts[C_HNODES] += need ? n_items : 0u;
  }
  if (threadIdx.x < P.T) {
    TreeCtl& c = B.ctl[threadIdx.x];
    const uint32_t cnt = s_cnt[threadIdx.x];
    if (need) c.n_free = s_free[threadIdx.x] - cnt;
    if (P.growth == 0)
      c.n_next += 2u * cnt;
    else
      c.n_open = s_open[threadIdx.x] + 2u * cnt;
  }
}

// Remaining candidates become leaves; learning rate; leaf-range update list.
__global__ __launch_bounds__(kCtlThreads) void k_ctl_finalize(CtlParams P, DeviceBufs B,
                                                              int advance,
                                                              uint32_t* __restrict__ round) {
  __shared__ uint32_t s_tmp[kCtlThreads / 64 + 1];
  // advance 1, depth-wise: the last step's children (the next level) are
  // what remains; they are read from the next-level list in place.
  // advance 2: the last step's partition applied its children's values and
  // its children step did not run: each item's two children become leaves
  // here, from the split's left sums and the parent's totals (what that step
  // would have given them), with no rows for the leaf-range list.
  const bool from_items = advance == 2;
  const bool lvl = (advance == 1 && P.growth == 0) || from_items;
  if (from_items) {
    const uint32_t ni = B.counts[C_ITEMS];
    for (uint32_t i = threadIdx.x; i < ni; i += blockDim.x) {
      const uint32_t t = B.pend[5 * i], lt = B.pend[5 * i + 1], rt = B.pend[5 * i + 2];
      const PartItem it = B.items[i];
      const long long lg = B.node_split[it.node].lg, lh = B.node_split[it.node].lh;
      const long long pg = B.node_sum[2 * it.node], ph = B.node_sum[2 * it.node + 1];
      const int out = P.out0 + (int)t;
      const int qg = B.qexp[2 * out], qh = B.qexp[2 * out + 1];
      make_leaf_from(P, B, t, it.node, lt, lg, lh, qg, qh, kNone);
      make_leaf_from(P, B, t, it.right_node, rt, pg - lg, ph - lh, qg, qh, kNone);
      B.node_count[it.node] = 0;
      B.node_count[it.right_node] = 0;
    }
    __syncthreads();
  }
  uint32_t carry_s = 0;
  for (uint32_t t = 0; t < P.T; t++) {
    const TreeCtl c = B.ctl[t];
    const uint32_t no = from_items ? 0u : lvl ? c.n_next : c.n_open;
    const uint64_t lb = lvl ? (uint64_t)t * 2 * P.W : (uint64_t)t * P.WO;
    const uint32_t* onode = (lvl ? B.next_node : B.open_node) + lb;
    const uint32_t* otn = (lvl ? B.next_tnode : B.open_tnode) + lb;
    const int out = P.out0 + (int)t;
    const int qg = B.qexp[2 * out], qh = B.qexp[2 * out + 1];
    for (uint32_t i = threadIdx.x; i < no; i += blockDim.x) {
      const uint32_t gid = t * P.MN + onode[i];
      // The tree is done: the next round's start resets every slot, so no
      // slot is returned to the free list.
      make_leaf_from(P, B, t, gid, otn[i], B.node_sum[2 * gid], B.node_sum[2 * gid + 1], qg, qh,
                     kNone);
    }
    __syncthreads();  // leaf values of the leaves made above are visible
    if (threadIdx.x == 0) {
      TreeCtl& cw = B.ctl[t];
      cw.n_open = 0;
      if (lvl) {
        cw.n_next = 0;
        cw.depth = c.depth + 1;
      }
    }
    // Leaf-range update list: every leaf's rows in seg_tile segments, laid
    // out by a scan over the partition ids.
    const uint32_t np = c.n_part;
    for (uint32_t base = 0; base < np; base += kCtlThreads) {
      const uint32_t p = base + threadIdx.x;
      const bool in = p < np;
      const uint32_t gid = t * P.MN + (in ? p : 0u);
      const float lv = B.leaf_val[gid];
      const uint32_t cnt = B.node_count[gid], b = B.pbegin[gid];
      const uint32_t alt = (B.pbuf[gid] & 1u) ? 0x80000000u : 0u;
      const uint32_t ns = in && !isnan(lv) ? (cnt + P.seg_tile - 1) / P.seg_tile : 0u;
      uint32_t tot;
      const uint32_t s0 = block_excl_scan(ns, s_tmp, &tot) + carry_s;
      carry_s += tot;
      for (uint32_t q = 0; q < ns; q++) {
        Segment sg;
        sg.node = gid;
        sg.begin = b + q * P.seg_tile;
        sg.count = min(P.seg_tile, cnt - q * P.seg_tile);
        sg.tree = t | alt;
        B.leaf_segs[s0 + q] = sg;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    B.counts[C_LSEGS] = carry_s;
    if (round) *round += 1u;  // nothing after this kernel in the round reads it
  }
}

__global__ void k_copy_u32(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src,
                           int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

size_t tree_out_layout(uint32_t T, uint32_t MNT, char* base, TreeOut* out) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* p = base ? base + off : nullptr;
    off += (bytes + 255) / 256 * 256;
    return p;
  };
  const size_t tn = (size_t)T * MNT;
  TreeOut o;
  o.ctl = (TreeCtl*)take(sizeof(TreeCtl) * T);
  o.counts = (uint32_t*)take(sizeof(uint32_t) * C_NCOUNTS);
  o.tr_split = (uint32_t*)take(4 * tn);
  o.tr_left = (uint32_t*)take(4 * tn);
  o.tr_right = (uint32_t*)take(4 * tn);
  o.tr_cats = (uint32_t*)take(4 * tn * kMaxCatWords);
  o.tr_thr = (float*)take(4 * tn);
  o.tr_leaf = (float*)take(4 * tn);
  o.tr_gain = (float*)take(4 * tn);
  o.tr_cover = (float*)take(4 * tn);
  o.tr_dleft = (uint8_t*)take(tn);
  o.tr_isleaf = (uint8_t*)take(tn);
  o.tr_stype = (uint8_t*)take(tn);
  if (out) *out = o;
  return off;
}

void launch_ctl_init(hipStream_t s, const CtlParams& p, const DeviceBufs& b,
                     const uint32_t* rows_per_tree, uint32_t* absmax, int n_q, int64_t n_rows,
                     int32_t* qexp) {
  if (n_q > kCtlThreads) throw std::runtime_error("too many quantisation scales for one block");
  hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(kCtlThreads), 0, s, p, b, rows_per_tree, absmax,
                     n_q, n_rows, qexp);
}
void launch_ctl_plan(hipStream_t s, const CtlParams& p, const DeviceBufs& b, bool advance) {
  hipLaunchKernelGGL(k_ctl_plan, dim3(1), dim3(kCtlThreads), 0, s, p, b, advance ? 1 : 0);
}
void launch_ctl_children(hipStream_t s, const CtlParams& p, const DeviceBufs& b) {
  hipLaunchKernelGGL(k_ctl_children, dim3(1), dim3(kCtlThreads), 0, s, p, b);
}
void launch_ctl_finalize(hipStream_t s, const CtlParams& p, const DeviceBufs& b, int advance,
                         uint32_t* round) {
  hipLaunchKernelGGL(k_ctl_finalize, dim3(1), dim3(kCtlThreads), 0, s, p, b, advance, round);
}
void launch_copy_u32(hipStream_t s, uint32_t* dst, const uint32_t* src, int64_t n) {
  if (n <= 0) return;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_copy_u32, dim3((unsigned)blocks), dim3(256), 0, s, dst, src, n);
}
void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, int64_t n) {
  if (n <= 0) return;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
}

}  // namespace bst
