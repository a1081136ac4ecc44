"""Pins for the oracle's Exclusive Feature Bundling (data/binned/bundling.rs,
dataset.rs:229-236,460-568,898-952) and the generator its row sample draws
from.

- xoshiro256++ against the reference implementation's published output, and
  rand_xoshiro's seed_from_u64 as four SplitMix64 words (SplitMix64 is pinned
  in test_oracle_sampling_metrics.py).
- rand 0.8.5's inclusive range sampling: the high word of the 128-bit
  product (closed form for power-of-two ranges) and the conservative
  rejection zone, on scripted draws where the exact zone (or no rejection)
  gives another value.
- The row sample (first and last tenth, a reservoir sample in between) on the
  reference's own test cases plus the tenths.
- The plan on the reference's own cases (too few sparse features, two
  exclusive ones) and on cases where a plausible mistake changes the bundles:
  the conflict budget ceil(1e-4 x sampled rows), the density order, the first
  bundle of smallest conflict, the 0.5 conflict-rate cut, numeric and dense
  features left out, the sparsity boundary in f32.
- The codes (offset + bin, offsets from 1, the first set feature of a row
  wins) against a numpy restatement of build_bundle, and the children of a
  split on a bundle: code 0 goes right whatever the category set (a
  categorical split's default_left is false), so each leaf is the closed form
  of the rows that reach it.

The exact reservoir stream for more than 10000 rows rests on the restated
rand 0.8.5 / rand_xoshiro 0.6 algorithms (crates the reference pins but does
not vendor); no reference vector covers it.
"""
import numpy as np
import pytest

from boosters_amd import _abi as A
from tests.conftest import make_cfg, make_ext, sparse_cat_data

M64 = 1 << 64


def _binning(enable=1, sparsity=0.9):
    return A.BinningConfig(256, sparsity, enable, 0, 200000)


# ------------------------------------------------------------ generators ---
def test_xoshiro256pp_published_vector(oracle):
    # xoshiro256plusplus.c (Blackman and Vigna) from the state {1, 2, 3, 4}:
    # the first ten outputs, the known-answer test rand_xoshiro ports.
    out = oracle.xoshiro256pp(10, state=[1, 2, 3, 4])
    assert [int(v) for v in out] == [
        41943041, 58720359, 3588806011781223, 3591011842654386, 9228616714210784205,
        9973669472204895162, 14011001112246962877, 12406186145184390807,
        15849039046786891736, 10450023813501588000]


def test_xoshiro_seed_from_u64_is_splitmix64_words(oracle):
    for seed in (0, 42, 2**63 + 5):
        words = oracle.splitmix64_stream(seed, 4)
        np.testing.assert_array_equal(oracle.xoshiro256pp(6, seed=seed),
                                      oracle.xoshiro256pp(6, state=words))


def test_sample_inclusive_power_of_two_takes_the_high_bits(oracle):
    # range 2^k: the conservative zone is (2^k << (63 - k)) - 1 = 2^63 - 1, so
    # a draw is kept when the low word d << k is below 2^63 (bit 63 - k of d
    # clear), and the value is the top k bits of the draw.  A low-word or
    # modulo rule gives the low bits; the exact zone (2^64 - 1 for a power of
    # two) would keep the draw with that bit set.
    rng = np.random.default_rng(0)
    for k in (1, 3, 17, 40):
        bit = 1 << (63 - k)
        for d in rng.integers(0, 2**63, 5, dtype=np.int64):
            keep = (int(d) * 2 + 1) & ~bit
            v, used = oracle.sample_inclusive([keep], 0, (1 << k) - 1)
            assert used == 1 and v == keep >> (64 - k)
            v, used = oracle.sample_inclusive([keep | bit, keep], 0, (1 << k) - 1)
            assert used == 2 and v == keep >> (64 - k)


def test_sample_inclusive_conservative_zone(oracle):
    # range 3: conservative zone (3 << 62) - 1 = 0xBFFF...F.  A draw whose low
    # word is 0xC000...0 is rejected (the exact zone 2^64 - 3 would accept it
    # and return 0); the next draw decides.  A low word equal to the zone is
    # accepted (the comparison is <=).
    inv3 = 0xAAAAAAAAAAAAAAAB
    rejected = (0xC000000000000000 * inv3) % M64
    at_zone = (((3 << 62) - 1) * inv3) % M64
    assert (rejected * 3) % M64 == 0xC000000000000000
    v, used = oracle.sample_inclusive([rejected, at_zone], 0, 2)
    assert used == 2 and v == (at_zone * 3) >> 64 == 1
    v, used = oracle.sample_inclusive([at_zone], 5, 7)
    assert used == 1 and v == 6  # low + the high word


# ------------------------------------------------------------ row sample ---
def test_efb_sample_rows_reference_cases(oracle):
    # bundling.rs test_sample_rows_small / test_sample_rows_large.
    np.testing.assert_array_equal(oracle.efb_sample_rows(50, 100), np.arange(50))
    rows = oracle.efb_sample_rows(100000, 1000)
    assert 900 <= len(rows) <= 1000 and np.all(np.diff(rows) > 0)
    # The first and last tenth of the sample are the first and last rows;
    # the rest lie strictly between them.
    assert set(range(100)) <= set(rows.tolist())
    assert set(range(99900, 100000)) <= set(rows.tolist())
    mid = rows[(rows >= 100) & (rows < 99900)]
    assert len(mid) == 800  # the reservoir keeps n_random distinct rows


def test_efb_sample_rows_reservoir_is_uniform(oracle):
    # A reservoir sample gives every middle row the same chance: over seeds,
    # each of the 10 middle rows of n = 30, max = 20 (16 kept) is in about
    # 16/26 of the samples... with max = 20: first 2, last 2, 16 of the 26
    # middle rows.
    counts = np.zeros(30)
    S = 400
    for seed in range(S):
        for r in oracle.efb_sample_rows(30, 20, seed):
            counts[r] += 1
    assert np.all(counts[:2] == S) and np.all(counts[28:] == S)
    p = 16 / 26
    sd = np.sqrt(S * p * (1 - p))
    assert np.all(np.abs(counts[2:28] - S * p) < 5 * sd)


# ------------------------------------------------------------------ plan ---
def _plan(oracle, X, is_cat, **kw):
    return oracle.Binned(X, is_cat, _binning(**kw)).bundles()


def _cat(n, rows, val=1.0):
    c = np.zeros(n, np.float32)
    c[list(rows)] = val
    return c


def test_plan_reference_cases(oracle):
    n = 100
    # test_create_bundle_plan_too_few_sparse: one sparse categorical.
    assert _plan(oracle, _cat(n, range(5))[None, :], np.ones(1, np.uint8)) == []
    # test_create_bundle_plan_no_conflicts: rows 0-4 and 5-9, one bundle.
    X = np.stack([_cat(n, range(5)), _cat(n, range(5, 10))])
    assert _plan(oracle, X, np.ones(2, np.uint8)) == [[0, 1]]


def test_plan_conflict_budget(oracle):
    # 100 sampled rows: budget ceil(1e-4 * 100) = 1 conflicting row.  A third
    # feature keeps the share of conflicting pairs at 1/3 (<= 0.5).  Order by
    # density (stable): 0, 2, 1.
    n = 100
    X = np.stack([_cat(n, [0, 1, 2, 3]), _cat(n, [3, 4, 5]), _cat(n, [50, 51, 52, 53])])
    assert _plan(oracle, X, np.ones(3, np.uint8)) == [[0, 2, 1]]
    X = np.stack([_cat(n, [0, 1, 2, 3]), _cat(n, [2, 3, 4]), _cat(n, [50, 51, 52, 53])])
    assert _plan(oracle, X, np.ones(3, np.uint8)) == [[0, 2]]
    # Two features that conflict at all: every pair conflicts (rate 1 > 0.5).
    X = np.stack([_cat(n, [0, 1, 2, 3]), _cat(n, [3, 4, 5])])
    assert _plan(oracle, X, np.ones(2, np.uint8)) == []


def test_plan_density_order_and_first_smallest_conflict(oracle):
    n = 200
    a = _cat(n, range(0, 9))       # density 0.045, conflicts with b on 2 rows
    b = _cat(n, range(7, 17))      # density 0.05 (denser: placed first)
    c = _cat(n, range(40, 45))     # no conflicts
    d = _cat(n, [7, 60, 61])       # conflicts with a and b on row 7
    X = np.stack([a, b, c, d])
    # 3 of 6 pairs conflict (0.5, not above the cut).  Order b, a, c, d: a
    # opens a bundle (2 conflicts with b > budget 1); c joins b's bundle
    # (increase 0 there and in a's: the earlier wins); d adds 1 in either,
    # both within budget: the earlier wins again; a is left alone.
    assert _plan(oracle, X, np.ones(4, np.uint8)) == [[1, 2, 3]]


def test_plan_leaves_out_numeric_dense_and_trivial(oracle):
    n = 100
    sparse_num = _cat(n, range(0, 3), 2.5)
    dense_cat = np.tile(np.array([0, 1, 2, 3], np.float32), 25)
    zero = np.zeros(n, np.float32)
    X = np.stack([_cat(n, range(10, 13)), sparse_num, dense_cat, zero, _cat(n, range(20, 25))])
    is_cat = np.array([1, 0, 1, 1, 1], np.uint8)
    assert _plan(oracle, X, is_cat) == [[4, 0]]  # denser first
    assert _plan(oracle, X, is_cat, enable=0) == []


def test_plan_high_conflict_rate_skips(oracle):
    # Three features pairwise conflicting on one row each: 3 of 3 pairs
    # conflict (rate 1 > 0.5), no plan even though each pair fits the budget.
    n = 100
    X = np.stack([_cat(n, [0, 1, 5]), _cat(n, [1, 2, 6]), _cat(n, [2, 0, 7])])
    assert _plan(oracle, X, np.ones(3, np.uint8)) == []


def test_plan_sparsity_boundary_is_f32(oracle):
    # density = 10 / 100 = 0.1 <= 1 - 0.9 (= 0.100000024 in f32): sparse.
    n = 100
    X = np.stack([_cat(n, range(10)), _cat(n, range(10, 20))])
    assert _plan(oracle, X, np.ones(2, np.uint8)) == [[0, 1]]
    X = np.stack([_cat(n, range(11)), _cat(n, range(11, 22))])
    assert _plan(oracle, X, np.ones(2, np.uint8)) == []


# ----------------------------------------------------------------- codes ---
def test_codes_and_decoded_bins(oracle):
    X, _, is_cat = sparse_cat_data(3000, 2, 4, card=5, conflicts=1, seed=3)
    b = oracle.Binned(X, is_cat, _binning())
    bundles = b.bundles()
    # One bundle of the four (one conflicting row, within the budget of
    # ceil(1e-4 x 3000) = 1), denser features first (a stable order).
    dens = {f: np.float32(np.count_nonzero(X[f])) / np.float32(X.shape[1]) for f in range(2, 6)}
    assert len(bundles) == 1
    assert bundles[0] == sorted(range(2, 6), key=lambda f: -dens[f])
    lay = b.layout()
    feats = bundles[0]
    nb = [b.feature_info(f)["n_bins"] for f in feats]
    assert lay["eff_to_orig"][0] == feats[0] and lay["n_bins"][0] == 1 + sum(nb)
    # numpy build_bundle: code = offset + bin (bin = category here), the
    # first feature in bundle order set on the row wins, offsets from 1.
    offs = np.cumsum([1] + nb[:-1])
    code = np.zeros(X.shape[1], np.int64)
    for f, o in zip(feats, offs):
        v = X[f]
        take = (code == 0) & (v != 0)
        code[take] = o + v[take].astype(np.int64)
    bins = b.bins()
    for f, o, k in zip(feats, offs, nb):
        own = (code >= o) & (code < o + k)
        np.testing.assert_array_equal(bins[f][own], code[own] - o)
        assert np.all(bins[f][~own] == 0)  # the bin of 0.0 (category 0)
    # The conflict row: set in both of the first two sparse features; the
    # bundle's first one keeps it, the other decodes to its default.
    both = (X[2] != 0) & (X[3] != 0)
    assert both.sum() == 1
    first = min((2, 3), key=feats.index)
    second = 5 - first
    assert bins[first][both][0] == X[first][both][0] and bins[second][both][0] == 0


def test_bundle_split_children_are_their_rows(oracle):
    # A stump on data where only the bundle matters: the tree stores a
    # categorical split on the bundle's first feature, its categories are
    # codes, and each leaf is -lr G / (H + lambda) of the rows that reach it
    # when code 0 goes right whatever the set (default_left false).
    X, y, is_cat = sparse_cat_data(4000, 0, 3, card=4, seed=5)
    b = oracle.Binned(X, is_cat, _binning())
    feats = b.bundles()[0]
    cfg = make_cfg(objective=A.OBJ_SQUARED, n_trees=1, max_depth=1, learning_rate=1.0)
    oracle.set_exp_mode(1)
    m = oracle.train(b, y, cfg, make_ext(0))
    arr = m.arrays
    assert arr["split_type"][0] == 1 and arr["split_index"][0] == feats[0]
    nb = [b.feature_info(f)["n_bins"] for f in feats]
    offs = np.cumsum([1] + nb[:-1])
    code = np.zeros(X.shape[1], np.int64)
    for f, o in zip(feats, offs):
        v = X[f]
        take = (code == 0) & (v != 0)
        code[take] = o + v[take].astype(np.int64)
    words = arr["cat_words"][arr["cat_start"][0]:arr["cat_start"][0] + arr["cat_size"][0]]
    in_set = np.array([(c >> 5) < len(words) and (int(words[c >> 5]) >> (c & 31)) & 1
                       for c in code], bool)
    train_left = in_set & (code != 0)
    # Squared loss: g = base - y in f32 (h = 1), summed in f64 by the histogram.
    g = (arr["base_score"][0] - y).astype(np.float64)
    lam = 1.0
    # Stored children: categorical splits swap them (training-left rows take
    # the inference-right child).
    left_leaf, right_leaf = arr["right"][0], arr["left"][0]
    for rows, node in ((train_left, left_leaf), (~train_left, right_leaf)):
        G = g[rows].sum()
        H = float(rows.sum())
        want = np.float32(-G / (H + lam)) if H > 0 else np.float32(0.0)
        assert abs(arr["leaf_value"][node] - want) <= 1e-6 * max(1.0, abs(float(want)))
