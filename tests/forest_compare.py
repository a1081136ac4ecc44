"""Forest comparison that tolerates split-gain near-ties (test helper).

Two trainers that sum the same gradients in a different order (the engine's
exact 64-bit fixed-point histograms vs the oracle's f64 sums, which the
reference also uses) can disagree where two candidate splits have gains equal
to about 1e-7: the f32-rounded gains then order differently.  Everything up
to such a tie must be identical, and at the first divergence the competing
gains must be within the stated relative tolerance; what follows the tie is
no longer comparable node by node (the accuracy gate checks it instead).
"""
from __future__ import annotations

import numpy as np

NODE_FIELDS = ("split_index", "split_threshold", "default_left", "is_leaf", "split_type")


def _tree(a, t):
    lo, hi = int(a["tree_offset"][t]), int(a["tree_offset"][t + 1])
    return lo, hi


def _cats(a, i):
    if not a["split_type"][i] or a["cat_size"][i] == 0:
        return ()
    s, n = int(a["cat_start"][i]), int(a["cat_size"][i])
    w = list(a["cat_words"][s:s + n])
    while w and w[-1] == 0:
        w.pop()
    return tuple(int(x) for x in w)


def _split_key(a, i):
    return (int(a["split_index"][i]), a["split_threshold"][i].view(np.uint32).item(),
            int(a["default_left"][i]), int(a["is_leaf"][i]), int(a["split_type"][i]), _cats(a, i))


def _parent_of(a, lo, hi, child):
    for i in range(lo, hi):
        if not a["is_leaf"][i] and (a["left"][i] == child or a["right"][i] == child):
            return i
    return None


def _close(x, y, rel):
    return abs(float(x) - float(y)) <= rel * max(abs(float(x)), abs(float(y)), 1e-30)


def compare(ga, oa, min_gain=0.0, gain_rel=1e-6, leaf_rtol=1e-6):
    """Compare forest arrays `ga` (engine) and `oa` (oracle) tree by tree.

    Returns a dict: `trees_equal` (trees identical before any divergence),
    `divergence` (None, or where and why the first tie occurred).  Raises
    AssertionError on a difference that is not a near-tie."""
    tg = len(ga["tree_offset"]) - 1
    to = len(oa["tree_offset"]) - 1
    np.testing.assert_array_equal(ga["base_score"], oa["base_score"])
    for t in range(min(tg, to)):
        assert ga["tree_group"][t] == oa["tree_group"][t], f"tree {t} group"
        g0, g1 = _tree(ga, t)
        o0, o1 = _tree(oa, t)
        m = min(g1 - g0, o1 - o0)
        first = None
        for i in range(m):
            gi, oi = g0 + i, o0 + i
            if _split_key(ga, gi) != _split_key(oa, oi):
                first = ("split", i)
                break
            if not ga["is_leaf"][gi] and (ga["left"][gi] != oa["left"][oi] or
                                          ga["right"][gi] != oa["right"][oi]):
                first = ("children", i)
                break
        if first is None and g1 - g0 == o1 - o0:
            lg = ga["leaf_value"][g0:g1]
            lo = oa["leaf_value"][o0:o1]
            np.testing.assert_allclose(lg, lo, rtol=leaf_rtol, atol=1e-12, err_msg=f"tree {t} leaves")
            inner = ~ga["is_leaf"][g0:g1].astype(bool)
            np.testing.assert_allclose(ga["gain"][g0:g1][inner], oa["gain"][o0:o1][inner],
                                       rtol=1e-5, err_msg=f"tree {t} gains")
            continue
        if first is None:  # same prefix, one tree longer: a leaf-count tie at the end
            first = ("children", m - 1)
        kind, i = first
        gi, oi = g0 + i, o0 + i
        if kind == "split":
            gl, ol = bool(ga["is_leaf"][gi]), bool(oa["is_leaf"][oi])
            if not gl and not ol:
                gg, og = ga["gain"][gi], oa["gain"][oi]
                assert _close(gg, og, gain_rel), (
                    f"tree {t} node {i}: different splits with gains {gg} vs {og} "
                    f"(engine {_split_key(ga, gi)}, oracle {_split_key(oa, oi)})")
                why = f"split tie, gains {gg:.9g} vs {og:.9g}"
            else:
                g = oa["gain"][oi] if gl else ga["gain"][gi]
                assert abs(float(g) - min_gain) <= gain_rel * max(1.0, abs(float(g))), (
                    f"tree {t} node {i}: leaf vs split with gain {g}")
                why = f"expand tie at min_gain, gain {g:.9g}"
        else:
            # Same split, children allocated in another order: two open nodes
            # popped in the other order (leaf-wise max-gain ties).  The node
            # ids where the trees part were allocated by different parents;
            # their gains must tie.
            c = int(min(ga["left"][gi], oa["left"][oi]))
            pg = _parent_of(ga, g0, g1, c)
            po = _parent_of(oa, o0, o1, c)
            assert pg is not None and po is not None, f"tree {t} node {i}: children differ"
            gg, og = ga["gain"][pg], oa["gain"][po]
            assert _close(gg, og, gain_rel), (
                f"tree {t}: expansion order differs at node id {c}, parent gains {gg} vs {og}")
            why = f"pop-order tie, gains {gg:.9g} vs {og:.9g}"
        return {"trees_equal": t, "divergence": {"tree": t, "node": i, "why": why}}
    assert tg == to, f"{tg} vs {to} trees"
    return {"trees_equal": tg, "divergence": None}
