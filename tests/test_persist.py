""".bstr / .bstr.json persistence through the C ABI (CPU only: no device calls).

The reference pins persistence by round trips and by corruption, truncation and
magic rejection (crates/boosters/tests/persist_proptest.rs:370-454); it holds
no byte-level fixture, and rmp-serde's exact bytes cannot be produced here.
These tests pin the format from its definition instead:

* CRC32C by its published check values (the Castagnoli "123456789" check and
  the iSCSI vectors of RFC 3720 B.4), computed by a bitwise implementation in
  this file, which then checks every trailer;
* the payload against an independent encoding: the schema of
  persist/schema.rs:27-149 is written out here as Python dicts (fields in
  declaration order, optional ones left out) and packed with the `msgpack`
  package, whose bytes must equal the library's payload;
* reading of files the library did not write: payloads packed here with
  fields reordered, unknown fields, nil options, vector leaves, integer and
  f32 encodings of floats, multi-frame zstd without content size;
* the error cases of the reference's tests and of Tree::validate.
"""
import ctypes as C
import json
import pickle
import struct

import msgpack
import numpy as np
import pytest

from boosters_amd import GBDTModel
from boosters_amd import _abi as A

# ----------------------------------------------------------------- CRC32C


def crc32c_bitwise(data: bytes) -> int:
    c = 0xFFFFFFFF
    for b in data:
        c ^= b
        for _ in range(8):
            c = (c >> 1) ^ (0x82F63B78 if c & 1 else 0)
    return c ^ 0xFFFFFFFF


def test_crc32c_reference_vectors():
    assert crc32c_bitwise(b"123456789") == 0xE3069283
    assert crc32c_bitwise(bytes(32)) == 0x8A9136AA
    assert crc32c_bitwise(b"\xff" * 32) == 0x62A8AB43
    assert crc32c_bitwise(bytes(range(32))) == 0x46DD794E
    assert crc32c_bitwise(bytes(range(31, -1, -1))) == 0x113FDB5C


# ----------------------------------------------------------------- builders


def tree(nodes, group=0):
    """nodes: list of dicts, node id = position.  Split nodes carry f, thr,
    l, r, dl (and cats for a categorical split); leaves carry v."""
    return dict(nodes=nodes, group=group)


def import_model(trees, n_groups=1, n_features=4, transform=0, base=None):
    """A bst_model through bst_model_import with a NULL context."""
    N = sum(len(t["nodes"]) for t in trees)
    words = []
    arr = {name: [] for name, _, _ in A.FOREST_FIELDS}
    offs = [0]
    for t in trees:
        for nd in t["nodes"]:
            leaf = "v" in nd
            arr["split_index"].append(nd.get("f", 0))
            arr["split_threshold"].append(nd.get("thr", 0.0))
            arr["left"].append(0 if leaf else nd["l"])
            arr["right"].append(0 if leaf else nd["r"])
            arr["default_left"].append(int(nd.get("dl", False)))
            arr["is_leaf"].append(int(leaf))
            arr["leaf_value"].append(nd.get("v", 0.0))
            arr["gain"].append(nd.get("gain", 0.0))
            arr["cover"].append(nd.get("cover", 0.0))
            bits = nd.get("bits")
            arr["split_type"].append(int(bits is not None))
            arr["cat_start"].append(0 if bits is None else len(words))
            arr["cat_size"].append(0 if bits is None else len(bits))
            if bits is not None:
                words.extend(bits)
        offs.append(offs[-1] + len(t["nodes"]))
        arr["tree_group"].append(t["group"])
    arr["tree_offset"] = offs
    arr["base_score"] = base if base is not None else [0.0] * n_groups
    arr["cat_words"] = words or [0]
    arrays = {name: np.ascontiguousarray(arr[name], dtype=dt) for name, dt, _ in A.FOREST_FIELDS}
    info = A.ModelInfo(len(trees), n_groups, n_features, N, len(words), transform, 0)
    h = C.c_void_p()
    A.check(A.lib().bst_model_import(None, C.byref(info), C.byref(A.forest_struct(arrays)), C.byref(h)))
    return GBDTModel(h, None)


def export(m):
    a = m.export()
    a["cat_words"] = a["cat_words"][: m.info.total_cat_words]
    return a


def same_model(a, b):
    assert (a.info.n_trees, a.info.n_groups, a.info.n_features, a.info.total_nodes,
            a.info.total_cat_words, a.info.output_transform) == (
        b.info.n_trees, b.info.n_groups, b.info.n_features, b.info.total_nodes,
        b.info.total_cat_words, b.info.output_transform)
    ea, eb = export(a), export(b)
    for k in ea:
        np.testing.assert_array_equal(ea[k], eb[k], err_msg=k)


# The reference's own example tree (persist_proptest.rs:380-385):
#   0 => num(0, 0.5, L) -> 1, 2;  1 => leaf(1.0);  2 => leaf(2.0)
def simple_model():
    return import_model([tree([dict(f=0, thr=0.5, l=1, r=2, dl=True), dict(v=1.0), dict(v=2.0)])],
                        n_features=2)


def random_tree(rng, F, n_leaves, group, with_stats, with_cats):
    nodes = [dict(v=0.0)]
    leaves = [0]
    while len(leaves) < n_leaves:
        j = leaves.pop(int(rng.integers(len(leaves))))
        l, r = len(nodes), len(nodes) + 1
        nd = dict(f=int(rng.integers(F)), thr=float(np.float32(rng.standard_normal())), l=l, r=r,
                  dl=bool(rng.integers(2)))
        if with_cats and rng.random() < 0.3:
            nw = int(rng.integers(1, 4))
            bits = [int(x) for x in rng.integers(0, 2**32, nw, dtype=np.uint64)]
            bits[-1] |= 1 << int(rng.integers(32))  # minimal width: last word non-zero
            nd["bits"] = bits
        if with_stats:
            nd["gain"] = float(np.float32(rng.random() * 10))
            nd["cover"] = float(np.float32(rng.random() * 100))
        nodes[j] = nd
        for c in (l, r):
            leaf = dict(v=float(np.float32(rng.standard_normal() * 0.1)))
            if with_stats:
                leaf["cover"] = float(np.float32(rng.random() * 50))
            nodes.append(leaf)
            leaves.append(c)
    return tree(nodes, group)


def random_model(seed, K=1, n_trees=7, F=9, transform=0):
    rng = np.random.default_rng(seed)
    trees = [random_tree(rng, F, int(rng.integers(1, 40)), t % K, bool(t % 3), True)
             for t in range(n_trees)]
    base = [float(np.float32(x)) for x in rng.standard_normal(K)]
    return import_model(trees, n_groups=K, n_features=F, transform=transform, base=base), trees


# ----------------------------------------------------------------- schema restated

TRANSFORMS = ["identity", "sigmoid", "softmax"]


def schema_dict(trees, K, F, transform, base):
    """GBDTModelSchema (persist/schema.rs:27-149) of the trees, as the
    reference's writer fills it (convert.rs:77-167, 417-430, 710-721)."""
    ts = []
    for t in trees:
        nodes = t["nodes"]
        d = {
            "num_nodes": len(nodes),
            "split_indices": [nd.get("f", 0) for nd in nodes],
            "thresholds": [float(np.float32(nd.get("thr", 0.0))) for nd in nodes],
            "children_left": [0 if "v" in nd else nd["l"] for nd in nodes],
            "children_right": [0 if "v" in nd else nd["r"] for nd in nodes],
            "default_left": [bool(nd.get("dl", False)) for nd in nodes],
            "leaf_values": {"type": "scalar",
                            "values": [float(np.float32(nd.get("v", 0.0))) for nd in nodes]},
        }
        cat_nodes = [i for i, nd in enumerate(nodes) if any(nd.get("bits") or [])]
        if cat_nodes:
            sets = []
            for i in cat_nodes:
                bits = nodes[i]["bits"]
                sets.append([w * 32 + b for w, word in enumerate(bits) for b in range(32) if word >> b & 1])
            d["categories"] = {"node_indices": cat_nodes, "category_sets": sets}
        if any(nd.get("gain", 0.0) or nd.get("cover", 0.0) for nd in nodes):
            d["gains"] = [float(np.float32(nd.get("gain", 0.0))) for nd in nodes]
            d["covers"] = [float(np.float32(nd.get("cover", 0.0))) for nd in nodes]
        ts.append(d)
    forest = {"trees": ts}
    if K > 1:
        forest["tree_groups"] = [t["group"] for t in trees]
    forest["n_groups"] = K
    forest["base_score"] = [float(np.float32(b)) for b in base]
    return {"meta": {"num_features": F, "num_groups": K}, "forest": forest,
            "output_transform": TRANSFORMS[transform]}


def split_file(b):
    assert b[:4] == b"BSTR"
    version, fmt, mtype = struct.unpack_from("<IBB", b, 4)
    assert b[10:32] == bytes(22)
    plen, crc, pad = struct.unpack_from("<QII", b, len(b) - 16)
    payload = b[32:-16]
    return dict(version=version, format=fmt, type=mtype, payload=payload, plen=plen, crc=crc, pad=pad)


def envelope(payload, fmt=1, version=2, mtype=1, crc=None, plen=None):
    head = b"BSTR" + struct.pack("<IBB", version, fmt, mtype) + bytes(22)
    crc = crc32c_bitwise(payload) if crc is None else crc
    plen = len(payload) if plen is None else plen
    return head + payload + struct.pack("<QII", plen, crc, 0)


def load(b):
    return GBDTModel.from_bytes(b)


def load_err(b):
    h = C.c_void_p()
    code = A.lib().bst_model_load_bstr(None, bytes(b), len(b), C.byref(h))
    assert not h.value
    return code, A.lib().bst_last_error().decode()


# ----------------------------------------------------------------- zstd via ctypes (tests only)

_Z = C.CDLL("libzstd.so.1")
_Z.ZSTD_compressBound.restype = C.c_size_t
_Z.ZSTD_compressBound.argtypes = [C.c_size_t]
_Z.ZSTD_decompress.restype = C.c_size_t
_Z.ZSTD_decompress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]
_Z.ZSTD_isError.restype = C.c_uint
_Z.ZSTD_isError.argtypes = [C.c_size_t]
_Z.ZSTD_createCCtx.restype = C.c_void_p
_Z.ZSTD_freeCCtx.argtypes = [C.c_void_p]
_Z.ZSTD_CCtx_setParameter.restype = C.c_size_t
_Z.ZSTD_CCtx_setParameter.argtypes = [C.c_void_p, C.c_int, C.c_int]
_Z.ZSTD_compress2.restype = C.c_size_t
_Z.ZSTD_compress2.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]
ZSTD_C_COMPRESSION_LEVEL, ZSTD_C_CONTENT_SIZE_FLAG = 100, 200


def zstd_frame(data: bytes, level=3, content_size=True) -> bytes:
    cc = _Z.ZSTD_createCCtx()
    try:
        assert not _Z.ZSTD_isError(_Z.ZSTD_CCtx_setParameter(cc, ZSTD_C_COMPRESSION_LEVEL, level))
        assert not _Z.ZSTD_isError(_Z.ZSTD_CCtx_setParameter(cc, ZSTD_C_CONTENT_SIZE_FLAG, int(content_size)))
        cap = _Z.ZSTD_compressBound(len(data))
        out = C.create_string_buffer(cap)
        n = _Z.ZSTD_compress2(cc, out, cap, data, len(data))
        assert not _Z.ZSTD_isError(n)
        return out.raw[:n]
    finally:
        _Z.ZSTD_freeCCtx(cc)


def zstd_decode(frame: bytes, size: int) -> bytes:
    out = C.create_string_buffer(size)
    n = _Z.ZSTD_decompress(out, size, frame, len(frame))
    assert not _Z.ZSTD_isError(n) and n == size
    return out.raw[:n]


# ----------------------------------------------------------------- writing


def test_layout_of_reference_example():
    m = simple_model()
    b = m.to_bytes(0)
    f = split_file(b)
    assert (f["version"], f["format"], f["type"], f["pad"]) == (2, 1, 1, 0)
    assert f["plen"] == len(f["payload"])
    assert f["crc"] == crc32c_bitwise(f["payload"])
    assert len(b) == 32 + len(f["payload"]) + 16


@pytest.mark.parametrize("seed,K,transform", [(0, 1, 0), (1, 1, 1), (2, 3, 2), (3, 16, 2)])
def test_payload_equals_independent_encoding(seed, K, transform):
    m, trees = random_model(seed, K=K, n_trees=3 * K + 2, transform=transform)
    f = split_file(m.to_bytes(0))
    base = export(m)["base_score"]
    want = schema_dict(trees, K, 9, transform, base)
    assert msgpack.unpackb(f["payload"]) == want
    assert msgpack.packb(want) == f["payload"]  # same bytes: minimal ints, f64 floats, str8 keys
    assert f["crc"] == crc32c_bitwise(f["payload"])


def test_key_order_follows_schema_declaration():
    m, _ = random_model(4, K=2, n_trees=4)
    d = msgpack.unpackb(split_file(m.to_bytes(0))["payload"])
    assert list(d) == ["meta", "forest", "output_transform"]
    assert list(d["meta"]) == ["num_features", "num_groups"]
    assert list(d["forest"]) == ["trees", "tree_groups", "n_groups", "base_score"]
    for t in d["forest"]["trees"]:
        keys = list(t)
        assert keys[:7] == ["num_nodes", "split_indices", "thresholds", "children_left",
                            "children_right", "default_left", "leaf_values"]
        assert keys[7:] in ([], ["categories"], ["gains", "covers"], ["categories", "gains", "covers"])
        assert list(t["leaf_values"]) == ["type", "values"]
    # One output group: no tree_groups (convert.rs:421-425).
    m1 = simple_model()
    d1 = msgpack.unpackb(split_file(m1.to_bytes(0))["payload"])
    assert list(d1["forest"]) == ["trees", "n_groups", "base_score"]
    assert d1["forest"]["trees"][0] == {
        "num_nodes": 3, "split_indices": [0, 0, 0], "thresholds": [0.5, 0.0, 0.0],
        "children_left": [1, 0, 0], "children_right": [2, 0, 0],
        "default_left": [True, False, False],
        "leaf_values": {"type": "scalar", "values": [0.0, 1.0, 2.0]}}


@pytest.mark.parametrize("level", [0, 1, 3, 19])
@pytest.mark.parametrize("seed,K", [(5, 1), (6, 4)])
def test_round_trip_binary(seed, K, level):
    m, _ = random_model(seed, K=K, n_trees=11)
    b = m.to_bytes(level)
    f = split_file(b)
    assert f["format"] == (1 if level == 0 else 2)
    m2 = load(b)
    same_model(m, m2)
    assert m2.to_bytes(level) == b  # save(load(save(m))) == save(m), byte for byte
    if level:
        assert f["payload"][:4] == b"\x28\xb5\x2f\xfd"  # zstd frame magic
        raw = split_file(m.to_bytes(0))["payload"]
        assert zstd_decode(f["payload"], len(raw)) == raw


@pytest.mark.parametrize("pretty", [False, True])
def test_round_trip_json(pretty):
    m, trees = random_model(7, K=3, n_trees=8, transform=2)
    js = m.to_json_bytes(pretty)
    env = json.loads(js)
    assert list(env) == ["bstr_version", "model_type", "model"]
    assert env["bstr_version"] == 2 and env["model_type"] == "gbdt"
    # Same schema values as the binary payload (floats round-trip exactly).
    assert env["model"] == msgpack.unpackb(split_file(m.to_bytes(0))["payload"])
    assert env["model"] == schema_dict(trees, 3, 9, 2, export(m)["base_score"])
    m2 = GBDTModel.from_json_bytes(js)
    same_model(m, m2)
    assert m2.to_json_bytes(pretty) == js
    if pretty:
        assert js.startswith(b'{\n  "bstr_version": 2,\n  "model_type": "gbdt",\n  "model": {\n')
    else:
        assert b"\n" not in js and b" " not in js


def test_json_numbers():
    m = import_model([tree([dict(f=1, thr=1e-7, l=1, r=2), dict(v=3.0), dict(v=-2.5e20)])], n_features=2,
                     base=[0.1])
    env = json.loads(m.to_json_bytes())
    t = env["model"]["forest"]["trees"][0]
    assert t["thresholds"][0] == float(np.float32(1e-7))
    assert t["leaf_values"]["values"][1:] == [3.0, float(np.float32(-2.5e20))]
    assert env["model"]["forest"]["base_score"] == [float(np.float32(0.1))]
    assert b'"values":[0.0,3.0,' in m.to_json_bytes()  # integral floats keep ".0"


def test_file_round_trip_and_pickle(tmp_path):
    m, _ = random_model(8, K=2, n_trees=5)
    p = str(tmp_path / "model.bstr")
    m.save(p)
    with open(p, "rb") as fh:
        assert fh.read() == m.to_bytes()
    same_model(m, GBDTModel.load(p))
    same_model(m, pickle.loads(pickle.dumps(m)))
    with pytest.raises(A.BoostersError):
        GBDTModel.load(str(tmp_path / "missing.bstr"))


def test_categories_written_as_lists():
    # Node 0 categorical with categories {1, 3, 40}: words [0b1010, 1 << 8].
    m = import_model([tree([dict(f=2, l=1, r=2, bits=[0b1010, 1 << 8]), dict(v=1.0), dict(v=2.0)])],
                     n_features=3)
    t = msgpack.unpackb(split_file(m.to_bytes(0))["payload"])["forest"]["trees"][0]
    assert t["categories"] == {"node_indices": [0], "category_sets": [[1, 3, 40]]}
    same_model(m, load(m.to_bytes(0)))


def test_category_bitsets_are_normalised():
    # Trailing zero words are dropped (categories_to_bitset, convert.rs:186-205);
    # a categorical node with no category is written without one and reads
    # back as a numeric split, as the reference's writer and reader do.
    m = import_model([tree([dict(f=0, l=1, r=2, bits=[5, 0, 0]), dict(f=1, l=3, r=4, bits=[0]),
                            dict(v=1.0), dict(v=2.0), dict(v=3.0)])], n_features=2)
    m2 = load(m.to_bytes(0))
    e = export(m2)
    assert list(e["cat_size"]) == [1, 0, 0, 0, 0]
    assert list(e["split_type"]) == [1, 0, 0, 0, 0]
    assert list(e["cat_words"]) == [5]


def test_stats_written_only_when_present():
    # gains/covers follow the frozen tree (mutable_tree.rs:333-354): present iff any is non-zero.
    a = import_model([tree([dict(f=0, thr=0.5, l=1, r=2), dict(v=1.0), dict(v=2.0)]),
                      tree([dict(f=0, thr=0.5, l=1, r=2, gain=3.0), dict(v=1.0, cover=4.0),
                            dict(v=2.0)])], n_features=1)
    ts = msgpack.unpackb(split_file(a.to_bytes(0))["payload"])["forest"]["trees"]
    assert "gains" not in ts[0] and "covers" not in ts[0]
    assert ts[1]["gains"] == [3.0, 0.0, 0.0] and ts[1]["covers"] == [0.0, 4.0, 0.0]


# ----------------------------------------------------------------- reading foreign payloads


def base_schema():
    return {
        "meta": {"num_features": 3, "num_groups": 1},
        "forest": {
            "trees": [{
                "num_nodes": 3, "split_indices": [2, 0, 0], "thresholds": [0.25, 0.0, 0.0],
                "children_left": [1, 0, 0], "children_right": [2, 0, 0],
                "default_left": [False, False, False],
                "leaf_values": {"type": "scalar", "values": [0.0, -1.0, 1.0]},
            }],
            "n_groups": 1, "base_score": [0.5],
        },
        "output_transform": "sigmoid",
    }


def test_reads_foreign_encodings():
    s = base_schema()
    ref = load(envelope(msgpack.packb(s)))
    e = export(ref)
    assert list(e["is_leaf"]) == [0, 1, 1] and list(e["leaf_value"]) == [0.0, -1.0, 1.0]
    assert ref.info.output_transform == A.TRANSFORM_SIGMOID and ref.info.n_features == 3
    # Reordered fields, unknown fields, explicit nils, ints and f32 for floats.
    t = s["forest"]["trees"][0]
    t2 = {"leaf_values": {"values": [0, -1, 1], "type": "scalar"}, "gains": None,
          "default_left": t["default_left"], "children_right": t["children_right"],
          "children_left": t["children_left"], "thresholds": [0.25, 0, 0], "split_indices": [2, 0, 0],
          "future_field": {"x": [1, 2, {"y": None}]}, "num_nodes": 3}
    s2 = {"output_transform": "sigmoid", "extra": b"\x00\x01",
          "forest": {"base_score": [0.5], "n_groups": 1, "tree_groups": None, "trees": [t2]},
          "meta": {"objective_name": None, "num_features": 3}}
    same_model(ref, load(envelope(msgpack.packb(s2, use_single_float=True))))


def test_reads_vector_leaves():
    # Vector leaves of width 1 read as scalar; of width n_groups expand into
    # one tree per group (convert.rs:475-565).
    s = base_schema()
    s["forest"]["trees"][0]["leaf_values"] = {"type": "vector", "values": [[0.0], [-1.0], [1.0]]}
    same_model(load(envelope(msgpack.packb(base_schema()))), load(envelope(msgpack.packb(s))))
    s = base_schema()
    s["meta"]["num_groups"] = 2
    s["forest"]["n_groups"] = 2
    s["forest"]["base_score"] = [0.0, 0.0]
    s["forest"]["tree_groups"] = [1]
    s["forest"]["trees"][0]["leaf_values"] = {"type": "vector", "values": [[0, 0], [1, 2], [3, 4]]}
    m = load(envelope(msgpack.packb(s)))
    e = export(m)
    assert m.info.n_trees == 2 and list(e["tree_group"]) == [0, 1]
    assert list(e["leaf_value"]) == [0, 1, 3, 0, 2, 4]
    s["forest"]["trees"][0]["leaf_values"]["values"] = [[0, 0, 0]] * 3
    assert load_err(envelope(msgpack.packb(s)))[0] == A.BST_E_FORMAT


def test_reads_zstd_frames_without_content_size():
    # The reference's streaming encoder writes frames without a content size;
    # zstd::decode_all also joins concatenated frames.
    raw = msgpack.packb(base_schema())
    ref = load(envelope(raw))
    same_model(ref, load(envelope(zstd_frame(raw, 3, content_size=False), fmt=2)))
    half = len(raw) // 2
    two = zstd_frame(raw[:half], 1, False) + zstd_frame(raw[half:], 5, True)
    same_model(ref, load(envelope(two, fmt=2)))
    assert load_err(envelope(two[:-3], fmt=2))[0] == A.BST_E_FORMAT  # truncated frame


def test_meta_round_trips():
    s = base_schema()
    s["meta"] = {"num_features": 3, "num_groups": 1, "feature_names": ["a", "b", "cé\"x"],
                 "feature_types": ["numeric", "categorical", "numeric"], "objective_name": "logistic"}
    payload = msgpack.packb(s)
    m = load(envelope(payload))
    assert split_file(m.to_bytes(0))["payload"] == payload
    js = json.loads(m.to_json_bytes())
    assert js["model"]["meta"] == s["meta"]
    assert GBDTModel.from_json_bytes(m.to_json_bytes()).to_bytes(0) == m.to_bytes(0)


# ----------------------------------------------------------------- rejections


def test_invalid_magic_rejected():
    code, msg = load_err(bytes([0, 1, 2, 3]) + bytes(100))
    assert code == A.BST_E_FORMAT and "magic" in msg


def test_corrupted_payload_rejected():
    b = bytearray(simple_model().to_bytes())
    for i in (40, 41, 42):
        b[i] ^= 0xFF
    code, msg = load_err(b)
    assert code == A.BST_E_FORMAT and "checksum" in msg


def test_truncated_file_rejected():
    b = simple_model().to_bytes()
    for cut in (len(b) // 2, 47, 31, 0):
        assert load_err(b[:cut])[0] == A.BST_E_FORMAT


@pytest.mark.parametrize("mutate,needle", [
    (lambda h: h[:4] + struct.pack("<I", 3) + h[8:], "version"),
    (lambda h: h[:8] + b"\x00" + h[9:], "Json"),
    (lambda h: h[:8] + b"\x07" + h[9:], "format"),
    (lambda h: h[:9] + b"\x02" + h[10:], "model type"),
    (lambda h: h[:9] + b"\x09" + h[10:], "model type"),
])
def test_header_rejections(mutate, needle):
    b = simple_model().to_bytes(0)
    code, msg = load_err(mutate(b[:32]) + b[32:])
    assert code == A.BST_E_FORMAT and needle in msg


def test_older_version_and_reserved_bytes_accepted():
    b = bytearray(simple_model().to_bytes(0))
    b[4:8] = struct.pack("<I", 1)
    b[20] = 0x7F  # reserved bytes are ignored (envelope.rs:165)
    same_model(simple_model(), load(bytes(b)))


def test_trailer_rejections():
    payload = msgpack.packb(base_schema())
    assert load_err(envelope(payload, plen=len(payload) + 1))[0] == A.BST_E_FORMAT
    assert load_err(envelope(payload, crc=crc32c_bitwise(payload) ^ 1))[0] == A.BST_E_FORMAT
    assert "missing payload" in load_err(envelope(b"", plen=5))[1]
    assert load_err(envelope(payload + b"\xc0"))[0] == A.BST_E_FORMAT  # trailing bytes


def _with(fn):
    s = base_schema()
    fn(s)
    return load_err(envelope(msgpack.packb(s)))


@pytest.mark.parametrize("edit", [
    lambda s: s["forest"]["trees"][0].update(thresholds=[0.25]),                    # length
    lambda s: s["forest"]["trees"][0].update(gains=[1.0]),                          # gains length
    lambda s: s["forest"]["trees"][0].update(children_right=[99, 0, 0]),            # child out of range
    lambda s: s["forest"]["trees"][0].update(children_left=[0 + 1, 1, 0],
                                             children_right=[2, 2, 0]),             # node 1 loops
    lambda s: s["forest"]["trees"][0].update(children_left=[1, 0, 0],
                                             children_right=[1, 0, 0]),             # shared child
    lambda s: s["forest"]["trees"][0].update(children_right=[0, 0, 0],
                                             children_left=[1, 0, 0]),              # right = root
    lambda s: s["forest"]["trees"][0].update(split_indices=[3, 0, 0]),              # feature >= num_features
    lambda s: s["forest"]["trees"][0].update(num_nodes=0, split_indices=[], thresholds=[],
                                             children_left=[], children_right=[], default_left=[],
                                             leaf_values={"type": "scalar", "values": []}),
    lambda s: s["forest"]["trees"][0].update(categories={"node_indices": [5], "category_sets": [[1]]}),
    lambda s: s["forest"]["trees"][0].update(categories={"node_indices": [0], "category_sets": []}),
    lambda s: s["forest"]["trees"][0].update(
        linear_coefficients={"node_indices": [1], "coefficients": [[0.5, 0.0, 1.0]]}),
    lambda s: s["forest"]["trees"][0].update(leaf_values={"type": "scalar", "values": [[1.0]] * 3}),
    lambda s: s["forest"]["trees"][0].update(leaf_values={"type": "tensor", "values": [1.0] * 3}),
    lambda s: s["forest"]["trees"][0].pop("default_left"),
    lambda s: s["forest"].update(base_score=[0.5, 0.5]),
    lambda s: s["forest"].update(tree_groups=[1]),
    lambda s: s["forest"].update(tree_groups=[0, 0]),
    lambda s: s["forest"].update(n_groups=0, base_score=[]),
    lambda s: s["forest"].update(n_groups=1025, base_score=[0.0] * 1025),
    lambda s: s.update(output_transform="tanh"),
    lambda s: s.pop("meta"),
    lambda s: s["meta"].update(feature_types=["numeric", "ordinal"]),
    lambda s: s["forest"]["trees"][0].update(default_left=[0, 0, 0]),               # ints for bools
    lambda s: s["forest"]["trees"][0].update(split_indices=[-1, 0, 0]),
    lambda s: s["forest"]["trees"][0].update(split_indices=[2**32, 0, 0]),
])
def test_schema_rejections(edit):
    code, msg = _with(edit)
    assert code == A.BST_E_FORMAT, msg
    assert msg


def test_category_ids_bounded_before_allocation():
    # A category id sizes its node's bitset ((id >> 5) + 1 words): ids at or
    # above 2^24 are refused before anything is allocated, and so are bitsets
    # far larger than the categories the file lists.
    def cat_model(sets):
        s = base_schema()
        t = s["forest"]["trees"][0]
        t["categories"] = {"node_indices": [0], "category_sets": sets}
        return s
    ok = load(envelope(msgpack.packb(cat_model([[1, 3, 2**16 - 1]]))))
    e = export(ok)
    assert list(e["split_type"]) == [1, 0, 0] and list(e["cat_size"])[0] == 2**11
    for bad_id in (2**24, 2**32 - 1):
        code, msg = load_err(envelope(msgpack.packb(cat_model([[bad_id]]))))
        assert code == A.BST_E_FORMAT and "category value" in msg
    # One tree per node with a single large id: the bitsets would need 2^19
    # words per node but the file lists one category per node.
    s = base_schema()
    t = s["forest"]["trees"][0]
    t["categories"] = {"node_indices": [0], "category_sets": [[2**24 - 1]]}
    code, msg = load_err(envelope(msgpack.packb(s)))
    assert code == A.BST_E_FORMAT and "too large" in msg


def test_import_rejects_cycles():
    # bst_model_import runs the same walk: a loop would never reach a leaf.
    with pytest.raises(A.BoostersError) as ei:
        import_model([tree([dict(f=0, thr=0.5, l=1, r=2), dict(f=0, thr=0.1, l=0, r=2), dict(v=1.0)])],
                     n_features=1)
    assert ei.value.code == A.BST_E_FORMAT


def test_invalid_json_rejected():
    bad = b'{"bstr_version": 1, "model_type": "gbdt", "model": "not_an_object"}'
    h = C.c_void_p()
    assert A.lib().bst_model_load_json(None, bad, len(bad), C.byref(h)) == A.BST_E_FORMAT
    for doc in [b'{"bstr_version": 3, "model_type": "gbdt", "model": %s}',
                b'{"bstr_version": 2, "model_type": "gblinear", "model": %s}',
                b'{"bstr_version": 2, "model": %s}',
                b'{"bstr_version": 2, "model_type": "gbdt", "model": %s} x',
                b'{"bstr_version": 2, "model_type": "gbdt", "model": %s']:
        js = doc % json.dumps(base_schema()).encode()
        assert A.lib().bst_model_load_json(None, js, len(js), C.byref(h)) == A.BST_E_FORMAT, doc
    good = b'{"bstr_version": 2, "model_type": "gbdt", "model": %s}' % json.dumps(base_schema()).encode()
    same_model(load(envelope(msgpack.packb(base_schema()))), GBDTModel.from_json_bytes(good))


def test_compression_level_range():
    m = simple_model()
    buf, size = C.POINTER(C.c_uint8)(), C.c_int64()
    for lvl in (-1, 23):
        assert A.lib().bst_model_save_bstr(m.h, lvl, C.byref(buf), C.byref(size)) == A.BST_E_CONFIG
    m.to_bytes(22)
