"""CPU: the Rust FFI crate (crates/boosters-mi355x-sys) mirrors the C header.

There is no Rust toolchain here, so the binding is checked mechanically:
every function include/boosters_amd.h declares is declared in the crate's
extern block with the same parameter types and return type, every struct has
the same fields in the same order with the same widths, and every constant
has the same value.  A dropped field, a swapped i32/u32 or a missing entry
point fails one of these.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "boosters_amd.h")
LIB_RS = os.path.join(ROOT, "crates", "boosters-mi355x-sys", "src", "lib.rs")
GLUE_RS = os.path.join(ROOT, "crates", "boosters-mi355x-sys", "integration", "mi355x.rs")

C_SCALAR = {
    "int32_t": "i32", "uint32_t": "u32", "int64_t": "i64", "uint64_t": "u64",
    "uint8_t": "u8", "uint16_t": "u16", "float": "f32", "double": "f64",
    "int": "c_int", "char": "c_char", "void": "c_void",
}


def _strip_c_comments(s):
    return re.sub(r"/\*.*?\*/", "", s, flags=re.S)


def _c_type_to_rust(t):
    """'const float*' -> '*const f32', 'bst_model**' -> '*mut *mut bst_model'."""
    t = t.strip()
    stars = t.count("*")
    base = t.replace("*", " ").split()
    const = "const" in base
    base = [b for b in base if b != "const"]
    assert len(base) == 1, t
    r = C_SCALAR.get(base[0], base[0])
    if stars == 0:
        return r
    # Only the innermost pointee carries the const in this header.
    out = ("*const " if const else "*mut ") + r
    for _ in range(stars - 1):
        out = "*mut " + out
    return out


def _split_param(p):
    p = p.strip()
    m = re.match(r"(.*?)([A-Za-z_][A-Za-z_0-9]*)\s*$", p)
    return m.group(1).strip(), m.group(2)


def header_functions():
    s = _strip_c_comments(open(HEADER).read())
    fns = {}
    for m in re.finditer(r"^([A-Za-z_][\w\s\*]*?)\b(bst_\w+)\s*\(([^;{]*?)\)\s*;", s, flags=re.M):
        ret, name, params = m.group(1).strip(), m.group(2), m.group(3)
        if params.strip() in ("", "void"):
            ptypes = []
        else:
            ptypes = [_c_type_to_rust(_split_param(p)[0]) for p in params.split(",")]
        fns[name] = (_c_type_to_rust(ret), ptypes)
    return fns


def header_structs():
    s = _strip_c_comments(open(HEADER).read())
    out = {}
    for m in re.finditer(r"typedef struct (\w+) \{(.*?)\}\s*\1;", s, flags=re.S):
        fields = []
        for line in m.group(2).split(";"):
            line = line.strip()
            if not line:
                continue
            arr = re.match(r"(.*?)(\w+)\[(\d+)\]$", line)
            if arr:
                fields.append((arr.group(2), "[%s; %s]" % (_c_type_to_rust(arr.group(1)), arr.group(3))))
            else:
                t, n = _split_param(line)
                fields.append((n, _c_type_to_rust(t)))
        out[m.group(1)] = fields
    return out


def header_constants():
    s = open(HEADER).read()
    return {m.group(1): int(m.group(2)) for m in
            re.finditer(r"#define (BST_\w+)\s+\(?(-?\d+)\)?", s)}


def _strip_rs_comments(s):
    return re.sub(r"//[^\n]*", "", s)


def rust_functions():
    s = _strip_rs_comments(open(LIB_RS).read())
    block = s[s.index('extern "C" {'):]
    fns = {}
    for m in re.finditer(r"pub fn (bst_\w+)\s*\((.*?)\)\s*(->\s*([^;]+?))?\s*;", block, flags=re.S):
        params = [p for p in m.group(2).split(",") if p.strip()]
        ptypes = [" ".join(p.split(":", 1)[1].split()) for p in params]
        ret = " ".join(m.group(4).split()) if m.group(4) else "c_void"
        fns[m.group(1)] = (ret, ptypes)
    return fns


def rust_structs():
    s = _strip_rs_comments(open(LIB_RS).read())
    out = {}
    for m in re.finditer(r"pub struct (bst_\w+)\s*\{(.*?)\}", s, flags=re.S):
        fields = []
        for f in m.group(2).split(","):
            f = f.strip()
            if not f or not f.startswith("pub "):
                continue
            n, t = f[4:].split(":", 1)
            fields.append((n.strip(), " ".join(t.split())))
        out[m.group(1)] = fields
    return out


def rust_constants():
    s = _strip_rs_comments(open(LIB_RS).read())
    return {m.group(1): int(m.group(2)) for m in
            re.finditer(r"pub const (BST_\w+): \w+ = (-?\d+);", s)}


def test_header_parses():
    fns = header_functions()
    assert len(fns) >= 50 and "bst_train" in fns and "bst_predict_device_noalloc" in fns
    assert fns["bst_train"][1][-1] == "*mut *mut bst_model"
    assert fns["bst_last_error"] == ("*const c_char", [])


def test_every_entry_point_is_bound_with_the_same_signature():
    h, r = header_functions(), rust_functions()
    assert sorted(h) == sorted(r), (sorted(set(h) - set(r)), sorted(set(r) - set(h)))
    for name, (ret, params) in h.items():
        rret, rparams = r[name]
        assert rret == ret, (name, rret, ret)
        assert rparams == params, (name, rparams, params)


def test_structs_match_field_for_field():
    h, r = header_structs(), rust_structs()
    assert set(h) == {"bst_binning_config", "bst_gbdt_config", "bst_ext_config", "bst_model_info",
                      "bst_forest_arrays", "bst_split_info"}
    for name, fields in h.items():
        assert r.get(name) == fields, name


def test_constants_match():
    h, r = header_constants(), rust_constants()
    h.pop("BOOSTERS_AMD_H", None)
    assert h == r


@pytest.mark.parametrize("seam", ["model.rs:149-187", "model.rs:238-258", "trainer.rs:162-401",
                                  "predictor.rs:197-247", "tree.rs:65-100"])
def test_glue_cites_the_reference_seams(seam):
    assert seam in open(GLUE_RS).read()


def test_glue_has_no_unimplemented_bodies():
    s = open(GLUE_RS).read()
    assert "unimplemented!" not in s and "todo!" not in s
    for fn in ("pub fn train(", "pub fn predict_into(", "fn export_forest(", "fn import_forest("):
        assert fn in s
