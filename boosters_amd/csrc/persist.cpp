// persist.cpp — model files: the binary .bstr envelope and its JSON form.
//
// Binary layout (persist/envelope.rs:13-25,111-130,194-216; binary.rs:52-146):
//   header  32 B  "BSTR", u32 LE schema version 2, format byte (1 MessagePack,
//                 2 zstd-compressed MessagePack), model type (1 = gbdt),
//                 22 zero bytes
//   payload       the GBDT model schema (persist/schema.rs:27-149): one
//                 MessagePack map per struct with string keys, fields in
//                 declaration order, optional fields left out when absent or
//                 empty; zstd frames over it for format 2
//   trailer 16 B  u64 LE stored payload length, u32 LE CRC32C (Castagnoli)
//                 of the stored payload, 4 zero bytes
// JSON: {"bstr_version":2,"model_type":"gbdt","model":{...}} with the same
// schema (persist/json.rs:50-75).
//
// Model <-> schema follows persist/convert.rs:77-167 (write) and 207-470
// (read): thresholds, leaf values, gains and covers widen f32 -> f64 on write
// and narrow back on read; a node is a leaf iff children_left == 0; category
// bitsets travel as sorted category lists; gains/covers are written when any
// of the tree's values is non-zero (mutable_tree.rs:333-354); tree_groups only
// when n_groups > 1; vector leaves of width 1 read as scalars and of width
// n_groups expand into one tree per group.
//
// The MessagePack and JSON codecs, CRC32C and the zstd binding are this
// file's own: MessagePack per the published specification, CRC32C by the
// reflected table method over polynomial 0x82F63B78, zstd through the
// system's libzstd.so.1 (loaded on first use).
#include <dlfcn.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "host_common.h"

using namespace bst;

namespace {

constexpr uint32_t kSchemaVersion = 2;
constexpr size_t kHeaderBytes = 32, kTrailerBytes = 16;
constexpr uint8_t kFormatJson = 0, kFormatMsgpack = 1, kFormatMsgpackZstd = 2;
constexpr uint8_t kModelGbdt = 1;
constexpr int kMaxNesting = 64;
constexpr size_t kMaxDecompressed = size_t(4) << 30;

[[noreturn]] void bad(const std::string& msg) { raise(BST_E_FORMAT, msg); }

// ---------------------------------------------------------------- CRC32C
struct Crc32cTable {
  uint32_t t[8][256];
  Crc32cTable() {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int k = 0; k < 8; k++) c = (c >> 1) ^ (0x82F63B78u & (0u - (c & 1u)));
      t[0][i] = c;
    }
    for (uint32_t i = 0; i < 256; i++)
      for (int s = 1; s < 8; s++) t[s][i] = (t[s - 1][i] >> 8) ^ t[0][t[s - 1][i] & 0xFF];
  }
};

uint32_t crc32c(const uint8_t* p, size_t n) {
  static const Crc32cTable T;
  uint32_t c = 0xFFFFFFFFu;
  while (n >= 8) {  // eight bytes per step (slicing by 8)
    uint64_t v;
    std::memcpy(&v, p, 8);
    v ^= c;
    c = T.t[7][v & 0xFF] ^ T.t[6][(v >> 8) & 0xFF] ^ T.t[5][(v >> 16) & 0xFF] ^
        T.t[4][(v >> 24) & 0xFF] ^ T.t[3][(v >> 32) & 0xFF] ^ T.t[2][(v >> 40) & 0xFF] ^
        T.t[1][(v >> 48) & 0xFF] ^ T.t[0][v >> 56];
    p += 8;
    n -= 8;
  }
  while (n--) c = (c >> 8) ^ T.t[0][(c ^ *p++) & 0xFF];
  return c ^ 0xFFFFFFFFu;
}

// ---------------------------------------------------------------- zstd
struct ZIn {
  const void* src;
  size_t size, pos;
};
struct ZOut {
  void* dst;
  size_t size, pos;
};

struct Zstd {
  bool ok = false;
  size_t (*bound)(size_t) = nullptr;
  size_t (*compress)(void*, size_t, const void*, size_t, int) = nullptr;
  unsigned (*is_error)(size_t) = nullptr;
  const char* (*error_name)(size_t) = nullptr;
  void* (*create_ds)() = nullptr;
  size_t (*free_ds)(void*) = nullptr;
  size_t (*init_ds)(void*) = nullptr;
  size_t (*decompress_stream)(void*, ZOut*, ZIn*) = nullptr;
  size_t (*ds_out_size)() = nullptr;
};

const Zstd& zstd(int code_if_missing) {
  static Zstd z;
  static std::once_flag once;
  std::call_once(once, [] {
    void* h = dlopen("libzstd.so.1", RTLD_NOW | RTLD_LOCAL);
    if (!h) return;
    auto sym = [&](const char* n) { return dlsym(h, n); };
    z.bound = (size_t(*)(size_t))sym("ZSTD_compressBound");
    z.compress = (size_t(*)(void*, size_t, const void*, size_t, int))sym("ZSTD_compress");
    z.is_error = (unsigned (*)(size_t))sym("ZSTD_isError");
    z.error_name = (const char* (*)(size_t))sym("ZSTD_getErrorName");
    z.create_ds = (void* (*)())sym("ZSTD_createDStream");
    z.free_ds = (size_t(*)(void*))sym("ZSTD_freeDStream");
    z.init_ds = (size_t(*)(void*))sym("ZSTD_initDStream");
    z.decompress_stream = (size_t(*)(void*, ZOut*, ZIn*))sym("ZSTD_decompressStream");
    z.ds_out_size = (size_t(*)())sym("ZSTD_DStreamOutSize");
    z.ok = z.bound && z.compress && z.is_error && z.error_name && z.create_ds && z.free_ds &&
           z.init_ds && z.decompress_stream && z.ds_out_size;
  });
  if (!z.ok) raise(code_if_missing, "zstd .bstr payloads need libzstd.so.1, which did not load");
  return z;
}

std::vector<uint8_t> zstd_compress(const std::vector<uint8_t>& in, int level) {
  const Zstd& z = zstd(BST_E_CONFIG);
  std::vector<uint8_t> out(z.bound(in.size()));
  const size_t r = z.compress(out.data(), out.size(), in.data(), in.size(), level);
  if (z.is_error(r)) raise(BST_E_INVALID, std::string("zstd compression failed: ") + z.error_name(r));
  out.resize(r);
  return out;
}

// Decodes every frame of `p` (zstd::decode_all semantics: concatenated frames
// are joined).
std::vector<uint8_t> zstd_decompress(const uint8_t* p, size_t n) {
  const Zstd& z = zstd(BST_E_FORMAT);
  void* ds = z.create_ds();
  if (!ds) raise(BST_E_OOM, "zstd stream allocation failed");
  std::unique_ptr<void, size_t (*)(void*)> hold(ds, z.free_ds);
  size_t r = z.init_ds(ds);
  if (z.is_error(r)) bad(std::string("zstd: ") + z.error_name(r));
  const size_t chunk = std::max<size_t>(z.ds_out_size(), 1 << 16);
  std::vector<uint8_t> out;
  size_t len = 0;
  ZIn in{p, n, 0};
  for (;;) {
    if (out.size() - len < chunk) out.resize(len + chunk);
    ZOut o{out.data() + len, out.size() - len, 0};
    r = z.decompress_stream(ds, &o, &in);
    if (z.is_error(r)) bad(std::string("zstd decompression failed: ") + z.error_name(r));
    len += o.pos;
    if (len > kMaxDecompressed) bad("zstd payload expands beyond 4 GiB");
    if (in.pos == in.size) {
      if (r == 0) break;                     // last frame complete
      if (o.pos < o.size) bad("truncated zstd frame");  // needs input that is not there
    }
  }
  out.resize(len);
  return out;
}

// ---------------------------------------------------------------- writers
// Both writers take the same calls: map(n)/key/.../end_map, arr(n)/.../end_arr
// and scalars.  The schema is emitted once, by write_schema below.
struct MsgpackWriter {
  std::vector<uint8_t>& b;
  void byte(uint8_t v) { b.push_back(v); }
  void be(uint64_t v, int bytes) {
    for (int i = bytes - 1; i >= 0; i--) b.push_back((uint8_t)(v >> (8 * i)));
  }
  void head(uint64_t n, uint8_t fix, uint8_t fix_max, uint8_t c8, uint8_t c16, uint8_t c32) {
    if (n <= fix_max) byte((uint8_t)(fix | n));
    else if (c8 && n <= 0xFF) { byte(c8); be(n, 1); }
    else if (n <= 0xFFFF) { byte(c16); be(n, 2); }
    else if (n <= 0xFFFFFFFFull) { byte(c32); be(n, 4); }
    else raise(BST_E_INVALID, "MessagePack container larger than 2^32 entries");
  }
  void map(size_t n) { head(n, 0x80, 15, 0, 0xDE, 0xDF); }
  void arr(size_t n) { head(n, 0x90, 15, 0, 0xDC, 0xDD); }
  void end_map() {}
  void end_arr() {}
  void str(const std::string& s) {
    head(s.size(), 0xA0, 31, 0xD9, 0xDA, 0xDB);
    b.insert(b.end(), s.begin(), s.end());
  }
  void key(const char* k) { str(k); }
  void u64(uint64_t v) {  // smallest unsigned form
    if (v < 128) byte((uint8_t)v);
    else if (v <= 0xFF) { byte(0xCC); be(v, 1); }
    else if (v <= 0xFFFF) { byte(0xCD); be(v, 2); }
    else if (v <= 0xFFFFFFFFull) { byte(0xCE); be(v, 4); }
    else { byte(0xCF); be(v, 8); }
  }
  void f64(double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    byte(0xCB);
    be(u, 8);
  }
  void boolean(bool v) { byte(v ? 0xC3 : 0xC2); }
};

struct JsonWriter {
  std::string& s;
  bool pretty;
  std::vector<bool> first;  // per open container: no element written yet
  bool after_key = false;
  void indent() {
    s += '\n';
    s.append(2 * first.size(), ' ');
  }
  void sep() {
    if (after_key) {
      after_key = false;
      return;
    }
    if (first.empty()) return;
    if (!first.back()) s += ',';
    first.back() = false;
    if (pretty) indent();
  }
  void open(char c) {
    sep();
    s += c;
    first.push_back(true);
  }
  void close(char c) {
    const bool empty = first.back();
    first.pop_back();
    if (pretty && !empty) indent();
    s += c;
  }
  void map(size_t) { open('{'); }
  void arr(size_t) { open('['); }
  void end_map() { close('}'); }
  void end_arr() { close(']'); }
  void quoted(const std::string& v) {
    s += '"';
    for (unsigned char c : v) {
      switch (c) {
        case '"': s += "\\\""; break;
        case '\\': s += "\\\\"; break;
        case '\n': s += "\\n"; break;
        case '\r': s += "\\r"; break;
        case '\t': s += "\\t"; break;
        case '\b': s += "\\b"; break;
        case '\f': s += "\\f"; break;
        default:
          if (c < 0x20) {
            char buf[8];
            std::snprintf(buf, sizeof buf, "\\u%04x", c);
            s += buf;
          } else {
            s += (char)c;
          }
      }
    }
    s += '"';
  }
  void str(const std::string& v) {
    sep();
    quoted(v);
  }
  void key(const char* k) {
    sep();
    quoted(k);
    s += pretty ? ": " : ":";
    after_key = true;
  }
  void u64(uint64_t v) {
    sep();
    s += std::to_string(v);
  }
  // Shortest decimal that reads back to the same double; integers keep a
  // ".0" so the value stays a float in the file; non-finite values -> null.
  void f64(double v) {
    sep();
    if (!std::isfinite(v)) {
      s += "null";
      return;
    }
    char buf[40];
    for (int prec = 1; prec <= 17; prec++) {
      std::snprintf(buf, sizeof buf, "%.*g", prec, v);
      if (std::strtod(buf, nullptr) == v) break;
    }
    std::string t(buf);
    if (t.find_first_of(".eE") == std::string::npos) t += ".0";
    s += t;
  }
  void boolean(bool v) {
    sep();
    s += v ? "true" : "false";
  }
};

// ---------------------------------------------------------------- readers
// Pull readers with one interface: begin_map()/next_key(k) and
// begin_arr()/next_elem() iterate containers, the scalar reads check types,
// skip() drops any value.  Sizes are checked against the bytes left before
// anything is reserved.
struct MsgpackReader {
  const uint8_t* p;
  const uint8_t* e;
  std::vector<uint64_t> left;  // entries left per open container
  int depth = 0;

  size_t rest() const { return (size_t)(e - p); }
  uint8_t peek() const {
    if (p >= e) bad("MessagePack payload ends early");
    return *p;
  }
  uint64_t be(int bytes) {
    if (rest() < (size_t)bytes) bad("MessagePack payload ends early");
    uint64_t v = 0;
    for (int i = 0; i < bytes; i++) v = (v << 8) | p[i];
    p += bytes;
    return v;
  }
  uint8_t take() {
    const uint8_t c = peek();
    p++;
    return c;
  }
  bool is_nil() const { return p < e && *p == 0xC0; }
  void nil() {
    if (take() != 0xC0) bad("expected nil");
  }
  uint64_t map_len(uint8_t c) {
    if ((c & 0xF0) == 0x80) return c & 0x0F;
    if (c == 0xDE) return be(2);
    if (c == 0xDF) return be(4);
    bad("expected a MessagePack map");
  }
  uint64_t arr_len(uint8_t c) {
    if ((c & 0xF0) == 0x90) return c & 0x0F;
    if (c == 0xDC) return be(2);
    if (c == 0xDD) return be(4);
    bad("expected a MessagePack array");
  }
  void push(uint64_t n, uint64_t min_bytes_each) {
    if (n > rest() / std::max<uint64_t>(min_bytes_each, 1)) bad("MessagePack length exceeds payload");
    if (++depth > kMaxNesting) bad("MessagePack nesting too deep");
    left.push_back(n);
  }
  void begin_map() { push(map_len(take()), 2); }
  void begin_arr() { push(arr_len(take()), 1); }
  size_t size_hint() const { return left.empty() ? 0 : (size_t)left.back(); }
  bool next_elem() {
    if (left.back() == 0) {
      left.pop_back();
      depth--;
      return false;
    }
    left.back()--;
    return true;
  }
  bool next_key(std::string& k) {
    if (!next_elem()) return false;
    k = str();
    return true;
  }
  std::string str() {
    const uint8_t c = take();
    uint64_t n;
    if ((c & 0xE0) == 0xA0) n = c & 0x1F;
    else if (c == 0xD9) n = be(1);
    else if (c == 0xDA) n = be(2);
    else if (c == 0xDB) n = be(4);
    else bad("expected a MessagePack string");
    if (n > rest()) bad("MessagePack string exceeds payload");
    std::string s((const char*)p, (size_t)n);
    p += n;
    return s;
  }
  // Any integer form; negative values are rejected (every integer of the
  // schema is unsigned).
  uint64_t u64() {
    const uint8_t c = take();
    if (c < 0x80) return c;
    switch (c) {
      case 0xCC: return be(1);
      case 0xCD: return be(2);
      case 0xCE: return be(4);
      case 0xCF: return be(8);
      case 0xD0: case 0xD1: case 0xD2: case 0xD3: {
        const int bytes = 1 << (c - 0xD0);
        const uint64_t raw = be(bytes);
        const int64_t v = bytes == 8 ? (int64_t)raw : ((int64_t)(raw << (64 - 8 * bytes)) >> (64 - 8 * bytes));
        if (v < 0) bad("negative integer where an unsigned one is required");
        return (uint64_t)v;
      }
      default: bad("expected an unsigned integer");
    }
  }
  // Floats of either width, or integers (serde's f64 accepts both).
  double f64() {
    const uint8_t c = peek();
    if (c == 0xCB) {
      p++;
      const uint64_t u = be(8);
      double d;
      std::memcpy(&d, &u, 8);
      return d;
    }
    if (c == 0xCA) {
      p++;
      const uint32_t u = (uint32_t)be(4);
      float f;
      std::memcpy(&f, &u, 4);
      return f;
    }
    if (c >= 0xE0) {  // negative fixint
      p++;
      return (double)(int8_t)c;
    }
    if (c >= 0xD0 && c <= 0xD3) {
      p++;
      const int bytes = 1 << (c - 0xD0);
      const uint64_t raw = be(bytes);
      return bytes == 8 ? (double)(int64_t)raw
                        : (double)((int64_t)(raw << (64 - 8 * bytes)) >> (64 - 8 * bytes));
    }
    return (double)u64();
  }
  bool boolean() {
    const uint8_t c = take();
    if (c == 0xC3) return true;
    if (c == 0xC2) return false;
    bad("expected a bool");
  }
  bool peek_container() const { return p < e && (((*p & 0xF0) == 0x90) || *p == 0xDC || *p == 0xDD); }
  bool peek_string() const {
    return p < e && (((*p & 0xE0) == 0xA0) || *p == 0xD9 || *p == 0xDA || *p == 0xDB);
  }
  void skip_bytes(uint64_t n) {
    if (n > rest()) bad("MessagePack value exceeds payload");
    p += n;
  }
  void skip(int d = 0) {
    if (d > kMaxNesting) bad("MessagePack nesting too deep");
    const uint8_t c = take();
    if (c < 0x80 || c >= 0xE0 || c == 0xC0 || c == 0xC2 || c == 0xC3) return;
    if ((c & 0xF0) == 0x80 || c == 0xDE || c == 0xDF) {
      p--;
      const uint64_t n = map_len(take());
      if (n > rest() / 2) bad("MessagePack length exceeds payload");
      for (uint64_t i = 0; i < 2 * n; i++) skip(d + 1);
      return;
    }
    if ((c & 0xF0) == 0x90 || c == 0xDC || c == 0xDD) {
      p--;
      const uint64_t n = arr_len(take());
      if (n > rest()) bad("MessagePack length exceeds payload");
      for (uint64_t i = 0; i < n; i++) skip(d + 1);
      return;
    }
    if ((c & 0xE0) == 0xA0) return skip_bytes(c & 0x1F);
    switch (c) {
      case 0xC4: case 0xD9: return skip_bytes(be(1));
      case 0xC5: case 0xDA: return skip_bytes(be(2));
      case 0xC6: case 0xDB: return skip_bytes(be(4));
      case 0xC7: { const uint64_t n = be(1); return skip_bytes(n + 1); }
      case 0xC8: { const uint64_t n = be(2); return skip_bytes(n + 1); }
      case 0xC9: { const uint64_t n = be(4); return skip_bytes(n + 1); }
      case 0xCA: return skip_bytes(4);
      case 0xCB: return skip_bytes(8);
      case 0xCC: case 0xD0: return skip_bytes(1);
      case 0xCD: case 0xD1: return skip_bytes(2);
      case 0xCE: case 0xD2: return skip_bytes(4);
      case 0xCF: case 0xD3: return skip_bytes(8);
      case 0xD4: return skip_bytes(2);
      case 0xD5: return skip_bytes(3);
      case 0xD6: return skip_bytes(5);
      case 0xD7: return skip_bytes(9);
      case 0xD8: return skip_bytes(17);
      default: bad("invalid MessagePack byte");
    }
  }
  void finish() {
    if (p != e) bad("trailing bytes after the MessagePack model");
  }
};

struct JsonReader {
  const char* p;
  const char* e;
  std::vector<bool> first;
  int depth = 0;

  void ws() {
    while (p < e && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) p++;
  }
  char peek() {
    ws();
    if (p >= e) bad("JSON ends early");
    return *p;
  }
  void expect(char c) {
    if (peek() != c) bad(std::string("JSON: expected '") + c + "'");
    p++;
  }
  bool word(const char* w) {
    ws();
    const size_t n = std::strlen(w);
    if ((size_t)(e - p) >= n && std::memcmp(p, w, n) == 0) {
      p += n;
      return true;
    }
    return false;
  }
  bool is_nil() { return peek() == 'n'; }
  void nil() {
    if (!word("null")) bad("JSON: expected null");
  }
  void open(char c) {
    expect(c);
    if (++depth > kMaxNesting) bad("JSON nesting too deep");
    first.push_back(true);
  }
  void begin_map() { open('{'); }
  void begin_arr() { open('['); }
  size_t size_hint() const { return 0; }
  bool next_any(char close) {
    if (peek() == close) {
      p++;
      first.pop_back();
      depth--;
      return false;
    }
    if (!first.back()) expect(',');
    first.back() = false;
    return true;
  }
  bool next_elem() { return next_any(']'); }
  bool next_key(std::string& k) {
    if (!next_any('}')) return false;
    k = str();
    expect(':');
    return true;
  }
  static void utf8(std::string& s, uint32_t cp) {
    if (cp < 0x80) s += (char)cp;
    else if (cp < 0x800) { s += (char)(0xC0 | (cp >> 6)); s += (char)(0x80 | (cp & 0x3F)); }
    else if (cp < 0x10000) {
      s += (char)(0xE0 | (cp >> 12));
      s += (char)(0x80 | ((cp >> 6) & 0x3F));
      s += (char)(0x80 | (cp & 0x3F));
    } else {
      s += (char)(0xF0 | (cp >> 18));
      s += (char)(0x80 | ((cp >> 12) & 0x3F));
      s += (char)(0x80 | ((cp >> 6) & 0x3F));
      s += (char)(0x80 | (cp & 0x3F));
    }
  }
  uint32_t hex4() {
    if (e - p < 4) bad("JSON: bad \\u escape");
    uint32_t v = 0;
    for (int i = 0; i < 4; i++) {
      const char c = *p++;
      v <<= 4;
      if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
      else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
      else bad("JSON: bad \\u escape");
    }
    return v;
  }
  std::string str() {
    expect('"');
    std::string s;
    for (;;) {
      if (p >= e) bad("JSON: unterminated string");
      const char c = *p++;
      if (c == '"') return s;
      if ((unsigned char)c < 0x20) bad("JSON: control character in string");
      if (c != '\\') {
        s += c;
        continue;
      }
      if (p >= e) bad("JSON: unterminated string");
      const char x = *p++;
      switch (x) {
        case '"': s += '"'; break;
        case '\\': s += '\\'; break;
        case '/': s += '/'; break;
        case 'b': s += '\b'; break;
        case 'f': s += '\f'; break;
        case 'n': s += '\n'; break;
        case 'r': s += '\r'; break;
        case 't': s += '\t'; break;
        case 'u': {
          uint32_t cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00) {
            if (e - p < 6 || p[0] != '\\' || p[1] != 'u') bad("JSON: lone surrogate");
            p += 2;
            const uint32_t lo = hex4();
            if (lo < 0xDC00 || lo >= 0xE000) bad("JSON: lone surrogate");
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          } else if (cp >= 0xDC00 && cp < 0xE000) {
            bad("JSON: lone surrogate");
          }
          utf8(s, cp);
          break;
        }
        default: bad("JSON: bad escape");
      }
    }
  }
  // A number token: [-]digits[.digits][(e|E)[+-]digits].
  std::string number(bool* integral) {
    ws();
    const char* b = p;
    if (p < e && *p == '-') p++;
    const char* d = p;
    while (p < e && *p >= '0' && *p <= '9') p++;
    if (p == d) bad("JSON: expected a number");
    *integral = true;
    if (p < e && *p == '.') {
      *integral = false;
      p++;
      const char* f = p;
      while (p < e && *p >= '0' && *p <= '9') p++;
      if (p == f) bad("JSON: bad number");
    }
    if (p < e && (*p == 'e' || *p == 'E')) {
      *integral = false;
      p++;
      if (p < e && (*p == '+' || *p == '-')) p++;
      const char* x = p;
      while (p < e && *p >= '0' && *p <= '9') p++;
      if (p == x) bad("JSON: bad number");
    }
    return std::string(b, p);
  }
  uint64_t u64() {
    bool integral;
    const std::string t = number(&integral);
    if (!integral || t[0] == '-') bad("JSON: expected an unsigned integer");
    errno = 0;
    const unsigned long long v = std::strtoull(t.c_str(), nullptr, 10);
    if (errno == ERANGE) bad("JSON: integer out of range");
    return v;
  }
  double f64() {
    bool integral;
    const std::string t = number(&integral);
    return std::strtod(t.c_str(), nullptr);
  }
  bool boolean() {
    if (word("true")) return true;
    if (word("false")) return false;
    bad("JSON: expected a bool");
  }
  bool peek_container() { return peek() == '['; }
  bool peek_string() { return peek() == '"'; }
  void skip(int d = 0) {
    if (d > kMaxNesting) bad("JSON nesting too deep");
    const char c = peek();
    if (c == '{') {
      p++;
      bool f = true;
      while (peek() != '}') {
        if (!f) expect(',');
        f = false;
        str();
        expect(':');
        skip(d + 1);
      }
      p++;
    } else if (c == '[') {
      p++;
      bool f = true;
      while (peek() != ']') {
        if (!f) expect(',');
        f = false;
        skip(d + 1);
      }
      p++;
    } else if (c == '"') {
      str();
    } else if (word("true") || word("false") || word("null")) {
    } else {
      bool integral;
      number(&integral);
    }
  }
  void finish() {
    ws();
    if (p != e) bad("trailing characters after the JSON model");
  }
};

// ---------------------------------------------------------------- schema
// One tree as the schema holds it (persist/schema.rs:80-110).
struct TreeSchema {
  uint64_t num_nodes = 0;
  std::vector<uint32_t> split, left, right;
  std::vector<double> thr;
  std::vector<uint8_t> dleft;
  int leaf_kind = 0;  // 1 scalar, 2 vector
  std::vector<double> leaf;
  std::vector<std::vector<double>> vleaf;
  std::vector<uint32_t> cat_nodes;
  std::vector<std::vector<uint32_t>> cat_sets;
  std::vector<uint32_t> lin_nodes;
  std::vector<size_t> lin_len;
  bool has_gains = false, has_covers = false;
  std::vector<double> gains, covers;
};

struct ModelSchema {
  uint64_t num_features = 0, num_groups = 1;
  bool has_names = false, has_types = false, has_objective = false;
  std::vector<std::string> names;
  std::vector<uint8_t> types;
  std::string objective;
  std::vector<TreeSchema> trees;
  bool has_tree_groups = false;
  std::vector<uint64_t> tree_groups;
  uint64_t n_groups = 0;
  std::vector<double> base_score;
  int transform = 0;
};

const char* kTransformNames[3] = {"identity", "sigmoid", "softmax"};

template <class W>
void write_u32s(W& w, const uint32_t* v, size_t n) {
  w.arr(n);
  for (size_t i = 0; i < n; i++) w.u64(v[i]);
  w.end_arr();
}

template <class W>
void write_f64s(W& w, const float* v, size_t n) {
  w.arr(n);
  for (size_t i = 0; i < n; i++) w.f64((double)v[i]);
  w.end_arr();
}

// GBDTModelSchema::from(&GBDTModel) (convert.rs:710-721) emitted field by field.
template <class W>
void write_schema(W& w, const bst_model& m) {
  const int64_t K = m.info.n_groups;
  w.map(3);
  w.key("meta");
  {
    const size_t nm = 2 + m.has_feature_names + m.has_feature_types + m.has_objective_name;
    w.map(nm);
    w.key("num_features");
    w.u64((uint64_t)m.info.n_features);
    w.key("num_groups");
    w.u64((uint64_t)(m.meta_num_groups >= 0 ? m.meta_num_groups : K));
    if (m.has_feature_names) {
      w.key("feature_names");
      w.arr(m.feature_names.size());
      for (auto& s : m.feature_names) w.str(s);
      w.end_arr();
    }
    if (m.has_feature_types) {
      w.key("feature_types");
      w.arr(m.feature_types.size());
      for (uint8_t t : m.feature_types) w.str(t ? "categorical" : "numeric");
      w.end_arr();
    }
    if (m.has_objective_name) {
      w.key("objective_name");
      w.str(m.objective_name);
    }
    w.end_map();
  }
  w.key("forest");
  {
    const int64_t T = m.info.n_trees;
    w.map(K > 1 ? 4 : 3);
    w.key("trees");
    w.arr((size_t)T);
    std::vector<uint32_t> lz, cats;
    for (int64_t t = 0; t < T; t++) {
      const int64_t b = m.tree_offset[t], n = m.tree_offset[t + 1] - b;
      // children_left == 0 marks a leaf (convert.rs:377).
      lz.assign(m.left.begin() + b, m.left.begin() + b + n);
      for (int64_t i = 0; i < n; i++)
        if (m.is_leaf[b + i]) lz[i] = 0;
      // Category lists of the nodes with a non-empty bitset (convert.rs:100-122).
      std::vector<uint32_t> cat_nodes;
      for (int64_t i = 0; i < n; i++) {
        const uint32_t sz = m.cat_size[b + i];
        bool any = false;
        for (uint32_t k = 0; k < sz && !any; k++) any = m.cat_words[m.cat_start[b + i] + k] != 0;
        if (any) cat_nodes.push_back((uint32_t)i);
      }
      bool stats = false;
      for (int64_t i = 0; i < n && !stats; i++) stats = m.gain[b + i] != 0.0f || m.cover[b + i] != 0.0f;
      w.map(7 + (cat_nodes.empty() ? 0 : 1) + (stats ? 2 : 0));
      w.key("num_nodes");
      w.u64((uint64_t)n);
      w.key("split_indices");
      write_u32s(w, m.split_index.data() + b, (size_t)n);
      w.key("thresholds");
      write_f64s(w, m.split_threshold.data() + b, (size_t)n);
      w.key("children_left");
      write_u32s(w, lz.data(), (size_t)n);
      w.key("children_right");
      write_u32s(w, m.right.data() + b, (size_t)n);
      w.key("default_left");
      w.arr((size_t)n);
      for (int64_t i = 0; i < n; i++) w.boolean(m.default_left[b + i] != 0);
      w.end_arr();
      w.key("leaf_values");
      w.map(2);
      w.key("type");
      w.str("scalar");
      w.key("values");
      write_f64s(w, m.leaf_value.data() + b, (size_t)n);
      w.end_map();
      if (!cat_nodes.empty()) {
        w.key("categories");
        w.map(2);
        w.key("node_indices");
        write_u32s(w, cat_nodes.data(), cat_nodes.size());
        w.key("category_sets");
        w.arr(cat_nodes.size());
        for (uint32_t i : cat_nodes) {
          cats.clear();
          const uint32_t* words = m.cat_words.data() + m.cat_start[b + i];
          for (uint32_t k = 0; k < m.cat_size[b + i]; k++)
            for (uint32_t bit = 0; bit < 32; bit++)
              if ((words[k] >> bit) & 1u) cats.push_back(k * 32 + bit);
          write_u32s(w, cats.data(), cats.size());
        }
        w.end_arr();
        w.end_map();
      }
      if (stats) {
        w.key("gains");
        write_f64s(w, m.gain.data() + b, (size_t)n);
        w.key("covers");
        write_f64s(w, m.cover.data() + b, (size_t)n);
      }
      w.end_map();
    }
    w.end_arr();
    if (K > 1) {
      w.key("tree_groups");
      write_u32s(w, m.tree_group.data(), (size_t)T);
    }
    w.key("n_groups");
    w.u64((uint64_t)K);
    w.key("base_score");
    write_f64s(w, m.base_score.data(), (size_t)K);
    w.end_map();
  }
  w.key("output_transform");
  const int tr = m.info.output_transform;
  w.str(kTransformNames[(tr >= 0 && tr <= 2) ? tr : 0]);
  w.end_map();
}

// -- reading: serde's derive rules (any field order, unknown fields ignored,
// duplicates and missing required fields rejected, Option fields accept nil).
[[noreturn]] void missing(const char* field) { bad(std::string("missing field `") + field + "`"); }

struct Seen {
  std::vector<std::string> keys;
  void mark(const std::string& k) {
    if (std::find(keys.begin(), keys.end(), k) != keys.end()) bad("duplicate field `" + k + "`");
    keys.push_back(k);
  }
  bool has(const char* k) const { return std::find(keys.begin(), keys.end(), k) != keys.end(); }
};

template <class R>
void read_u32s(R& r, std::vector<uint32_t>& v) {
  r.begin_arr();
  v.clear();
  v.reserve(r.size_hint());
  while (r.next_elem()) {
    const uint64_t x = r.u64();
    if (x > 0xFFFFFFFFull) bad("value does not fit u32");
    v.push_back((uint32_t)x);
  }
}

template <class R>
void read_u64s(R& r, std::vector<uint64_t>& v) {
  r.begin_arr();
  v.clear();
  v.reserve(r.size_hint());
  while (r.next_elem()) v.push_back(r.u64());
}

template <class R>
void read_f64s(R& r, std::vector<double>& v) {
  r.begin_arr();
  v.clear();
  v.reserve(r.size_hint());
  while (r.next_elem()) v.push_back(r.f64());
}

template <class R>
void read_leaf_values(R& r, TreeSchema& t) {
  std::string k, type;
  Seen seen;
  bool vec_form = false, num_form = false;
  r.begin_map();
  while (r.next_key(k)) {
    seen.mark(k);
    if (k == "type") {
      type = r.str();
    } else if (k == "values") {
      r.begin_arr();
      while (r.next_elem()) {
        if (r.peek_container()) {
          vec_form = true;
          t.vleaf.emplace_back();
          read_f64s(r, t.vleaf.back());
        } else {
          num_form = true;
          t.leaf.push_back(r.f64());
        }
      }
    } else {
      r.skip();
    }
  }
  if (!seen.has("type")) missing("type");
  if (!seen.has("values")) missing("values");
  if (type == "scalar") {
    if (vec_form) bad("scalar leaf_values hold arrays");
    t.leaf_kind = 1;
  } else if (type == "vector") {
    if (num_form) bad("vector leaf_values hold numbers");
    t.leaf_kind = 2;
  } else {
    bad("unknown leaf_values type `" + type + "`");
  }
}

template <class R>
void read_tree(R& r, TreeSchema& t) {
  std::string k;
  Seen seen;
  r.begin_map();
  while (r.next_key(k)) {
    seen.mark(k);
    if (k == "num_nodes") {
      t.num_nodes = r.u64();
      if (t.num_nodes > 0xFFFFFFFFull) bad("num_nodes does not fit u32");
    } else if (k == "split_indices") {
      read_u32s(r, t.split);
    } else if (k == "thresholds") {
      read_f64s(r, t.thr);
    } else if (k == "children_left") {
      read_u32s(r, t.left);
    } else if (k == "children_right") {
      read_u32s(r, t.right);
    } else if (k == "default_left") {
      r.begin_arr();
      t.dleft.reserve(r.size_hint());
      while (r.next_elem()) t.dleft.push_back(r.boolean() ? 1 : 0);
    } else if (k == "leaf_values") {
      read_leaf_values(r, t);
    } else if (k == "categories") {
      Seen s2;
      std::string k2;
      r.begin_map();
      while (r.next_key(k2)) {
        s2.mark(k2);
        if (k2 == "node_indices") {
          read_u32s(r, t.cat_nodes);
        } else if (k2 == "category_sets") {
          r.begin_arr();
          while (r.next_elem()) {
            t.cat_sets.emplace_back();
            read_u32s(r, t.cat_sets.back());
          }
        } else {
          r.skip();
        }
      }
      if (!s2.has("node_indices")) missing("node_indices");
      if (!s2.has("category_sets")) missing("category_sets");
    } else if (k == "linear_coefficients") {
      Seen s2;
      std::string k2;
      r.begin_map();
      while (r.next_key(k2)) {
        s2.mark(k2);
        if (k2 == "node_indices") {
          read_u32s(r, t.lin_nodes);
        } else if (k2 == "coefficients") {
          r.begin_arr();
          std::vector<double> tmp;
          while (r.next_elem()) {
            read_f64s(r, tmp);
            t.lin_len.push_back(tmp.size());
          }
        } else {
          r.skip();
        }
      }
      if (!s2.has("node_indices")) missing("node_indices");
      if (!s2.has("coefficients")) missing("coefficients");
    } else if (k == "gains" || k == "covers") {
      const bool g = k == "gains";
      if (r.is_nil()) {
        r.nil();
      } else {
        (g ? t.has_gains : t.has_covers) = true;
        read_f64s(r, g ? t.gains : t.covers);
      }
    } else {
      r.skip();
    }
  }
  for (const char* f : {"num_nodes", "split_indices", "thresholds", "children_left", "children_right",
                        "default_left", "leaf_values"})
    if (!seen.has(f)) missing(f);
}

template <class R>
void read_schema(R& r, ModelSchema& ms) {
  std::string k;
  Seen seen;
  r.begin_map();
  while (r.next_key(k)) {
    seen.mark(k);
    if (k == "meta") {
      Seen s2;
      std::string k2;
      r.begin_map();
      while (r.next_key(k2)) {
        s2.mark(k2);
        if (k2 == "num_features") {
          ms.num_features = r.u64();
        } else if (k2 == "num_groups") {
          ms.num_groups = r.u64();
        } else if (k2 == "feature_names") {
          if (r.is_nil()) { r.nil(); continue; }
          ms.has_names = true;
          r.begin_arr();
          while (r.next_elem()) ms.names.push_back(r.str());
        } else if (k2 == "feature_types") {
          if (r.is_nil()) { r.nil(); continue; }
          ms.has_types = true;
          r.begin_arr();
          while (r.next_elem()) {
            const std::string ft = r.str();
            if (ft == "numeric") ms.types.push_back(0);
            else if (ft == "categorical") ms.types.push_back(1);
            else bad("unknown feature type `" + ft + "`");
          }
        } else if (k2 == "objective_name") {
          if (r.is_nil()) { r.nil(); continue; }
          ms.has_objective = true;
          ms.objective = r.str();
        } else {
          r.skip();
        }
      }
      if (!s2.has("num_features")) missing("num_features");
    } else if (k == "forest") {
      Seen s2;
      std::string k2;
      r.begin_map();
      while (r.next_key(k2)) {
        s2.mark(k2);
        if (k2 == "trees") {
          r.begin_arr();
          while (r.next_elem()) {
            ms.trees.emplace_back();
            read_tree(r, ms.trees.back());
          }
        } else if (k2 == "tree_groups") {
          if (r.is_nil()) { r.nil(); continue; }
          ms.has_tree_groups = true;
          read_u64s(r, ms.tree_groups);
        } else if (k2 == "n_groups") {
          ms.n_groups = r.u64();
        } else if (k2 == "base_score") {
          read_f64s(r, ms.base_score);
        } else {
          r.skip();
        }
      }
      for (const char* f : {"trees", "n_groups", "base_score"})
        if (!s2.has(f)) missing(f);
    } else if (k == "output_transform") {
      // Unit variants travel as their snake_case names; an integer variant
      // index is accepted as well.
      if (r.peek_string()) {
        const std::string v = r.str();
        int found = -1;
        for (int i = 0; i < 3; i++)
          if (v == kTransformNames[i]) found = i;
        if (found < 0) bad("unknown output_transform `" + v + "`");
        ms.transform = found;
      } else {
        const uint64_t v = r.u64();
        if (v > 2) bad("unknown output_transform index");
        ms.transform = (int)v;
      }
    } else {
      r.skip();
    }
  }
  for (const char* f : {"meta", "forest", "output_transform"})
    if (!seen.has(f)) missing(f);
}

// Appends one schema tree with scalar leaves to the model (TryFrom<TreeSchema>,
// convert.rs:207-410).
// Largest category id a model file may hold (exclusive).
constexpr uint32_t kMaxCategoryId = 1u << 24;

// cat_entries: category entries of the file seen so far (bounds the bitsets).
void append_tree(bst_model& m, const TreeSchema& t, const std::vector<double>& leaf, uint32_t group,
                 size_t& cat_entries) {
  const size_t n = (size_t)t.num_nodes;
  const size_t b = m.split_index.size();
  m.tree_offset.push_back((int64_t)(b + n));
  m.tree_group.push_back(group);
  for (size_t i = 0; i < n; i++) {
    m.split_index.push_back(t.split[i]);
    m.split_threshold.push_back((float)t.thr[i]);
    m.left.push_back(t.left[i]);
    m.right.push_back(t.right[i]);
    m.default_left.push_back(t.dleft[i]);
    m.is_leaf.push_back(t.left[i] == 0 ? 1 : 0);
    m.split_type.push_back(0);
    m.leaf_value.push_back((float)leaf[i]);
    m.gain.push_back(t.has_gains ? (float)t.gains[i] : 0.0f);
    m.cover.push_back(t.has_covers ? (float)t.covers[i] : 0.0f);
    m.cat_start.push_back(0);
    m.cat_size.push_back(0);
  }
  // Category lists -> bitsets of max/32 + 1 words; a later entry for the same
  // node replaces the earlier one.
  for (size_t j = 0; j < t.cat_nodes.size(); j++) {
    const uint32_t node = t.cat_nodes[j];
    const auto& cats = t.cat_sets[j];
    m.split_type[b + node] = 1;
    m.cat_start[b + node] = (uint32_t)m.cat_words.size();
    if (cats.empty()) {
      m.cat_size[b + node] = 0;
      continue;
    }
    const uint32_t mx = *std::max_element(cats.begin(), cats.end());
    // Bounds before allocating: a category id sizes its node's bitset, so a
    // hostile file could ask for 512 MiB per node with one entry.  Ids are
    // capped at 2^24 (trained models hold bin indices, below 2^16), and the
    // bitsets of a whole file at 2^16 words plus 64 words per listed entry.
    if (mx >= kMaxCategoryId)
      bad("category value " + std::to_string(mx) + " exceeds the supported maximum " +
          std::to_string(kMaxCategoryId - 1));
    const uint32_t nw = (mx >> 5) + 1;
    cat_entries += cats.size();
    if (m.cat_words.size() + nw > 65536 + 64 * cat_entries)
      bad("category bitsets are too large for the number of categories listed");
    const size_t w0 = m.cat_words.size();
    m.cat_words.resize(w0 + nw, 0u);
    for (uint32_t c : cats) m.cat_words[w0 + (c >> 5)] |= 1u << (c & 31);
    m.cat_size[b + node] = nw;
  }
}

std::unique_ptr<bst_model> model_from_schema(ModelSchema& ms) {
  // ForestSchema -> Forest (convert.rs:432-470).
  if (ms.n_groups < 1 || ms.n_groups > (uint64_t)kMaxOutputs)
    bad("n_groups must be in 1.." + std::to_string(kMaxOutputs));
  const uint32_t K = (uint32_t)ms.n_groups;
  if (ms.base_score.size() != K) bad("Forest schema base_score must have length n_groups");
  if (ms.has_tree_groups && ms.tree_groups.size() != ms.trees.size())
    bad("Forest schema tree_groups must have same length as trees");
  auto m = std::make_unique<bst_model>();
  m->base_score.reserve(K);
  for (double v : ms.base_score) m->base_score.push_back((float)v);
  m->tree_offset.push_back(0);
  size_t cat_entries = 0;
  for (size_t ti = 0; ti < ms.trees.size(); ti++) {
    TreeSchema& t = ms.trees[ti];
    const uint64_t group = ms.has_tree_groups ? ms.tree_groups[ti] : 0;
    if (group >= K) bad("Forest schema tree group out of range");
    const size_t n = (size_t)t.num_nodes;
    if (t.split.size() != n || t.thr.size() != n || t.left.size() != n || t.right.size() != n ||
        t.dleft.size() != n)
      bad("Tree schema arrays must all have length num_nodes");
    if (t.has_gains && t.gains.size() != n) bad("Tree schema gains must have length num_nodes");
    if (t.has_covers && t.covers.size() != n) bad("Tree schema covers must have length num_nodes");
    if (t.cat_nodes.size() != t.cat_sets.size())
      bad("Tree schema categories.node_indices and categories.category_sets must have same length");
    for (uint32_t node : t.cat_nodes)
      if (node >= n) bad("Tree schema categories contains out-of-bounds node index");
    if (t.lin_nodes.size() != t.lin_len.size())
      bad("Tree schema linear_coefficients.node_indices and linear_coefficients.coefficients must "
          "have same length");
    for (size_t j = 0; j < t.lin_nodes.size(); j++) {
      if (t.lin_nodes[j] >= n) bad("Tree schema linear_coefficients contains out-of-bounds node index");
      if (t.lin_len[j] > 1 && (t.lin_len[j] - 1) % 2 != 0)
        bad("Tree schema linear_coefficients packed array must have length 1 + 2*k");
    }
    if (!t.lin_nodes.empty())
      raise(BST_E_FORMAT, "linear-in-leaves trees are not supported by this engine");
    if (t.leaf_kind == 1) {
      if (t.leaf.size() != n) bad("Tree schema leaf_values must have length num_nodes");
      append_tree(*m, t, t.leaf, (uint32_t)group, cat_entries);
    } else {
      if (t.vleaf.size() != n) bad("Tree schema leaf_values must have length num_nodes");
      const size_t k = t.vleaf.empty() ? 0 : t.vleaf[0].size();
      for (auto& v : t.vleaf)
        if (k == 0 || v.size() != k) bad("VectorLeaf values must have consistent non-zero length");
      std::vector<double> col(n);
      if (k == 1) {
        for (size_t i = 0; i < n; i++) col[i] = t.vleaf[i][0];
        append_tree(*m, t, col, (uint32_t)group, cat_entries);
      } else if (k == K) {
        for (uint32_t g = 0; g < K; g++) {
          for (size_t i = 0; i < n; i++) col[i] = t.vleaf[i][g];
          append_tree(*m, t, col, g, cat_entries);
        }
      } else {
        bad("VectorLeaf length must be 1 or equal to forest n_groups");
      }
    }
  }
  bst_model_info& info = m->info;
  info.n_trees = (int64_t)m->tree_group.size();
  info.n_groups = K;
  info.n_features = (int64_t)ms.num_features;
  info.total_nodes = (int64_t)m->split_index.size();
  info.total_cat_words = (int64_t)m->cat_words.size();
  info.output_transform = ms.transform;
  info.objective = -1;  // the file names no objective the engine can train
  if (ms.num_features > (uint64_t)INT64_MAX) bad("num_features out of range");
  m->meta_num_groups = ms.num_groups == K ? -1 : (int64_t)ms.num_groups;
  m->has_feature_names = ms.has_names;
  m->feature_names = std::move(ms.names);
  m->has_feature_types = ms.has_types;
  m->feature_types = std::move(ms.types);
  m->has_objective_name = ms.has_objective;
  m->objective_name = std::move(ms.objective);
  check_forest(*m);
  return m;
}

void put_le(std::vector<uint8_t>& b, uint64_t v, int bytes) {
  for (int i = 0; i < bytes; i++) b.push_back((uint8_t)(v >> (8 * i)));
}
uint64_t get_le(const uint8_t* p, int bytes) {
  uint64_t v = 0;
  for (int i = bytes - 1; i >= 0; i--) v = (v << 8) | p[i];
  return v;
}

std::vector<uint8_t> encode_bstr(const bst_model& m, int level) {
  if (level < 0 || level > 22)
    raise(BST_E_CONFIG, "compression level must be 0 (uncompressed) or a zstd level 1..22");
  std::vector<uint8_t> payload;
  MsgpackWriter w{payload};
  write_schema(w, m);
  if (level > 0) payload = zstd_compress(payload, level);
  std::vector<uint8_t> out;
  out.reserve(kHeaderBytes + payload.size() + kTrailerBytes);
  out.insert(out.end(), {'B', 'S', 'T', 'R'});
  put_le(out, kSchemaVersion, 4);
  out.push_back(level > 0 ? kFormatMsgpackZstd : kFormatMsgpack);
  out.push_back(kModelGbdt);
  out.resize(kHeaderBytes, 0);
  out.insert(out.end(), payload.begin(), payload.end());
  put_le(out, payload.size(), 8);
  put_le(out, crc32c(payload.data(), payload.size()), 4);
  put_le(out, 0, 4);
  return out;
}

std::unique_ptr<bst_model> decode_bstr(const uint8_t* buf, size_t size) {
  // BstrHeader::read_from (envelope.rs:133-170), read_binary_schema (binary.rs:101-146).
  if (size < kHeaderBytes) bad("file too small for a .bstr header");
  if (std::memcmp(buf, "BSTR", 4) != 0) bad("invalid magic bytes: not a .bstr file");
  const uint32_t version = (uint32_t)get_le(buf + 4, 4);
  if (version > kSchemaVersion)
    bad("unsupported schema version " + std::to_string(version) + " (max supported " +
        std::to_string(kSchemaVersion) + ")");
  const uint8_t format = buf[8], type = buf[9];
  if (format > kFormatMsgpackZstd) bad("unknown format byte " + std::to_string(format));
  if (type < 1 || type > 3) bad("unsupported model type " + std::to_string(type));
  if (format == kFormatJson) bad("binary header format 0x00 (Json) is not valid for .bstr");
  if (type != kModelGbdt) bad("unsupported model type " + std::to_string(type) + " (expected gbdt)");
  if (size < kHeaderBytes + kTrailerBytes) bad("file too small");
  const uint8_t* payload = buf + kHeaderBytes;
  const size_t plen = size - kHeaderBytes - kTrailerBytes;
  const uint8_t* tr = buf + size - kTrailerBytes;
  const uint64_t want_len = get_le(tr, 8);
  const uint32_t want_crc = (uint32_t)get_le(tr + 8, 4);
  if (plen == 0 && want_len != 0) bad("missing payload");
  if (want_len != plen)
    bad("payload length mismatch: trailer says " + std::to_string(want_len) + ", actual " +
        std::to_string(plen));
  const uint32_t crc = crc32c(payload, plen);
  if (crc != want_crc) bad("checksum mismatch");
  std::vector<uint8_t> raw;
  const uint8_t* mp = payload;
  size_t mlen = plen;
  if (format == kFormatMsgpackZstd) {
    raw = zstd_decompress(payload, plen);
    mp = raw.data();
    mlen = raw.size();
  }
  MsgpackReader r{mp, mp + mlen, {}, 0};
  ModelSchema ms;
  read_schema(r, ms);
  r.finish();
  return model_from_schema(ms);
}

std::string encode_json(const bst_model& m, bool pretty) {
  std::string s;
  JsonWriter w{s, pretty, {}, false};
  w.map(3);
  w.key("bstr_version");
  w.u64(kSchemaVersion);
  w.key("model_type");
  w.str("gbdt");
  w.key("model");
  write_schema(w, m);
  w.end_map();
  return s;
}

std::unique_ptr<bst_model> decode_json(const char* buf, size_t size) {
  // read_json_from (persist/mod.rs:126-143): envelope, version, model type.
  JsonReader r{buf, buf + size, {}, 0};
  std::string k, type;
  uint64_t version = 0;
  Seen seen;
  ModelSchema ms;
  r.begin_map();
  while (r.next_key(k)) {
    seen.mark(k);
    if (k == "bstr_version") {
      version = r.u64();
    } else if (k == "model_type") {
      type = r.str();
    } else if (k == "model") {
      // Parsed as a GBDT schema in the same pass; the envelope's version and
      // model type are checked once the envelope is complete.
      read_schema(r, ms);
    } else {
      r.skip();
    }
  }
  r.finish();
  for (const char* f : {"bstr_version", "model_type", "model"})
    if (!seen.has(f)) missing(f);
  if (version > kSchemaVersion)
    bad("unsupported schema version " + std::to_string(version) + " (max supported " +
        std::to_string(kSchemaVersion) + ")");
  if (type != "gbdt") bad("unknown model type `" + type + "`");
  return model_from_schema(ms);
}

}  // namespace

namespace bst {

void check_forest(const bst_model& m) {
  const bst_model_info& I = m.info;
  const int64_t T = I.n_trees, N = I.total_nodes, W = I.total_cat_words, K = I.n_groups;
  if (K < 1 || K > kMaxOutputs) bad("n_groups must be in 1.." + std::to_string(kMaxOutputs));
  if (T < 0 || N < 0 || W < 0 || I.n_features < 0) bad("negative model sizes");
  if ((int64_t)m.base_score.size() != K) bad("base_score must have n_groups entries");
  if ((int64_t)m.tree_offset.size() != T + 1 || (int64_t)m.tree_group.size() != T)
    bad("tree offsets / groups must have n_trees (+1) entries");
  for (auto* v : {&m.split_index, &m.left, &m.right, &m.cat_start, &m.cat_size})
    if ((int64_t)v->size() != N) bad("node arrays must have total_nodes entries");
  for (auto* v : {&m.split_threshold, &m.leaf_value, &m.gain, &m.cover})
    if ((int64_t)v->size() != N) bad("node arrays must have total_nodes entries");
  for (auto* v : {&m.default_left, &m.is_leaf, &m.split_type})
    if ((int64_t)v->size() != N) bad("node arrays must have total_nodes entries");
  if ((int64_t)m.cat_words.size() != W) bad("cat_words must have total_cat_words entries");
  if (m.tree_offset[0] != 0 || m.tree_offset[T] != N) bad("bad tree offsets");
  std::vector<uint8_t> seen;
  std::vector<uint32_t> stack;
  for (int64_t t = 0; t < T; t++) {
    const int64_t b = m.tree_offset[t], e = m.tree_offset[t + 1];
    if (e <= b) bad("tree " + std::to_string(t) + " is empty or has bad offsets");
    if ((int64_t)m.tree_group[t] >= K) bad("tree group out of range");
    const uint64_t n = (uint64_t)(e - b);
    seen.assign(n, 0);
    stack.assign(1, 0u);
    // Every node reachable from the root is visited exactly once: a second
    // visit is a cycle or a shared subtree, both rejected.
    while (!stack.empty()) {
      const uint32_t v = stack.back();
      stack.pop_back();
      if (seen[v]) bad("tree " + std::to_string(t) + ": node " + std::to_string(v) + " reached twice");
      seen[v] = 1;
      const int64_t g = b + v;
      if (m.is_leaf[g]) continue;
      const uint32_t l = m.left[g], r = m.right[g];
      if (l == v || r == v) bad("tree " + std::to_string(t) + ": self loop at node " + std::to_string(v));
      if (l >= n || r >= n) bad("tree " + std::to_string(t) + ": child id out of range");
      if ((int64_t)m.split_index[g] >= I.n_features || m.split_index[g] >= (1u << 24))
        bad("tree " + std::to_string(t) + ": split feature out of range");
      if (m.split_type[g] && (uint64_t)m.cat_start[g] + m.cat_size[g] > (uint64_t)W)
        bad("tree " + std::to_string(t) + ": category bitset out of range");
      stack.push_back(r);
      stack.push_back(l);
    }
    for (int64_t g = b; g < e; g++)
      if (m.cat_size[g] && (uint64_t)m.cat_start[g] + m.cat_size[g] > (uint64_t)W)
        bad("tree " + std::to_string(t) + ": category bitset out of range");
  }
}

}  // namespace bst

extern "C" {

int bst_model_save_bstr(const bst_model* model, int32_t compression_level, uint8_t** out,
                        int64_t* size) {
  return guard([&] {
    need(model && out && size, BST_E_INVALID, "null argument");
    std::vector<uint8_t> bytes = encode_bstr(*model, compression_level);
    auto* p = (uint8_t*)std::malloc(std::max<size_t>(bytes.size(), 1));
    if (!p) raise(BST_E_OOM, "host allocation failed");
    std::memcpy(p, bytes.data(), bytes.size());
    *out = p;
    *size = (int64_t)bytes.size();
    return BST_OK;
  });
}

int bst_model_load_bstr(bst_ctx* ctx, const uint8_t* buf, int64_t size, bst_model** out) {
  return guard([&] {
    need(out && (buf || size == 0) && size >= 0, BST_E_INVALID, "null argument");
    (void)ctx;
    *out = decode_bstr(buf, (size_t)size).release();
    return BST_OK;
  });
}

int bst_model_save_bstr_file(const bst_model* model, const char* path, int32_t compression_level) {
  return guard([&] {
    need(model && path, BST_E_INVALID, "null argument");
    const std::vector<uint8_t> bytes = encode_bstr(*model, compression_level);
    FILE* f = std::fopen(path, "wb");
    if (!f) raise(BST_E_INVALID, std::string("cannot open ") + path + " for writing");
    const size_t wrote = std::fwrite(bytes.data(), 1, bytes.size(), f);
    const int closed = std::fclose(f);
    if (wrote != bytes.size() || closed != 0) raise(BST_E_INVALID, std::string("write to ") + path + " failed");
    return BST_OK;
  });
}

int bst_model_load_bstr_file(bst_ctx* ctx, const char* path, bst_model** out) {
  return guard([&] {
    need(path && out, BST_E_INVALID, "null argument");
    (void)ctx;
    FILE* f = std::fopen(path, "rb");
    if (!f) raise(BST_E_INVALID, std::string("cannot open ") + path);
    std::vector<uint8_t> bytes;
    uint8_t chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof chunk, f)) > 0) bytes.insert(bytes.end(), chunk, chunk + got);
    const bool err = std::ferror(f) != 0;
    std::fclose(f);
    if (err) raise(BST_E_INVALID, std::string("read of ") + path + " failed");
    *out = decode_bstr(bytes.data(), bytes.size()).release();
    return BST_OK;
  });
}

int bst_model_save_json(const bst_model* model, int32_t pretty, char** out, int64_t* size) {
  return guard([&] {
    need(model && out && size, BST_E_INVALID, "null argument");
    const std::string s = encode_json(*model, pretty != 0);
    auto* p = (char*)std::malloc(s.size() + 1);
    if (!p) raise(BST_E_OOM, "host allocation failed");
    std::memcpy(p, s.c_str(), s.size() + 1);
    *out = p;
    *size = (int64_t)s.size();
    return BST_OK;
  });
}

int bst_model_load_json(bst_ctx* ctx, const char* buf, int64_t size, bst_model** out) {
  return guard([&] {
    need(out && (buf || size == 0) && size >= 0, BST_E_INVALID, "null argument");
    (void)ctx;
    *out = decode_json(buf, (size_t)size).release();
    return BST_OK;
  });
}

void bst_buffer_free(void* p) { std::free(p); }

}  // extern "C"
