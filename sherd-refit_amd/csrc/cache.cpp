// The fragment cache `<out>/cache/<name>.sherd` (crates/sherd-core/src/fragment/cache.rs) read on
// the host, and MatchData::own (crates/sherd-core/src/fragment/samples.rs:465-560) built from it:
// the arrays the pair stages read, ready to be uploaded into the device objects.
//
// The file is safetensors: an 8-byte little-endian header length, a JSON header naming every
// tensor's dtype, shape and byte range, then the tensor bytes. The reference's metadata travels as
// one JSON object (a string) under `__metadata__.sherd` (cache.rs:44-56). The reader checks what
// cache.rs's reader checks, with its messages: the format, version and algorithm fields
// (cache.rs:561-587), every tensor's dtype and shape, that triangles index V, that labels describe
// F, that the breakline frames describe brk_P and brk_sub indexes it, that sp/fp describe S/Pf
// and index F, and that margin_idx indexes S (cache.rs:329-438).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/sherd_hip.h"

namespace sherd {
int fail(int code, const std::string &msg);
}

namespace {

constexpr uint32_t CACHE_VERSION = 7;             // crates/sherd-core/src/lib.rs:152
constexpr const char *ALGO_REF = "2026-09-06/9d4b9d3";  // crates/sherd-core/src/lib.rs:130
constexpr const char *FORMAT = "sherd-cache";      // cache.rs:76

// ---- a small JSON reader (objects, arrays, strings, numbers, literals) ------------------------

struct Json {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  bool b = false;
  double num = 0.0;
  std::string raw;  // a number's text (for exact integers), or a string's value
  std::vector<Json> arr;
  std::vector<std::pair<std::string, Json>> obj;
  const Json *get(const std::string &k) const {
    for (auto &kv : obj)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

struct Parser {
  const char *p, *end;
  std::string err;
  void ws() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
  }
  bool lit(const char *s) {
    size_t n = std::strlen(s);
    if ((size_t)(end - p) < n || std::strncmp(p, s, n) != 0) return false;
    p += n;
    return true;
  }
  static void utf8(uint32_t c, std::string &o) {
    if (c < 0x80) {
      o += (char)c;
    } else if (c < 0x800) {
      o += (char)(0xC0 | (c >> 6));
      o += (char)(0x80 | (c & 0x3F));
    } else if (c < 0x10000) {
      o += (char)(0xE0 | (c >> 12));
      o += (char)(0x80 | ((c >> 6) & 0x3F));
      o += (char)(0x80 | (c & 0x3F));
    } else {
      o += (char)(0xF0 | (c >> 18));
      o += (char)(0x80 | ((c >> 12) & 0x3F));
      o += (char)(0x80 | ((c >> 6) & 0x3F));
      o += (char)(0x80 | (c & 0x3F));
    }
  }
  bool hex4(uint32_t *v) {
    if (end - p < 4) return false;
    *v = 0;
    for (int k = 0; k < 4; ++k) {
      char c = *p++;
      *v <<= 4;
      if (c >= '0' && c <= '9') *v |= (uint32_t)(c - '0');
      else if (c >= 'a' && c <= 'f') *v |= (uint32_t)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') *v |= (uint32_t)(c - 'A' + 10);
      else return false;
    }
    return true;
  }
  bool str(std::string &o) {
    if (p >= end || *p != '"') return false;
    ++p;
    while (p < end && *p != '"') {
      char c = *p++;
      if (c != '\\') {
        o += c;
        continue;
      }
      if (p >= end) return false;
      char e = *p++;
      switch (e) {
        case '"': o += '"'; break;
        case '\\': o += '\\'; break;
        case '/': o += '/'; break;
        case 'b': o += '\b'; break;
        case 'f': o += '\f'; break;
        case 'n': o += '\n'; break;
        case 'r': o += '\r'; break;
        case 't': o += '\t'; break;
        case 'u': {
          uint32_t v;
          if (!hex4(&v)) return false;
          if (v >= 0xD800 && v < 0xDC00) {
            uint32_t lo;
            if (!lit("\\u") || !hex4(&lo)) return false;
            v = 0x10000 + ((v - 0xD800) << 10) + (lo - 0xDC00);
          }
          utf8(v, o);
          break;
        }
        default: return false;
      }
    }
    if (p >= end) return false;
    ++p;
    return true;
  }
  bool value(Json &v, int depth) {
    if (depth > 64) return false;
    ws();
    if (p >= end) return false;
    char c = *p;
    if (c == '{') {
      ++p;
      v.kind = Json::OBJ;
      ws();
      if (p < end && *p == '}') {
        ++p;
        return true;
      }
      for (;;) {
        ws();
        std::string k;
        if (!str(k)) return false;
        ws();
        if (p >= end || *p != ':') return false;
        ++p;
        Json x;
        if (!value(x, depth + 1)) return false;
        v.obj.emplace_back(std::move(k), std::move(x));
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == '}') {
          ++p;
          return true;
        }
        return false;
      }
    }
    if (c == '[') {
      ++p;
      v.kind = Json::ARR;
      ws();
      if (p < end && *p == ']') {
        ++p;
        return true;
      }
      for (;;) {
        Json x;
        if (!value(x, depth + 1)) return false;
        v.arr.push_back(std::move(x));
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == ']') {
          ++p;
          return true;
        }
        return false;
      }
    }
    if (c == '"') {
      v.kind = Json::STR;
      return str(v.raw);
    }
    if (lit("true")) {
      v.kind = Json::BOOL;
      v.b = true;
      return true;
    }
    if (lit("false")) {
      v.kind = Json::BOOL;
      return true;
    }
    if (lit("null")) return true;
    const char *s = p;
    if (p < end && (*p == '-' || *p == '+')) ++p;
    while (p < end && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' || *p == '-' || *p == '+')) ++p;
    if (p == s) return false;
    v.kind = Json::NUM;
    v.raw.assign(s, p);
    char *stop = nullptr;
    v.num = std::strtod(v.raw.c_str(), &stop);  // correctly rounded: a shortest round-trip text reads back exactly
    return stop && *stop == '\0';
  }
};

bool parse_json(const char *s, size_t n, Json &out) {
  Parser ps{s, s + n, {}};
  if (!ps.value(out, 0)) return false;
  ps.ws();
  return ps.p == ps.end;
}

// ---- the cache --------------------------------------------------------------------------------

struct Tensor {
  std::string dtype;
  std::vector<uint64_t> shape;
  const uint8_t *data = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct sherd_cache {
  std::vector<uint8_t> buf;
  std::map<std::string, Tensor> tensors;
  sherd_cache_info info{};
  // The stored arrays, decoded (little-endian by the safetensors specification).
  std::vector<float> V, brk_P, brk_ns, brk_nf, brk_f, S, Pf;
  std::vector<uint32_t> F, brk_sub, sp, fp, margin_idx;
  std::vector<uint8_t> labels, brk_valid;
};

namespace {

int bad(const std::string &m) { return sherd::fail(SHERD_E_FORMAT, "cache: " + m); }

std::string shape_text(const Tensor &t) {
  std::string s = t.dtype + " [";
  for (size_t k = 0; k < t.shape.size(); ++k) s += (k ? ", " : "") + std::to_string(t.shape[k]);
  return s + "]";
}

template <class T>
void decode(const Tensor &t, std::vector<T> &out) {
  out.resize(t.bytes / sizeof(T));
  if (!out.empty()) std::memcpy(out.data(), t.data, out.size() * sizeof(T));  // x86-64 and gfx hosts are little-endian
}

// One tensor of `dtype` with `cols` columns (0: one dimension); its rows in *rows.
int tensor(sherd_cache *c, const char *name, const char *dtype, size_t elem, int cols, size_t *rows) {
  auto it = c->tensors.find(name);
  if (it == c->tensors.end()) return bad(std::string("tensor ") + name + ": not in the file");
  const Tensor &t = it->second;
  bool ok = t.dtype == dtype && (cols ? (t.shape.size() == 2 && t.shape[1] == (uint64_t)cols) : t.shape.size() == 1);
  if (ok) {
    uint64_t n = t.shape[0] * (cols ? (uint64_t)cols : 1);
    ok = n * elem == t.bytes;
  }
  if (!ok)
    return bad(std::string("tensor ") + name + " is " + shape_text(t) + ", expected " + dtype +
               (cols ? " [n, " + std::to_string(cols) + "]" : std::string(" [n]")));
  *rows = (size_t)t.shape[0];
  return SHERD_OK;
}

double num(const Json *j, double dflt = 0.0) { return j && j->kind == Json::NUM ? j->num : dflt; }

int parse(sherd_cache *c) {
  const std::vector<uint8_t> &b = c->buf;
  if (b.size() < 8) return bad("shorter than a safetensors header");
  uint64_t n = 0;
  for (int k = 7; k >= 0; --k) n = (n << 8) | b[(size_t)k];
  if (n > b.size() - 8) return bad("the header length " + std::to_string(n) + " runs past the file");
  Json head;
  if (!parse_json((const char *)b.data() + 8, (size_t)n, head) || head.kind != Json::OBJ)
    return bad("reading the header: not a JSON object");
  const uint8_t *data = b.data() + 8 + n;
  size_t data_len = b.size() - 8 - (size_t)n;
  // Every tensor's byte range lies in the data section, and together they cover it exactly
  // (safetensors refuses a buffer that is not header plus data).
  std::vector<std::pair<uint64_t, uint64_t>> ranges;
  const Json *meta = nullptr;
  for (auto &kv : head.obj) {
    if (kv.first == "__metadata__") {
      meta = &kv.second;
      continue;
    }
    const Json &d = kv.second;
    const Json *dt = d.get("dtype"), *sh = d.get("shape"), *off = d.get("data_offsets");
    if (!dt || dt->kind != Json::STR || !sh || sh->kind != Json::ARR || !off || off->kind != Json::ARR ||
        off->arr.size() != 2)
      return bad("tensor " + kv.first + ": a malformed header entry");
    Tensor t;
    t.dtype = dt->raw;
    for (auto &s : sh->arr) {
      if (s.kind != Json::NUM || s.num < 0 || s.num != std::floor(s.num)) return bad("tensor " + kv.first + ": a bad shape");
      t.shape.push_back((uint64_t)s.num);
    }
    double a = num(&off->arr[0], -1), e = num(&off->arr[1], -1);
    if (!(a >= 0 && e >= a && e <= (double)data_len)) return bad("tensor " + kv.first + ": its bytes lie outside the file");
    t.data = data + (size_t)a;
    t.bytes = (size_t)(e - a);
    ranges.emplace_back((uint64_t)a, (uint64_t)e);
    c->tensors[kv.first] = t;
  }
  std::sort(ranges.begin(), ranges.end());
  uint64_t at = 0;
  for (auto &r : ranges) {
    if (r.first != at) return bad("the tensors' byte ranges overlap or leave a gap");
    at = r.second;
  }
  if (at != data_len) return bad("the tensors do not cover the data section");
  // The metadata block: one JSON object as a string under `sherd` (cache.rs:44-56, 561-587).
  const Json *js = meta ? meta->get("sherd") : nullptr;
  if (!js || js->kind != Json::STR) return bad("no `sherd` metadata key");
  Json m;
  if (!parse_json(js->raw.data(), js->raw.size(), m) || m.kind != Json::OBJ) return bad("parsing the metadata");
  const Json *fmt = m.get("format"), *ver = m.get("cache_version"), *algo = m.get("algo_ref");
  if (!fmt || fmt->kind != Json::STR || fmt->raw != FORMAT)
    return bad("format `" + std::string(fmt && fmt->kind == Json::STR ? fmt->raw : "") + "`, expected `" + FORMAT + "`");
  if (!ver || ver->kind != Json::NUM || ver->num != (double)CACHE_VERSION)
    return bad("cache_version " + (ver ? ver->raw : std::string("missing")) + ", expected " + std::to_string(CACHE_VERSION));
  if (!algo || algo->kind != Json::STR || algo->raw != ALGO_REF)
    return bad("algo_ref `" + std::string(algo && algo->kind == Json::STR ? algo->raw : "") + "`, expected `" + ALGO_REF + "`");
  sherd_cache_info &I = c->info;
  const Json *name = m.get("name");
  std::snprintf(I.name, sizeof I.name, "%s", name && name->kind == Json::STR ? name->raw.c_str() : "");
  I.cache_version = CACHE_VERSION;
  I.target_faces = (uint32_t)num(m.get("target_faces"));
  I.face_budget = (uint32_t)num(m.get("face_budget"));
  I.n_boundary = (uint32_t)num(m.get("n_boundary"));
  const Json *wt = m.get("watertight");
  I.watertight = wt && wt->kind == Json::BOOL && wt->b ? 1 : 0;
  I.thick = num(m.get("thick"), NAN);
  I.thick_mode = num(m.get("thick_mode"), NAN);
  I.res = num(m.get("res"), NAN);
  I.area = num(m.get("area"), NAN);
  I.area0 = num(m.get("area0"), NAN);
  I.frac_area = num(m.get("frac_area"), NAN);
  const Json *bp = m.get("brk_params"), *mp = m.get("md_params");
  if (!bp || bp->kind != Json::OBJ) return bad("no brk_params");
  if (!mp || mp->kind != Json::OBJ) return bad("no md_params");
  I.brk_t = num(bp->get("t"), NAN);
  I.md_t = num(mp->get("t"), NAN);
  // The tensors (cache.rs:329-438), in the reader's order and with its checks.
  size_t nv, nf, k;
  int rc;
  if ((rc = tensor(c, "V", "F32", 4, 3, &nv)) || (rc = tensor(c, "F", "U32", 4, 3, &nf))) return rc;
  decode(c->tensors["V"], c->V);
  decode(c->tensors["F"], c->F);
  for (uint32_t i : c->F)
    if (i >= nv) return bad("triangle index " + std::to_string(i) + " is outside V (" + std::to_string(nv) + ")");
  {
    auto it = c->tensors.find("labels");
    size_t nl;
    if (it == c->tensors.end() || it->second.dtype != "U8" || tensor(c, "labels", "U8", 1, 0, &nl) || nl != nf)
      return bad("tensor labels is " + (it == c->tensors.end() ? std::string("missing") : shape_text(it->second)) +
                 ", expected U8 [" + std::to_string(nf) + "]");
    decode(it->second, c->labels);
    for (uint8_t l : c->labels)
      if (l > 2) return bad("label " + std::to_string(l));
  }
  size_t kns, knf, kf, ksub, kval;
  if ((rc = tensor(c, "brk_P", "F32", 4, 3, &k)) || (rc = tensor(c, "brk_ns", "F32", 4, 3, &kns)) ||
      (rc = tensor(c, "brk_nf", "F32", 4, 3, &knf)) || (rc = tensor(c, "brk_f", "F32", 4, 3, &kf)) ||
      (rc = tensor(c, "brk_sub", "U32", 4, 0, &ksub)) || (rc = tensor(c, "brk_valid", "BOOL", 1, 0, &kval)))
    return rc;
  if (kns != k || knf != k || kf != k) return bad("the brk_* frames do not describe brk_P");
  if (kval != k) return bad("brk_valid does not describe brk_P");
  decode(c->tensors["brk_P"], c->brk_P);
  decode(c->tensors["brk_ns"], c->brk_ns);
  decode(c->tensors["brk_nf"], c->brk_nf);
  decode(c->tensors["brk_f"], c->brk_f);
  decode(c->tensors["brk_sub"], c->brk_sub);
  decode(c->tensors["brk_valid"], c->brk_valid);
  for (uint32_t i : c->brk_sub)
    if (i >= k) return bad("brk_sub " + std::to_string(i) + " is outside brk_P (" + std::to_string(k) + ")");
  size_t ns, nsp, npf, nfp, nm;
  if ((rc = tensor(c, "S", "F32", 4, 3, &ns)) || (rc = tensor(c, "sp", "U32", 4, 0, &nsp)) ||
      (rc = tensor(c, "Pf", "F32", 4, 3, &npf)) || (rc = tensor(c, "fp", "U32", 4, 0, &nfp)) ||
      (rc = tensor(c, "margin_idx", "U32", 4, 0, &nm)))
    return rc;
  if (nsp != ns || nfp != npf) return bad("sp/fp do not describe S/Pf");
  decode(c->tensors["S"], c->S);
  decode(c->tensors["sp"], c->sp);
  decode(c->tensors["Pf"], c->Pf);
  decode(c->tensors["fp"], c->fp);
  decode(c->tensors["margin_idx"], c->margin_idx);
  for (const auto *v : {&c->sp, &c->fp})
    for (uint32_t i : *v)
      if (i >= nf) return bad("sample face " + std::to_string(i) + " is outside F (" + std::to_string(nf) + ")");
  for (uint32_t i : c->margin_idx)
    if (i >= ns) return bad("margin_idx " + std::to_string(i) + " is outside S (" + std::to_string(ns) + ")");
  I.nv = nv;
  I.nf = nf;
  I.n_brk = k;
  I.n_sub = ksub;
  I.n_s = ns;
  I.n_pf = npf;
  I.n_margin = nm;
  return SHERD_OK;
}

// samples.rs:601-615: the split of reg_points between the fracture samples and the margin
// (numpy's half-to-even rounding; the fracture side keeps at least one point).
void registration_split(size_t n_frac, size_t n_margin, size_t reg_points, size_t *take_f, size_t *take_m) {
  if (reg_points == 0 || n_frac + n_margin <= reg_points) {
    *take_f = n_frac;
    *take_m = n_margin;
    return;
  }
  double share = (double)n_frac * (double)reg_points / (double)(n_frac + n_margin);
  size_t r = (size_t)std::nearbyint(share);  // the default rounding mode: ties to even
  size_t tf = r < 1 ? 1 : r;
  if (tf > n_frac) tf = n_frac;
  size_t tm = reg_points > tf ? reg_points - tf : 0;
  if (tm > n_margin) tm = n_margin;
  *take_f = tf;
  *take_m = tm;
}

}  // namespace

extern "C" {

int sherd_cache_from_bytes(const uint8_t *bytes, size_t n, sherd_cache **out) {
  if (!out || (!bytes && n)) return sherd::fail(SHERD_E_ARG, "null argument");
  *out = nullptr;
  std::unique_ptr<sherd_cache> c(new sherd_cache());
  c->buf.assign(bytes, bytes + n);
  int rc = parse(c.get());
  if (rc) return rc;
  *out = c.release();
  return SHERD_OK;
}

int sherd_cache_read(const char *path, sherd_cache **out) {
  if (!path || !out) return sherd::fail(SHERD_E_ARG, "null argument");
  *out = nullptr;
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) return bad(std::string(path) + ": cannot be opened");
  std::streamsize n = f.tellg();
  f.seekg(0);
  std::unique_ptr<sherd_cache> c(new sherd_cache());
  c->buf.resize((size_t)n);
  if (n > 0 && !f.read((char *)c->buf.data(), n)) return bad(std::string(path) + ": cannot be read");
  int rc = parse(c.get());
  if (rc) return rc;
  *out = c.release();
  return SHERD_OK;
}

void sherd_cache_free(sherd_cache *c) { delete c; }

int sherd_cache_info_get(const sherd_cache *c, sherd_cache_info *out) {
  if (!c || !out) return sherd::fail(SHERD_E_ARG, "null argument");
  *out = c->info;
  return SHERD_OK;
}

int sherd_cache_arrays(const sherd_cache *c, const sherd_cache_arrays_out *o) {
  if (!c || !o) return sherd::fail(SHERD_E_ARG, "null argument");
  auto put = [](void *dst, const void *src, size_t bytes) {
    if (dst && bytes) std::memcpy(dst, src, bytes);
  };
  put(o->V, c->V.data(), c->V.size() * 4);
  put(o->F, c->F.data(), c->F.size() * 4);
  put(o->labels, c->labels.data(), c->labels.size());
  put(o->brk_P, c->brk_P.data(), c->brk_P.size() * 4);
  put(o->brk_ns, c->brk_ns.data(), c->brk_ns.size() * 4);
  put(o->brk_nf, c->brk_nf.data(), c->brk_nf.size() * 4);
  put(o->brk_f, c->brk_f.data(), c->brk_f.size() * 4);
  put(o->brk_sub, c->brk_sub.data(), c->brk_sub.size() * 4);
  put(o->brk_valid, c->brk_valid.data(), c->brk_valid.size());
  put(o->S, c->S.data(), c->S.size() * 4);
  put(o->sp, c->sp.data(), c->sp.size() * 4);
  put(o->Pf, c->Pf.data(), c->Pf.size() * 4);
  put(o->fp, c->fp.data(), c->fp.size() * 4);
  put(o->margin_idx, c->margin_idx.data(), c->margin_idx.size() * 4);
  return SHERD_OK;
}

int sherd_match_data(const sherd_cache *c, size_t reg_points, sherd_match_info *info, const sherd_match_arrays *o) {
  if (!c || !info) return sherd::fail(SHERD_E_ARG, "null argument");
  const sherd_cache_info &I = c->info;
  // MatchData::own is the cached arrays at the fragment's own t; arrays built at another t are
  // rebuilt by the reference (samples.rs:562-575), which is preprocessing and not this library's.
  if (!(I.brk_t == I.thick && I.md_t == I.thick))
    return sherd::fail(SHERD_E_LIMIT, "the cached arrays were built at another t than the fragment's own");
  size_t tf, tm;
  registration_split(I.n_pf, I.n_margin, reg_points, &tf, &tm);
  std::memset(info, 0, sizeof *info);
  info->n_reg = tf + tm;
  info->n_reg_frac = tf;
  info->n_reg_margin = tm;
  info->n_pf = I.n_pf;
  info->n_brk = I.n_brk;
  info->n_sub = I.n_sub;
  info->n_margin = I.n_margin;
  info->n_s = I.n_s;
  info->nf = I.nf;
  info->t = I.thick;
  info->frac_area = I.frac_area;
  info->watertight = I.watertight;
  if (!o) return SHERD_OK;
  // WorkingMesh::from_parts (types.rs:100-112): face_geometry in f64 over the f32 vertices,
  // normals narrowed to f32 (geometry.rs:70-96, the length floored at 1e-12).
  std::vector<float> fn(3 * I.nf);
  for (size_t f = 0; f < I.nf; ++f) {
    const float *a = &c->V[3 * (size_t)c->F[3 * f]], *b = &c->V[3 * (size_t)c->F[3 * f + 1]],
                *cc = &c->V[3 * (size_t)c->F[3 * f + 2]];
    double e1[3], e2[3];
    for (int k = 0; k < 3; ++k) {
      e1[k] = (double)b[k] - (double)a[k];
      e2[k] = (double)cc[k] - (double)a[k];
    }
    double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    double len = std::sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
    double den = len > 1e-12 ? len : 1e-12;
    for (int k = 0; k < 3; ++k) fn[3 * f + k] = (float)(n[k] / den);
  }
  auto row = [](double *dst, size_t i, const float *src) {
    if (!dst) return;
    for (int k = 0; k < 3; ++k) dst[3 * i + k] = (double)src[k];
  };
  if (o->face_normals)
    for (size_t f = 0; f < I.nf; ++f) row(o->face_normals, f, &fn[3 * f]);
  // Surface samples with SN = FN[sp] and S_frac = labels[sp] == fracture.
  for (size_t i = 0; i < I.n_s; ++i) {
    row(o->s, i, &c->S[3 * i]);
    row(o->s_n, i, &fn[3 * (size_t)c->sp[i]]);
    if (o->s_frac) o->s_frac[i] = c->labels[c->sp[i]] == 1 ? 1 : 0;
  }
  // pc_frac = (Pf, FN[fp]); the margin = (S[margin_idx], SN[margin_idx]).
  for (size_t i = 0; i < I.n_pf; ++i) {
    row(o->pf, i, &c->Pf[3 * i]);
    row(o->pf_n, i, &fn[3 * (size_t)c->fp[i]]);
  }
  for (size_t i = 0; i < I.n_margin; ++i) {
    uint32_t j = c->margin_idx[i];
    row(o->margin_p, i, &c->S[3 * (size_t)j]);
    row(o->margin_n, i, &fn[3 * (size_t)c->sp[j]]);
  }
  // pc_reg: a prefix of each (samples.rs:585-593).
  for (size_t i = 0; i < tf; ++i) {
    row(o->reg_p, i, &c->Pf[3 * i]);
    row(o->reg_n, i, &fn[3 * (size_t)c->fp[i]]);
  }
  for (size_t i = 0; i < tm; ++i) {
    uint32_t j = c->margin_idx[i];
    row(o->reg_p, tf + i, &c->S[3 * (size_t)j]);
    row(o->reg_n, tf + i, &fn[3 * (size_t)c->sp[j]]);
  }
  // The breakline: points, shell normals, in-plane axes, the subset; the tangent ns x f in f64
  // (breakline.rs:180-182, unnarrowed as R §5.1 reads it) and the dihedral
  // acos(clamp(ns . nf)) in degrees (breakline.rs:189-195).
  for (size_t i = 0; i < I.n_brk; ++i) {
    row(o->brk_p, i, &c->brk_P[3 * i]);
    row(o->brk_ns, i, &c->brk_ns[3 * i]);
    row(o->brk_f, i, &c->brk_f[3 * i]);
    double ns[3], f[3], nf[3];
    for (int k = 0; k < 3; ++k) {
      ns[k] = c->brk_ns[3 * i + k];
      f[k] = c->brk_f[3 * i + k];
      nf[k] = c->brk_nf[3 * i + k];
    }
    if (o->brk_t) {
      o->brk_t[3 * i] = ns[1] * f[2] - ns[2] * f[1];
      o->brk_t[3 * i + 1] = ns[2] * f[0] - ns[0] * f[2];
      o->brk_t[3 * i + 2] = ns[0] * f[1] - ns[1] * f[0];
    }
    if (o->brk_dih) {
      double d = (ns[0] * nf[0] + ns[1] * nf[1]) + ns[2] * nf[2];
      d = d < -1.0 ? -1.0 : (d > 1.0 ? 1.0 : d);
      o->brk_dih[i] = std::acos(d) * (180.0 / M_PI);
    }
  }
  if (o->brk_sub) std::memcpy(o->brk_sub, c->brk_sub.data(), c->brk_sub.size() * 4);
  return SHERD_OK;
}

}  // extern "C"
