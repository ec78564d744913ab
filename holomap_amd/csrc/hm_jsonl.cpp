// holomap_gpu_jsonl -- the clusterer's JSON-lines front door
// (crates/holomap-clusterer/src/main.rs:10-31, protocol.rs:17-94,
// pipeline.rs:79-141) over the serving loop of libholomap_gpu.so.
//
// One request per stdin line, one response per stdout line, in request
// order, until EOF. A request is the clusterer's:
//   {"protocol_version": 1, "vectors": [[f32, ...], ...],
//    "params": {"n_components": u, "n_neighbors": u, "min_cluster_size": u, "seed": u}}
// and is checked as run_pipeline checks it, with its messages: the protocol
// version, empty input, inconsistent or zero dimensions; then the vectors are
// L2-normalised in f32 (pipeline.rs:151-162, on the device) and, when
// n_components > 0, handed to the serving loop as a cosine kNN-graph build
// with k = n_neighbors (the reduction's first stages, pipeline.rs:190-199),
// where fit_transform's own checks apply (its messages, prefixed
// "holomap reduction failed: " as pipeline.rs:113 does). Requests read while
// others run are admitted into concurrent device jobs (continuous batching);
// responses still leave in request order.
//
// The response carries the graph rather than cluster labels (the layout and
// HDBSCAN stages are not on this path):
//   {"protocol_version":1,"n":N,"n_neighbors":K,"knn_indices":[[...]],
//    "knn_dists":[[...]],"sigmas":[...],"rhos":[...],
//    "graph":{"rows":[...],"cols":[...],"vals":[...]}}
// (n_components == 0: no reduction runs, "n_neighbors":0 and empty arrays).
// Errors are in-band, as protocol.rs:67-93 prescribes:
//   {"protocol_version":1,"assignments":[],"error":"..."}
// Floats are written with 9 significant digits, which read back to the same
// f32. Numbers are read with strtod then rounded to f32.
#include <poll.h>
#include <unistd.h>

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/holomap_gpu.h"

namespace {

constexpr uint32_t kProtocolVersion = 1;  // protocol.rs:17

// ---- a small JSON reader: the value kinds a request holds ----
struct JVal {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  bool b = false;
  double num = 0.0;
  std::string numtext;  // the number as written (integer checks)
  std::string str;
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;
  const JVal* get(const char* key) const {
    for (auto& kv : obj)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
};

struct Parser {
  const char* p;
  const char* end;
  std::string err;
  int depth = 0;
  void ws() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
  }
  bool fail(const std::string& m) {
    if (err.empty()) err = m;
    return false;
  }
  bool value(JVal& v) {
    ws();
    if (p >= end) return fail("EOF while parsing a value");
    if (++depth > 64) return fail("recursion limit exceeded");
    bool ok;
    switch (*p) {
      case '{': ok = object(v); break;
      case '[': ok = array(v); break;
      case '"': v.kind = JVal::STR; ok = string(v.str); break;
      case 't': ok = lit("true", v, JVal::BOOL, true); break;
      case 'f': ok = lit("false", v, JVal::BOOL, false); break;
      case 'n': ok = lit("null", v, JVal::NUL, false); break;
      default: ok = number(v); break;
    }
    --depth;
    return ok;
  }
  bool lit(const char* w, JVal& v, JVal::Kind k, bool b) {
    const size_t n = strlen(w);
    if ((size_t)(end - p) < n || strncmp(p, w, n) != 0) return fail("expected value");
    p += n;
    v.kind = k;
    v.b = b;
    return true;
  }
  bool number(JVal& v) {
    const char* s = p;
    if (p < end && *p == '-') ++p;
    if (p >= end || !(*p >= '0' && *p <= '9')) return fail("expected value");
    while (p < end && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' ||
                       *p == '+' || *p == '-'))
      ++p;
    v.kind = JVal::NUM;
    v.numtext.assign(s, p);
    char* e = nullptr;
    errno = 0;
    v.num = strtod(v.numtext.c_str(), &e);
    if (e != v.numtext.c_str() + v.numtext.size()) return fail("invalid number");
    return true;
  }
  bool string(std::string& out) {
    ++p;  // opening quote
    while (p < end && *p != '"') {
      if (*p == '\\') {
        if (++p >= end) break;
        switch (*p) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u':
            if (end - p < 5) return fail("invalid escape");
            out += '?';  // (no string of a request is interpreted)
            p += 4;
            break;
          default: out += *p; break;
        }
        ++p;
      } else {
        out += *p++;
      }
    }
    if (p >= end) return fail("EOF while parsing a string");
    ++p;
    return true;
  }
  bool array(JVal& v) {
    v.kind = JVal::ARR;
    ++p;
    ws();
    if (p < end && *p == ']') { ++p; return true; }
    for (;;) {
      v.arr.emplace_back();
      if (!value(v.arr.back())) return false;
      ws();
      if (p < end && *p == ',') { ++p; continue; }
      if (p < end && *p == ']') { ++p; return true; }
      return fail("expected `,` or `]`");
    }
  }
  bool object(JVal& v) {
    v.kind = JVal::OBJ;
    ++p;
    ws();
    if (p < end && *p == '}') { ++p; return true; }
    for (;;) {
      ws();
      if (p >= end || *p != '"') return fail("key must be a string");
      std::string key;
      if (!string(key)) return false;
      ws();
      if (p >= end || *p != ':') return fail("expected `:`");
      ++p;
      v.obj.emplace_back(key, JVal());
      if (!value(v.obj.back().second)) return false;
      ws();
      if (p < end && *p == ',') { ++p; continue; }
      if (p < end && *p == '}') { ++p; return true; }
      return fail("expected `,` or `}`");
    }
  }
};

// An unsigned integer field (u32 / usize / u64 in protocol.rs).
bool get_uint(const JVal& o, const char* key, uint64_t maxv, uint64_t* out, std::string* err) {
  const JVal* v = o.get(key);
  if (!v) {
    *err = std::string("missing field `") + key + "`";
    return false;
  }
  if (v->kind != JVal::NUM || v->numtext.find_first_of(".eE-") != std::string::npos) {
    *err = std::string("invalid type for `") + key + "`, expected an unsigned integer";
    return false;
  }
  errno = 0;
  const unsigned long long u = strtoull(v->numtext.c_str(), nullptr, 10);
  if (errno || u > maxv) {
    *err = std::string("invalid value for `") + key + "`";
    return false;
  }
  *out = u;
  return true;
}

std::string json_escape(const std::string& s) {
  std::string o;
  for (char c : s) {
    if (c == '"' || c == '\\') { o += '\\'; o += c; }
    else if (c == '\n') o += "\\n";
    else if ((unsigned char)c < 0x20) o += ' ';
    else o += c;
  }
  return o;
}

std::string error_response(const std::string& msg) {
  return "{\"protocol_version\":" + std::to_string(kProtocolVersion) +
         ",\"assignments\":[],\"error\":\"" + json_escape(msg) + "\"}";
}

void put_f32(std::string& o, float v) {
  char b[32];
  snprintf(b, sizeof b, "%.9g", (double)v);
  o += b;
}

// The messages of HolomapError's Display (src/api.rs:46-63), for the checks
// fit_transform makes on what this path builds (src/api.rs:113-144).
std::string holomap_error(hm_status st, uint64_t n, uint32_t k, uint64_t bad) {
  std::string m;
  switch (st) {
    case HM_ETOOFEW:
      m = "need more than n_neighbors=" + std::to_string(k) + " points, got " + std::to_string(n);
      break;
    case HM_EINVAL: m = "invalid parameter: n_neighbors must be >= 2"; break;
    case HM_ENONFINITE: m = "data contains a non-finite value at index " + std::to_string(bad); break;
    case HM_EUNSUPPORTED:
      m = "n_neighbors > " + std::to_string(HM_MAX_NEIGHBORS) + " is not supported on the GPU path";
      break;
    case HM_EBUDGET: m = "the request exceeds the device memory budget"; break;
    default: m = "device failure (status " + std::to_string(st) + ")"; break;
  }
  return "holomap reduction failed: " + m;
}

struct Pending {
  bool ready = false;
  std::string response;
  uint64_t id = 0;  // server request id when !ready
  uint64_t n = 0;
  uint32_t k = 0;
};

struct Frontdoor {
  hm_ctx* ctx = nullptr;
  hm_server* srv = nullptr;
  std::deque<Pending> pending;

  bool ensure_device(std::string* err) {
    if (srv) return true;
    if (hm_ctx_create(&ctx, 0, 0) != HM_OK) {
      *err = "no usable GPU";
      return false;
    }
    hm_server_config cfg;
    hm_server_default_config(&cfg);
    if (hm_server_create(ctx, &cfg, &srv) != HM_OK) {
      *err = "server create failed";
      return false;
    }
    return true;
  }

  // run_pipeline's checks (pipeline.rs:79-115), then the graph request.
  void handle(const std::string& line) {
    Pending pe;
    pe.ready = true;
    Parser ps{line.data(), line.data() + line.size()};
    JVal root;
    if (!ps.value(root)) {
      pe.response = error_response("bad request: " + ps.err);
      pending.push_back(pe);
      return;
    }
    ps.ws();
    std::string err;
    uint64_t version = 0, ncomp = 0, nnb = 0, mcs = 0, seed = 0;
    const JVal* vecs = root.kind == JVal::OBJ ? root.get("vectors") : nullptr;
    const JVal* params = root.kind == JVal::OBJ ? root.get("params") : nullptr;
    if (ps.p != ps.end) err = "trailing characters";
    else if (root.kind != JVal::OBJ) err = "invalid type: expected struct Request";
    else if (!get_uint(root, "protocol_version", 0xffffffffull, &version, &err)) {}
    else if (!vecs) err = "missing field `vectors`";
    else if (vecs->kind != JVal::ARR) err = "invalid type for `vectors`, expected a sequence";
    else if (!params) err = "missing field `params`";
    else if (params->kind != JVal::OBJ) err = "invalid type for `params`, expected struct Params";
    else if (!get_uint(*params, "n_components", ~0ull, &ncomp, &err)) {}
    else if (!get_uint(*params, "n_neighbors", ~0ull, &nnb, &err)) {}
    else if (!get_uint(*params, "min_cluster_size", ~0ull, &mcs, &err)) {}
    else if (!get_uint(*params, "seed", ~0ull, &seed, &err)) {}
    if (err.empty()) {
      for (auto& row : vecs->arr) {
        if (row.kind != JVal::ARR) { err = "invalid type: expected a sequence of f32"; break; }
        for (auto& x : row.arr)
          if (x.kind != JVal::NUM) { err = "invalid type: expected f32"; break; }
        if (!err.empty()) break;
      }
    }
    if (!err.empty()) {
      pe.response = error_response("bad request: " + err);
      pending.push_back(pe);
      return;
    }
    if (version != kProtocolVersion) {
      pe.response = error_response("unsupported protocol_version " + std::to_string(version));
    } else if (vecs->arr.empty()) {
      pe.response = error_response("empty input");
    } else {
      const size_t dim = vecs->arr[0].arr.size();
      bool ragged = false;
      for (auto& row : vecs->arr) ragged |= row.arr.size() != dim;
      if (ragged) pe.response = error_response("vector dimensions inconsistent");
      else if (dim == 0) pe.response = error_response("zero-dimensional vectors");
      else if (ncomp == 0) {
        // no reduction runs (pipeline.rs:102-105): no graph
        pe.response = "{\"protocol_version\":1,\"n\":" + std::to_string(vecs->arr.size()) +
                      ",\"n_neighbors\":0,\"knn_indices\":[],\"knn_dists\":[],\"sigmas\":[],"
                      "\"rhos\":[],\"graph\":{\"rows\":[],\"cols\":[],\"vals\":[]}}";
      } else {
        submit(*vecs, dim, nnb, pe);
      }
    }
    pending.push_back(pe);
  }

  void submit(const JVal& vecs, size_t dim, uint64_t nnb, Pending& pe) {
    const uint64_t n = vecs.arr.size();
    std::vector<float> x(n * dim), xn(n * dim);
    for (uint64_t i = 0; i < n; ++i)
      for (size_t f = 0; f < dim; ++f) x[i * dim + f] = (float)vecs.arr[i].arr[f].num;
    std::string err;
    if (!ensure_device(&err)) {
      pe.response = error_response("holomap reduction failed: " + err);
      return;
    }
    if (hm_l2_normalise(ctx, x.data(), n, (uint32_t)dim, xn.data()) != HM_OK) {
      pe.response = error_response(std::string("holomap reduction failed: ") + hm_last_error(ctx));
      return;
    }
    // fit_transform's checks (src/api.rs:113-144) in its order
    const uint32_t k = nnb > 0xffffffffull ? 0xffffffffu : (uint32_t)nnb;
    uint64_t bad = 0;
    hm_status v = HM_OK;
    if (nnb < 2) v = HM_EINVAL;
    else if (n <= nnb) v = HM_ETOOFEW;
    else v = hm_validate(xn.data(), n * dim, (uint32_t)dim, k, &bad);
    if (v == HM_OK) {
      uint64_t id = 0;
      v = hm_server_submit(srv, xn.data(), n, (uint32_t)dim, k, HM_COSINE, &id);
      if (v == HM_OK) {
        pe.ready = false;
        pe.id = id;
        pe.n = n;
        pe.k = k;
        return;
      }
    }
    pe.response = error_response(holomap_error(v, n, k, bad));
  }

  // The graph of a finished request.
  std::string graph_response(const Pending& pe) {
    uint64_t nn = 0, nnz = 0;
    uint32_t kk = 0;
    hm_status st = hm_server_result_info(srv, pe.id, &nn, &kk, &nnz);
    if (st != HM_OK) return error_response(holomap_error(hm_server_fetch(srv, pe.id, nullptr), pe.n, pe.k, 0));
    std::vector<uint32_t> idx(nn * kk), rows(nnz ? nnz : 1), cols(nnz ? nnz : 1);
    std::vector<float> dist(nn * kk), sig(nn), rho(nn), vals(nnz ? nnz : 1);
    hm_graph_out o;
    memset(&o, 0, sizeof o);
    o.indices = idx.data();
    o.dists = dist.data();
    o.sigmas = sig.data();
    o.rhos = rho.data();
    o.rows = rows.data();
    o.cols = cols.data();
    o.vals = vals.data();
    o.cap = nnz ? nnz : 1;
    st = hm_server_fetch(srv, pe.id, &o);
    if (st != HM_OK) return error_response(holomap_error(st, pe.n, pe.k, 0));
    std::string r = "{\"protocol_version\":1,\"n\":" + std::to_string(nn) +
                    ",\"n_neighbors\":" + std::to_string(kk) + ",\"knn_indices\":[";
    for (uint64_t i = 0; i < nn; ++i) {
      r += i ? ",[" : "[";
      for (uint32_t t = 0; t < kk; ++t) {
        if (t) r += ',';
        r += std::to_string(idx[i * kk + t]);
      }
      r += ']';
    }
    r += "],\"knn_dists\":[";
    for (uint64_t i = 0; i < nn; ++i) {
      r += i ? ",[" : "[";
      for (uint32_t t = 0; t < kk; ++t) {
        if (t) r += ',';
        put_f32(r, dist[i * kk + t]);
      }
      r += ']';
    }
    r += "],\"sigmas\":[";
    for (uint64_t i = 0; i < nn; ++i) { if (i) r += ','; put_f32(r, sig[i]); }
    r += "],\"rhos\":[";
    for (uint64_t i = 0; i < nn; ++i) { if (i) r += ','; put_f32(r, rho[i]); }
    r += "],\"graph\":{\"rows\":[";
    for (uint64_t e = 0; e < o.nnz; ++e) { if (e) r += ','; r += std::to_string(rows[e]); }
    r += "],\"cols\":[";
    for (uint64_t e = 0; e < o.nnz; ++e) { if (e) r += ','; r += std::to_string(cols[e]); }
    r += "],\"vals\":[";
    for (uint64_t e = 0; e < o.nnz; ++e) { if (e) r += ','; put_f32(r, vals[e]); }
    r += "]}}";
    return r;
  }

  // Advance the server; write the responses at the head of the queue that
  // are ready (request order).
  void flush() {
    if (srv) {
      uint32_t done = 0;
      (void)hm_server_progress(srv, &done);
    }
    while (!pending.empty()) {
      Pending& h = pending.front();
      if (!h.ready) {
        const int st = hm_server_state(srv, h.id);
        if (st == 0 || st == 1) break;  // queued or running
        h.response = st == 2 ? graph_response(h)
                             : error_response(holomap_error(hm_server_fetch(srv, h.id, nullptr),
                                                            h.n, h.k, 0));
        h.ready = true;
      }
      fwrite(h.response.data(), 1, h.response.size(), stdout);
      fputc('\n', stdout);
      pending.pop_front();
    }
    fflush(stdout);
  }

  ~Frontdoor() {
    if (srv) hm_server_destroy(srv);
    if (ctx) hm_ctx_destroy(ctx);
  }
};

bool blank(const std::string& l) {
  for (char c : l)
    if (c != ' ' && c != '\t' && c != '\r') return false;
  return true;
}

}  // namespace

int main() {
  Frontdoor fd;
  std::string buf;
  bool eof = false;
  char chunk[1 << 16];
  while (!eof || !fd.pending.empty()) {
    // read what stdin holds; wait for it only when nothing is in flight
    if (!eof) {
      struct pollfd p = {0, POLLIN, 0};
      const int r = poll(&p, 1, fd.pending.empty() ? -1 : 1);
      if (r > 0) {
        const ssize_t got = read(0, chunk, sizeof chunk);
        if (got <= 0) eof = true;
        else buf.append(chunk, (size_t)got);
      }
      size_t nl;
      while ((nl = buf.find('\n')) != std::string::npos) {
        std::string line = buf.substr(0, nl);
        buf.erase(0, nl + 1);
        if (!blank(line)) fd.handle(line);
      }
      if (eof && !buf.empty()) {  // a last line without its newline
        if (!blank(buf)) fd.handle(buf);
        buf.clear();
      }
    }
    fd.flush();
  }
  return 0;
}
