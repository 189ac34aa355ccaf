// Tuning tables and the SpMM autotuner behind the C ABI.
//
// A table maps keys to a variant and knobs; it is stored as JSON in the format the
// Python autotuner writes (segscan_amd/autotune.py), so both read and write the same
// files. Keys:
//   bucket  "<arch>|rocm<major>.<minor>|spmm|<dtype>|N<n>|skew<c>|<size>"
//   matrix  bucket key + "|m<rows>x<cols>:<nnz>:<longest row>"
// n = N rounded up to a power of two; c = 0 for even rows (longest row <= 4 x mean + 32),
// else 1 + floor(log2(max / mean) / 4) capped at 3; size "s" below 2^20 nonzeros, else
// "l". The arch and ROCm version are part of every key: entries tuned under another
// ROCm release or for another GPU are dropped when a table is loaded. Bucket knobs may
// hold "waves_per_cu" in place of a block factor, resolved per matrix at lookup.
//
// The autotuner replaces the reference's brute-force parameter sweeps
// (test-cg/src/params.md, driven by test-cg/src/test.py:42-59): it times a fixed
// candidate list (the reference's knob points mapped to wave64 plus gfx950 points) with
// HIP events on the caller's stream, a spin kernel queued before each start event so
// that the host's issue cost overlaps it, and keeps the fastest.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>
#include <hip/hip_version.h>

#include "../../include/segscan.h"
#include "abi_internal.hpp"

namespace segscan {

// Spins for `ticks` of the 100 MHz constant clock (always terminates).
__global__ void tune_spin_kernel(int64_t ticks) {
  const int64_t t0 = wall_clock64();
  while (wall_clock64() - t0 < ticks) {
  }
}

}  // namespace segscan

struct segscan_tune_table {
  struct Entry {
    std::string raw;  // the value's JSON text, written back as is
    int variant = SEGSCAN_SPMM_AUTO;
    segscan_knobs_t knobs{};
    int waves_per_cu = 0;
    bool has_knobs = false;
  };
  std::map<std::string, Entry> entries;
};

namespace {

using segscan::set_error;

segscan_status_t einval(const char* msg) { return set_error(SEGSCAN_EINVAL, msg); }

const char* kArch = "gfx950";

std::string prefix() {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%s|rocm%d.%d", kArch, HIP_VERSION_MAJOR, HIP_VERSION_MINOR);
  return buf;
}

const char* const kVariantNames[] = {"ebpr", "rbpr", "ebsr", "rbsr", "plain_ebpr", "plain_rbpr",
                                     "plain_ebsr", "plain_rbsr", "auto"};
const char* const kKnobNames[] = {"group_factor", "tile_factor", "thread_per_block", "block_numer",
                                  "block_denom", "coarsen_factor", "thread_nz", "dtile_size",
                                  "dtile_num", "block_dim_y", "write_policy"};

int32_t* knob_field(segscan_knobs_t& k, int i) {
  int32_t* f[] = {&k.group_factor, &k.tile_factor, &k.thread_per_block, &k.block_numer, &k.block_denom,
                  &k.coarsen_factor, &k.thread_nz, &k.dtile_size, &k.dtile_num, &k.block_dim_y, &k.write_policy};
  return f[i];
}

// ---- a small JSON reader: objects, arrays, strings, numbers, literals ----

struct Json {
  const char* s;
  size_t n, i = 0;
  bool ok = true;
  explicit Json(const std::string& t) : s(t.data()), n(t.size()) {}
  void ws() {
    while (i < n && (s[i] == ' ' || s[i] == '\n' || s[i] == '\r' || s[i] == '\t')) ++i;
  }
  bool eat(char c) {
    ws();
    if (i < n && s[i] == c) {
      ++i;
      return true;
    }
    return false;
  }
  bool string(std::string* out) {
    ws();
    if (i >= n || s[i] != '"') return ok = false;
    ++i;
    std::string r;
    while (i < n && s[i] != '"') {
      if (s[i] == '\\') {
        if (i + 1 >= n) return ok = false;
        const char e = s[i + 1];
        if (e == 'u') {
          if (i + 5 >= n) return ok = false;
          r += '?';  // keys and names here are ASCII
          i += 6;
          continue;
        }
        r += (e == 'n') ? '\n' : (e == 't') ? '\t' : e;
        i += 2;
        continue;
      }
      r += s[i++];
    }
    if (i >= n) return ok = false;
    ++i;
    if (out) *out = r;
    return true;
  }
  bool number(double* out) {
    ws();
    const size_t b = i;
    while (i < n && (std::strchr("+-0123456789.eE", s[i]) != nullptr)) ++i;
    if (b == i) return ok = false;
    if (out) *out = std::strtod(std::string(s + b, i - b).c_str(), nullptr);
    return true;
  }
  // skips any value; [b, e) is its text
  bool value(size_t* b, size_t* e) {
    ws();
    *b = i;
    if (i >= n) return ok = false;
    const char c = s[i];
    if (c == '"') {
      if (!string(nullptr)) return false;
    } else if (c == '{' || c == '[') {
      const char close = c == '{' ? '}' : ']';
      ++i;
      if (!eat(close)) {
        do {
          size_t vb, ve;
          if (c == '{') {
            if (!string(nullptr) || !eat(':')) return ok = false;
          }
          if (!value(&vb, &ve)) return false;
        } while (eat(','));
        if (!eat(close)) return ok = false;
      }
    } else if (std::strncmp(s + i, "true", 4) == 0) {
      i += 4;
    } else if (std::strncmp(s + i, "false", 5) == 0) {
      i += 5;
    } else if (std::strncmp(s + i, "null", 4) == 0) {
      i += 4;
    } else if (!number(nullptr)) {
      return false;
    }
    *e = i;
    return true;
  }
};

// Parses an entry's object: "variant" (string) and "knobs" (object of integers or null).
bool parse_entry(const std::string& text, segscan_tune_table::Entry& out) {
  Json j(text);
  if (!j.eat('{')) return false;
  out.raw = text;
  bool have_variant = false;
  if (j.eat('}')) return false;
  do {
    std::string key;
    if (!j.string(&key) || !j.eat(':')) return false;
    if (key == "variant") {
      std::string v;
      if (!j.string(&v)) return false;
      for (int k = 0; k <= SEGSCAN_SPMM_AUTO; ++k)
        if (v == kVariantNames[k]) {
          out.variant = k;
          have_variant = true;
        }
      if (!have_variant) return false;
    } else if (key == "knobs") {
      j.ws();
      if (std::strncmp(j.s + j.i, "null", 4) == 0) {
        j.i += 4;
        continue;
      }
      if (!j.eat('{')) return false;
      out.has_knobs = true;
      if (!j.eat('}')) {
        do {
          std::string kn;
          double v = 0;
          if (!j.string(&kn) || !j.eat(':') || !j.number(&v)) return false;
          if (kn == "waves_per_cu") {
            out.waves_per_cu = static_cast<int>(v);
            continue;
          }
          bool known = false;
          for (int f = 0; f < 11; ++f)
            if (kn == kKnobNames[f]) {
              *knob_field(out.knobs, f) = static_cast<int32_t>(v);
              known = true;
            }
          if (!known) return false;
        } while (j.eat(','));
        if (!j.eat('}')) return false;
      }
    } else {
      size_t b, e;
      if (!j.value(&b, &e)) return false;
    }
  } while (j.eat(','));
  return j.eat('}') && have_variant;
}

// ---- keys ----

struct RowStats {
  int64_t rows, nnz, max;
};

int skew_class(const RowStats& st) {
  const double mean = st.rows > 0 ? static_cast<double>(st.nnz) / static_cast<double>(st.rows) : 0.0;
  if (mean <= 0.0 || static_cast<double>(st.max) <= 4.0 * mean + 32.0) return 0;
  return 1 + std::min(2, static_cast<int>(std::floor(std::log2(static_cast<double>(st.max) / mean) / 4.0)));
}

std::string bucket_key(const RowStats& st, int64_t N, segscan_dtype_t dt) {
  int64_t nb = 1;
  while (nb < N) nb <<= 1;
  const char* d = dt == SEGSCAN_F32 ? "f32" : dt == SEGSCAN_BF16 ? "bf16" : "f16";
  char buf[160];
  std::snprintf(buf, sizeof(buf), "%s|spmm|%s|N%lld|skew%d|%s", prefix().c_str(), d, (long long)nb, skew_class(st),
                st.nnz < (int64_t{1} << 20) ? "s" : "l");
  return buf;
}

std::string matrix_key(const RowStats& st, int64_t ncol, int64_t N, segscan_dtype_t dt) {
  char buf[96];
  std::snprintf(buf, sizeof(buf), "|m%lldx%lld:%lld:%lld", (long long)st.rows, (long long)ncol, (long long)st.nnz,
                (long long)st.max);
  return bucket_key(st, N, dt) + buf;
}

// "waves_per_cu" -> a grid of min(chunks / groups per workgroup, 256 CUs x waves x 64 /
// workgroup) workgroups as the block factor 1 / (rows / grid) (autotune.resolve_knobs).
segscan_knobs_t resolve(const segscan_tune_table::Entry& e, const RowStats& st) {
  segscan_knobs_t k = e.knobs;
  if (e.waves_per_cu <= 0) return k;
  const int gf = k.group_factor;
  const int64_t g = gf > 0 ? (int64_t{1} << gf) : 1;
  const int64_t chunk = static_cast<int64_t>(k.thread_nz) * g;
  const int64_t nchunks = chunk > 0 ? (std::max<int64_t>(1, st.nnz) + chunk - 1) / chunk : 1;
  const int64_t tpb = std::max(1, k.thread_per_block);
  const int64_t gpb = std::max<int64_t>(1, tpb / g);
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((nchunks + gpb - 1) / gpb,
                                                                 256 * int64_t{e.waves_per_cu} * 64 / tpb));
  k.block_numer = 1;
  k.block_denom = static_cast<int32_t>(std::max<int64_t>(1, st.rows / blocks));
  return k;
}

std::string entry_json(int variant, const segscan_knobs_t& k, double ms) {
  std::string r = "{\"knobs\": {";
  bool first = true;
  segscan_knobs_t kk = k;
  for (int f = 0; f < 11; ++f) {
    const int32_t v = *knob_field(kk, f);
    if (v == 0) continue;
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%s\"%s\": %d", first ? "" : ", ", kKnobNames[f], v);
    r += buf;
    first = false;
  }
  char tail[96];
  std::snprintf(tail, sizeof(tail), "}, \"ms\": %.5f, \"variant\": \"%s\"}", ms, kVariantNames[variant]);
  return r + tail;
}

segscan_knobs_t mk(int gf, int tf, int tpb, int bn, int bd, int cf, int tnz, int policy) {
  segscan_knobs_t k{};
  k.group_factor = gf;
  k.tile_factor = tf;
  k.thread_per_block = tpb;
  k.block_numer = bn;
  k.block_denom = bd;
  k.coarsen_factor = cf;
  k.thread_nz = tnz;
  k.write_policy = policy;
  return k;
}

// The candidate list (autotune.spmm_candidates, same points and order).
void candidates(int64_t N, segscan_dtype_t dt, const RowStats& st,
                std::vector<std::pair<int, segscan_knobs_t>>& out) {
  const int64_t rows = std::max<int64_t>(1, st.rows), nnz = std::max<int64_t>(1, st.nnz);
  const int vec = dt == SEGSCAN_F32 ? 4 : 8;
  int cf = vec;
  while (cf > 1 && N % cf) cf /= 2;
  const int64_t lanes = std::min<int64_t>(64, (N + cf - 1) / cf);
  int g = 1;
  while (g < lanes) g *= 2;
  const segscan_knobs_t none{};
  for (int v = SEGSCAN_SPMM_PLAIN_EBPR; v <= SEGSCAN_SPMM_PLAIN_RBSR; ++v) out.push_back({v, none});
  for (int v = SEGSCAN_SPMM_EBPR; v <= SEGSCAN_SPMM_RBSR; ++v) out.push_back({v, none});
  const int eb_pts[][2] = {{5, 5}, {4, 4}, {3, 3}, {2, 2}, {6, 6}, {4, 5}, {3, 5}};
  for (auto& q : eb_pts) out.push_back({SEGSCAN_SPMM_EBPR, mk(q[0], q[1], 256, 1, 16, 0, 0, 0)});
  const int rb_pts[][2] = {{3, 3}, {2, 2}, {4, 4}, {5, 5}, {6, 6}, {2, 3}};
  for (auto& q : rb_pts) out.push_back({SEGSCAN_SPMM_RBPR, mk(q[0], q[1], 256, 1, 2, 0, 0, 0)});
  std::vector<int> cfs = {1, 2, cf};
  std::sort(cfs.begin(), cfs.end());
  cfs.erase(std::unique(cfs.begin(), cfs.end()), cfs.end());
  for (int c : cfs)
    for (int tf : {5, 6}) out.push_back({SEGSCAN_SPMM_RBSR, mk(0, tf, 256, 0, 0, c, 0, 0)});
  std::vector<std::pair<int, int>> shapes;
  for (int c = cf, gg = g; c >= 1 && gg <= 64 && shapes.size() < 3; c /= 2, gg *= 2) shapes.push_back({c, gg});
  for (auto& sh : shapes) {
    const int c = sh.first, gg = sh.second;
    int gf = 0;
    while ((1 << gf) < gg) ++gf;
    for (int tnz : {2, 4, 8, 16, 32, 64}) {
      const int64_t chunk = static_cast<int64_t>(tnz) * gg;
      if (chunk < 32 || chunk > 2048) continue;
      const int64_t nchunks = (nnz + chunk - 1) / chunk;
      for (int tpb : {256, 512}) {
        const int64_t gpb = tpb / gg;
        for (int waves : {4, 8, 16}) {
          const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((nchunks + gpb - 1) / gpb,
                                                                         256 * int64_t{waves} * 64 / tpb));
          const int bd = static_cast<int>(std::max<int64_t>(1, rows / std::max<int64_t>(1, blocks)));
          const int pols[2] = {chunk >= 64 ? SEGSCAN_WRITE_OWNER : SEGSCAN_WRITE_LESSATOM,
                               chunk >= 64 ? SEGSCAN_WRITE_LESSATOM : SEGSCAN_WRITE_ATOMIC};
          for (int pol : pols) out.push_back({SEGSCAN_SPMM_EBSR, mk(gf ? gf : -1, 0, tpb, 1, bd, c, tnz, pol)});
        }
      }
    }
  }
}

segscan_status_t stats_of(const segscan_csr_t* A, bool read_back, void* stream, RowStats* st) {
  st->rows = A->nrow;
  st->nnz = A->nnz;
  st->max = A->max_row_nnz;
  if (st->max > 0 || A->nnz == 0) return SEGSCAN_OK;
  if (!read_back) return einval("max_row_nnz must be given (the longest row is part of the table key)");
  std::vector<int32_t> rp(static_cast<size_t>(A->nrow) + 1);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = hipMemcpyAsync(rp.data(), A->rowptr, rp.size() * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return set_error(SEGSCAN_EHIP, hipGetErrorString(e));
  int64_t mx = 0;
  for (int64_t r = 0; r < A->nrow; ++r) mx = std::max<int64_t>(mx, static_cast<int64_t>(rp[r + 1]) - rp[r]);
  st->max = mx;
  return SEGSCAN_OK;
}

}  // namespace

extern "C" {

const char* segscan_tune_prefix(void) {
  static const std::string p = prefix();
  return p.c_str();
}

segscan_status_t segscan_tune_table_create(segscan_tune_table_t* table) {
  if (!table) return einval("table is NULL");
  *table = new segscan_tune_table();
  return SEGSCAN_OK;
}

segscan_status_t segscan_tune_table_destroy(segscan_tune_table_t table) {
  delete table;
  return SEGSCAN_OK;
}

segscan_status_t segscan_tune_table_load(segscan_tune_table_t table, const char* path, int64_t* loaded,
                                         int64_t* dropped) {
  if (!table || !path) return einval("table or path is NULL");
  FILE* f = std::fopen(path, "rb");
  if (!f) return einval("cannot open the tuning table file");
  std::string text;
  char buf[65536];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
  std::fclose(f);
  Json j(text);
  int64_t nl = 0, nd = 0;
  std::map<std::string, segscan_tune_table::Entry> fresh;
  const std::string pre = prefix() + "|";
  if (!j.eat('{')) return einval("tuning table: not a JSON object");
  if (!j.eat('}')) {
    do {
      std::string key;
      size_t b, e;
      if (!j.string(&key) || !j.eat(':') || !j.value(&b, &e)) return einval("tuning table: malformed JSON");
      if (key.compare(0, pre.size(), pre) != 0) {
        ++nd;  // tuned for another GPU or ROCm release
        continue;
      }
      segscan_tune_table::Entry en;
      if (!parse_entry(text.substr(b, e - b), en)) return einval("tuning table: malformed entry");
      fresh[key] = en;
      ++nl;
    } while (j.eat(','));
    if (!j.eat('}')) return einval("tuning table: malformed JSON");
  }
  for (auto& kv : fresh) table->entries[kv.first] = kv.second;
  if (loaded) *loaded = nl;
  if (dropped) *dropped = nd;
  set_error(SEGSCAN_OK, "");
  return SEGSCAN_OK;
}

segscan_status_t segscan_tune_table_save(segscan_tune_table_t table, const char* path) {
  if (!table || !path) return einval("table or path is NULL");
  FILE* f = std::fopen(path, "wb");
  if (!f) return einval("cannot write the tuning table file");
  std::fputs("{", f);
  bool first = true;
  for (auto& kv : table->entries) {
    std::fprintf(f, "%s\n \"%s\": %s", first ? "" : ",", kv.first.c_str(), kv.second.raw.c_str());
    first = false;
  }
  std::fputs("\n}\n", f);
  const bool bad = std::ferror(f) != 0;
  std::fclose(f);
  if (bad) return einval("write error on the tuning table file");
  return SEGSCAN_OK;
}

segscan_status_t segscan_tune_table_lookup(segscan_tune_table_t table, const segscan_csr_t* A, int64_t N,
                                           segscan_dtype_t b_dtype, segscan_spmm_variant_t* variant,
                                           segscan_knobs_t* knobs, int32_t* found) {
  if (!table || !A || !variant || !knobs) return einval("NULL argument");
  if (N < 1) return einval("N must be positive");
  if (b_dtype != SEGSCAN_F32 && b_dtype != SEGSCAN_BF16 && b_dtype != SEGSCAN_F16) return einval("unknown dtype");
  RowStats st;
  segscan_status_t s = stats_of(A, false, nullptr, &st);
  if (s != SEGSCAN_OK) return s;
  const std::string keys[2] = {matrix_key(st, A->ncol, N, b_dtype), bucket_key(st, N, b_dtype)};
  for (int i = 0; i < 2; ++i) {
    auto it = table->entries.find(keys[i]);
    if (it == table->entries.end()) continue;
    *variant = static_cast<segscan_spmm_variant_t>(it->second.variant);
    *knobs = resolve(it->second, st);
    if (found) *found = i == 0 ? 2 : 1;
    return SEGSCAN_OK;
  }
  *variant = SEGSCAN_SPMM_AUTO;
  std::memset(knobs, 0, sizeof(*knobs));
  if (found) *found = 0;
  return SEGSCAN_OK;
}

segscan_status_t segscan_autotune_workspace_size(const segscan_csr_t* A, int64_t N, segscan_dtype_t b_dtype,
                                                 size_t* bytes) {
  if (!A || !bytes) return einval("NULL argument");
  RowStats st{A->nrow, A->nnz, A->max_row_nnz};
  std::vector<std::pair<int, segscan_knobs_t>> cands;
  candidates(N, b_dtype, st, cands);
  cands.push_back({SEGSCAN_SPMM_AUTO, segscan_knobs_t{}});
  size_t most = 0;
  for (auto& c : cands) {
    size_t b = 0;
    if (segscan_spmm_workspace_size(A, N, b_dtype, static_cast<segscan_spmm_variant_t>(c.first),
                                    c.first == SEGSCAN_SPMM_AUTO ? nullptr : &c.second, &b) == SEGSCAN_OK)
      most = std::max(most, b);
  }
  *bytes = most;
  set_error(SEGSCAN_OK, "");
  return SEGSCAN_OK;
}

segscan_status_t segscan_autotune_spmm(segscan_tune_table_t table, const segscan_csr_t* A, int64_t N,
                                       const void* B, int64_t ldb, segscan_dtype_t b_dtype, float* C, int64_t ldc,
                                       int32_t reps, void* workspace, size_t workspace_bytes, void* stream,
                                       segscan_spmm_variant_t* variant, segscan_knobs_t* knobs, float* ms) {
  if (!A || !variant || !knobs) return einval("NULL argument");
  if (reps < 1 || reps > 1000) return einval("reps must be in [1, 1000]");
  if (N < 1) return einval("N must be positive");
  if (b_dtype != SEGSCAN_F32 && b_dtype != SEGSCAN_BF16 && b_dtype != SEGSCAN_F16) return einval("unknown dtype");
  hipStream_t s = static_cast<hipStream_t>(stream);
  RowStats st;
  segscan_status_t status = stats_of(A, true, stream, &st);
  if (status != SEGSCAN_OK) return status;
  segscan_csr_t Ad = *A;
  Ad.max_row_nnz = st.max;
  std::vector<std::pair<int, segscan_knobs_t>> cands;
  cands.push_back({SEGSCAN_SPMM_AUTO, segscan_knobs_t{}});
  candidates(N, b_dtype, st, cands);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e != hipSuccess) return set_error(SEGSCAN_EHIP, hipGetErrorString(e));
  // spin ticks (100 MHz clock) queued before each start event: 0.3 ms
  const int64_t spin = 30000;
  float best = INFINITY;
  int best_v = SEGSCAN_SPMM_AUTO;
  segscan_knobs_t best_k{};
  segscan_status_t last_err = SEGSCAN_OK;
  for (auto& c : cands) {
    const segscan_knobs_t* kp = c.first == SEGSCAN_SPMM_AUTO ? nullptr : &c.second;
    const auto v = static_cast<segscan_spmm_variant_t>(c.first);
    size_t need = 0;
    if (segscan_spmm_workspace_size(&Ad, N, b_dtype, v, kp, &need) != SEGSCAN_OK) continue;  // not compiled
    if (need > workspace_bytes) continue;
    // one untimed call (also the check that the candidate runs at all)
    status = segscan_spmm_csr(&Ad, N, B, ldb, b_dtype, C, ldc, v, kp, 0, workspace, workspace_bytes, stream);
    if (status == SEGSCAN_EUNSUPPORTED) continue;
    if (status != SEGSCAN_OK) {
      last_err = status;
      break;
    }
    std::vector<float> t;
    for (int r = 0; r < reps && status == SEGSCAN_OK; ++r) {
      hipLaunchKernelGGL(segscan::tune_spin_kernel, dim3(1), dim3(1), 0, s, spin);
      e = hipEventRecord(e0, s);
      if (e == hipSuccess)
        status = segscan_spmm_csr(&Ad, N, B, ldb, b_dtype, C, ldc, v, kp, 0, workspace, workspace_bytes, stream);
      if (e == hipSuccess && status == SEGSCAN_OK) e = hipEventRecord(e1, s);
      if (e == hipSuccess && status == SEGSCAN_OK) e = hipEventSynchronize(e1);
      float dt = 0.f;
      if (e == hipSuccess && status == SEGSCAN_OK) e = hipEventElapsedTime(&dt, e0, e1);
      if (e != hipSuccess) break;
      t.push_back(dt);
    }
    if (e != hipSuccess || status != SEGSCAN_OK) {
      last_err = e != hipSuccess ? SEGSCAN_EHIP : status;
      break;
    }
    std::nth_element(t.begin(), t.begin() + t.size() / 2, t.end());
    const float med = t[t.size() / 2];
    if (med < best) {
      best = med;
      best_v = c.first;
      best_k = c.second;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (last_err != SEGSCAN_OK) {
    if (last_err == SEGSCAN_EHIP) return set_error(SEGSCAN_EHIP, hipGetErrorString(e));
    return last_err;
  }
  if (!std::isfinite(best)) return set_error(SEGSCAN_EWORKSPACE, "no candidate fits the workspace given");
  *variant = static_cast<segscan_spmm_variant_t>(best_v);
  *knobs = best_k;
  if (ms) *ms = best;
  if (table) {
    segscan_tune_table::Entry en;
    en.variant = best_v;
    en.knobs = best_k;
    en.has_knobs = true;
    en.raw = entry_json(best_v, best_k, best);
    table->entries[matrix_key(st, A->ncol, N, b_dtype)] = en;
  }
  set_error(SEGSCAN_OK, "");
  return SEGSCAN_OK;
}

}  // extern "C"
