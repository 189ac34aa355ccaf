// Matrix Market reader: the reference's read_mtx_file
// (test-cg/include/dataloader/dataloader.hpp:22-107) restated on the host.
//  * banner "%%MatrixMarket matrix coordinate <field> <symmetry>", '%' comments,
//    size line "nrow ncol nnz", then nnz entries "row col [value [imag]]", 1-based;
//  * symmetric matrices get every off-diagonal entry mirrored, then the list is
//    sorted by (row, col) and exact duplicates removed (dataloader.hpp:69-86);
//    general matrices are sorted without de-duplication (:87-89);
//  * CSR arrays are built from the sorted list (:92-106).
// Differences by default: values can be kept (SEGSCAN_MTX_KEEP_VALUES; the
// reference drops them and later fills ones, dataloader.hpp:154-155),
// skew-symmetric entries are mirrored with a negated value and hermitian ones
// with their real part (the reference mirrors only "symmetric", mm_is_symmetric),
// every mirrored entry is kept (the reference's CSR loop stops at the declared
// entry count, dataloader.hpp:95, which drops the tail of a symmetric matrix),
// and malformed files return an error instead of calling exit().
// SEGSCAN_MTX_REFERENCE_QUIRKS restores the reference's mirroring and truncation.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/segscan.h"

namespace {

struct Entry {
  int32_t r, c;
  float v;
};

thread_local std::string g_mtx_error;

segscan_status_t mtx_fail(const std::string& msg) {
  g_mtx_error = msg;
  return SEGSCAN_EINVAL;
}

std::string lower(std::string s) {
  for (auto& ch : s) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
  return s;
}

}  // namespace

extern "C" {

segscan_status_t segscan_mtx_read(const char* path, uint32_t flags, segscan_host_csr_t* out) {
  if (!path || !out) return mtx_fail("NULL argument");
  if (flags & ~(SEGSCAN_MTX_KEEP_VALUES | SEGSCAN_MTX_REFERENCE_QUIRKS)) return mtx_fail("unknown flag bits");
  const bool keep_values = (flags & SEGSCAN_MTX_KEEP_VALUES) != 0;
  const bool quirks = (flags & SEGSCAN_MTX_REFERENCE_QUIRKS) != 0;
  std::memset(out, 0, sizeof(*out));
  FILE* f = std::fopen(path, "r");
  if (!f) return mtx_fail(std::string("cannot open ") + path);
  char line[4096];
  if (!std::fgets(line, sizeof(line), f)) {
    std::fclose(f);
    return mtx_fail("empty file");
  }
  char banner[64] = {0}, object[64] = {0}, format[64] = {0}, field[64] = {0}, symmetry[64] = {0};
  if (std::sscanf(line, "%63s %63s %63s %63s %63s", banner, object, format, field, symmetry) != 5 ||
      std::string(banner) != "%%MatrixMarket") {
    std::fclose(f);
    return mtx_fail("missing %%MatrixMarket banner");
  }
  const std::string obj = lower(object), fmt = lower(format), fld = lower(field), sym = lower(symmetry);
  if (obj != "matrix" || fmt != "coordinate") {
    std::fclose(f);
    return mtx_fail("only coordinate matrices are supported");
  }
  const bool pattern = fld == "pattern";
  const bool complex_field = fld == "complex";
  if (!(pattern || complex_field || fld == "real" || fld == "integer" || fld == "double")) {
    std::fclose(f);
    return mtx_fail("unknown field " + fld);
  }
  const bool symmetric = sym == "symmetric";
  if (!(symmetric || sym == "skew-symmetric" || sym == "hermitian" || sym == "general")) {
    std::fclose(f);
    return mtx_fail("unknown symmetry " + sym);
  }
  const bool skew = !quirks && sym == "skew-symmetric", herm = !quirks && sym == "hermitian";
  long long nrow = -1, ncol = -1, nnz_decl = -1;
  while (std::fgets(line, sizeof(line), f)) {
    if (line[0] == '%') continue;
    bool blank = true;
    for (char* p = line; *p; ++p)
      if (!std::isspace(static_cast<unsigned char>(*p))) blank = false;
    if (blank) continue;
    if (std::sscanf(line, "%lld %lld %lld", &nrow, &ncol, &nnz_decl) != 3) {
      std::fclose(f);
      return mtx_fail("bad size line");
    }
    break;
  }
  if (nrow < 0 || ncol < 0 || nnz_decl < 0 || nrow >= INT32_MAX || ncol >= INT32_MAX) {
    std::fclose(f);
    return mtx_fail("bad or missing sizes");
  }
  std::vector<Entry> coords;
  coords.reserve(static_cast<size_t>(nnz_decl) * ((symmetric || skew || herm) ? 2 : 1));
  for (long long i = 0; i < nnz_decl; ++i) {
    long long r, c;
    double v = 1.0, im = 0.0;
    int got;
    if (pattern) got = std::fscanf(f, "%lld %lld", &r, &c);
    else if (complex_field) got = std::fscanf(f, "%lld %lld %lf %lf", &r, &c, &v, &im);
    else got = std::fscanf(f, "%lld %lld %lf", &r, &c, &v);
    const int want = pattern ? 2 : complex_field ? 4 : 3;
    if (got != want) {
      std::fclose(f);
      return mtx_fail("not enough entries (entry " + std::to_string(i) + ")");
    }
    if (r < 1 || r > nrow || c < 1 || c > ncol) {
      std::fclose(f);
      return mtx_fail("entry out of range at " + std::to_string(i));
    }
    const float fv = keep_values ? static_cast<float>(v) : 1.0f;
    coords.push_back({static_cast<int32_t>(r - 1), static_cast<int32_t>(c - 1), fv});
    if ((symmetric || skew || herm) && r != c) {
      const float mv = keep_values ? (skew ? -fv : fv) : 1.0f;
      coords.push_back({static_cast<int32_t>(c - 1), static_cast<int32_t>(r - 1), mv});
    }
  }
  std::fclose(f);
  std::stable_sort(coords.begin(), coords.end(),
                   [](const Entry& a, const Entry& b) { return a.r != b.r ? a.r < b.r : a.c < b.c; });
  if (symmetric || skew || herm) {
    coords.erase(std::unique(coords.begin(), coords.end(),
                             [](const Entry& a, const Entry& b) { return a.r == b.r && a.c == b.c; }),
                 coords.end());
  }
  if (quirks && coords.size() > static_cast<size_t>(nnz_decl)) coords.resize(static_cast<size_t>(nnz_decl));
  if (coords.size() > static_cast<size_t>(INT32_MAX)) return mtx_fail("too many entries for int32 indices");
  out->nrow = nrow;
  out->ncol = ncol;
  out->nnz = static_cast<int64_t>(coords.size());
  out->rowptr = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * static_cast<size_t>(nrow + 1)));
  out->colind = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * std::max<size_t>(coords.size(), 1)));
  out->vals = static_cast<float*>(std::malloc(sizeof(float) * std::max<size_t>(coords.size(), 1)));
  if (!out->rowptr || !out->colind || !out->vals) {
    segscan_host_csr_free(out);
    return mtx_fail("out of host memory");
  }
  size_t pos = 0;
  out->rowptr[0] = 0;
  for (long long row = 0; row < nrow; ++row) {
    while (pos < coords.size() && coords[pos].r == row) {
      out->colind[pos] = coords[pos].c;
      out->vals[pos] = coords[pos].v;
      ++pos;
    }
    out->rowptr[row + 1] = static_cast<int32_t>(pos);
  }
  g_mtx_error.clear();
  return SEGSCAN_OK;
}

void segscan_host_csr_free(segscan_host_csr_t* csr) {
  if (!csr) return;
  std::free(csr->rowptr);
  std::free(csr->colind);
  std::free(csr->vals);
  csr->rowptr = nullptr;
  csr->colind = nullptr;
  csr->vals = nullptr;
}

const char* segscan_mtx_last_error(void) { return g_mtx_error.c_str(); }

}  // extern "C"
