// mtx_io.cpp -- MatrixMarket coordinate files to CSR, on the host.
//
// Replaces the Python loader of the reference's SuiteSparse harness
// (tests/test_spmv.py:103-176, load_mtx_to_csr_torch) with the same reading
// rules:
//   * the banner's 4th / 5th tokens give field and symmetry (default real /
//     general); '%' lines are comments; the first other non-empty line is
//     "rows cols [entries]"; only the first `entries` data lines are read;
//   * pattern entries are 1.0; otherwise the value is the third token read
//     as a double (for complex files this is the real part);
//   * entries outside the matrix are dropped; symmetric and hermitian files
//     mirror off-diagonal entries with the same value, skew-symmetric with the
//     negated value;
//   * repeated (row, col) pairs are summed in fp64 in file order (a mirrored
//     entry counts at the position of the line that produced it);
//   * rows come out with ascending columns.
// The sort is a stable counting sort by row then by column, so the summation
// order of duplicates is the file order, as in the reference's dict update.
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "flagsparse_hip.h"

namespace {

thread_local char g_err[512];

int err(int code, const char* fmt, const char* a = "", long long b = 0) {
  std::snprintf(g_err, sizeof(g_err), fmt, a, b);
  return code;
}

struct Parsed {
  int64_t n_rows = 0, n_cols = 0;
  std::vector<int64_t> row_ptr;
  std::vector<int64_t> cols;
  std::vector<double> vals;
  int32_t field = FS_MTX_REAL, symmetry = FS_MTX_GENERAL;
};

bool lower_eq(const char* s, size_t n, const char* lit) {
  if (std::strlen(lit) != n) return false;
  for (size_t i = 0; i < n; ++i) {
    char c = s[i];
    if (c >= 'A' && c <= 'Z') c = (char)(c - 'A' + 'a');
    if (c != lit[i]) return false;
  }
  return true;
}

// whitespace-separated tokens of one line
void split(const std::string& line, std::vector<std::pair<size_t, size_t>>& tok) {
  tok.clear();
  size_t i = 0, n = line.size();
  while (i < n) {
    while (i < n && (line[i] == ' ' || line[i] == '\t' || line[i] == '\r' || line[i] == '\n' ||
                     line[i] == '\v' || line[i] == '\f'))
      ++i;
    if (i >= n) break;
    size_t j = i;
    while (j < n && !(line[j] == ' ' || line[j] == '\t' || line[j] == '\r' || line[j] == '\n' ||
                      line[j] == '\v' || line[j] == '\f'))
      ++j;
    tok.emplace_back(i, j - i);
    i = j;
  }
}

bool parse_int(const std::string& line, std::pair<size_t, size_t> t, long long* out) {
  std::string s = line.substr(t.first, t.second);
  char* end = nullptr;
  errno = 0;
  const long long v = std::strtoll(s.c_str(), &end, 10);
  if (errno || end == s.c_str() || *end) return false;
  *out = v;
  return true;
}

bool parse_double(const std::string& line, std::pair<size_t, size_t> t, double* out) {
  std::string s = line.substr(t.first, t.second);
  char* end = nullptr;
  const double v = std::strtod(s.c_str(), &end);
  if (end == s.c_str() || *end) return false;
  *out = v;
  return true;
}

bool read_line(FILE* f, std::string& line) {
  line.clear();
  char buf[4096];
  while (std::fgets(buf, sizeof(buf), f)) {
    line += buf;
    if (!line.empty() && line.back() == '\n') return true;
  }
  return !line.empty();
}

struct Triple {
  int64_t r, c;
  double v;
};

int parse(const char* path, Parsed* P) {
  FILE* f = std::fopen(path, "r");
  if (!f) return err(FS_ERR_VALUE, "fs_mtx_read: cannot open %s", path);
  std::string line;
  std::vector<std::pair<size_t, size_t>> tok;
  bool have_header = false;
  long long rows = 0, cols = 0, entries = 0, seen = 0;
  std::vector<Triple> trip;
  int rc = FS_OK;
  while (read_line(f, line)) {
    split(line, tok);
    if (tok.empty()) continue;  // blank line
    const char* first = line.c_str() + tok[0].first;
    if (tok[0].second >= 14 && std::strncmp(first, "%%MatrixMarket", 14) == 0) {
      if (tok.size() >= 5) {
        const char* fd = line.c_str() + tok[3].first;
        const char* sy = line.c_str() + tok[4].first;
        if (lower_eq(fd, tok[3].second, "pattern")) P->field = FS_MTX_PATTERN;
        else if (lower_eq(fd, tok[3].second, "complex")) P->field = FS_MTX_COMPLEX;
        else if (lower_eq(fd, tok[3].second, "integer")) P->field = FS_MTX_INTEGER;
        else P->field = FS_MTX_REAL;
        if (lower_eq(sy, tok[4].second, "symmetric")) P->symmetry = FS_MTX_SYMMETRIC;
        else if (lower_eq(sy, tok[4].second, "hermitian")) P->symmetry = FS_MTX_HERMITIAN;
        else if (lower_eq(sy, tok[4].second, "skew-symmetric")) P->symmetry = FS_MTX_SKEW;
        else P->symmetry = FS_MTX_GENERAL;
      }
      continue;
    }
    if (first[0] == '%') continue;
    if (!have_header) {
      if (tok.size() < 2 || !parse_int(line, tok[0], &rows) || !parse_int(line, tok[1], &cols) ||
          (tok.size() > 2 && !parse_int(line, tok[2], &entries))) {
        rc = err(FS_ERR_VALUE, "fs_mtx_read: bad size line in %s", path);
        break;
      }
      if (tok.size() <= 2) entries = 0;
      if (rows < 0 || cols < 0 || entries < 0) {
        rc = err(FS_ERR_VALUE, "fs_mtx_read: negative sizes in %s", path);
        break;
      }
      have_header = true;
      try {
        trip.reserve((size_t)(P->symmetry == FS_MTX_GENERAL ? entries : 2 * entries));
      } catch (const std::bad_alloc&) {
        rc = err(FS_ERR_VALUE, "fs_mtx_read: %s declares too many entries", path);
        break;
      }
      continue;
    }
    if (seen >= entries) break;  // only the declared number of entries
    ++seen;
    long long r1 = 0, c1 = 0;
    double v = 1.0;
    if (tok.size() < 2 || !parse_int(line, tok[0], &r1) || !parse_int(line, tok[1], &c1)) {
      rc = err(FS_ERR_VALUE, "fs_mtx_read: %s: bad entry line %lld", path, seen);
      break;
    }
    if (P->field != FS_MTX_PATTERN) {
      if (tok.size() < 3 || !parse_double(line, tok[2], &v)) {
        rc = err(FS_ERR_VALUE, "fs_mtx_read: %s: bad value on entry line %lld", path, seen);
        break;
      }
    }
    const int64_t r = r1 - 1, c = c1 - 1;
    if (r < 0 || r >= rows || c < 0 || c >= cols) continue;
    trip.push_back({r, c, v});
    if (r != c && c < rows && r < cols) {
      if (P->symmetry == FS_MTX_SYMMETRIC || P->symmetry == FS_MTX_HERMITIAN)
        trip.push_back({c, r, v});
      else if (P->symmetry == FS_MTX_SKEW)
        trip.push_back({c, r, -v});
    }
  }
  std::fclose(f);
  if (rc) return rc;
  if (!have_header) return err(FS_ERR_VALUE, "fs_mtx_read: cannot parse the header of %s", path);
  P->n_rows = rows;
  P->n_cols = cols;
  // stable counting sort by column, then by row: rows ascending, columns
  // ascending inside a row, equal pairs in file order
  const size_t m = trip.size();
  std::vector<int64_t> cnt((size_t)cols + 1, 0);
  for (const Triple& t : trip) ++cnt[(size_t)t.c + 1];
  for (size_t i = 1; i < cnt.size(); ++i) cnt[i] += cnt[i - 1];
  std::vector<Triple> by_col(m);
  for (const Triple& t : trip) by_col[(size_t)cnt[(size_t)t.c]++] = t;
  std::vector<int64_t> rcnt((size_t)rows + 1, 0);
  for (const Triple& t : by_col) ++rcnt[(size_t)t.r + 1];
  for (size_t i = 1; i < rcnt.size(); ++i) rcnt[i] += rcnt[i - 1];
  std::vector<Triple> sorted(m);
  for (const Triple& t : by_col) sorted[(size_t)rcnt[(size_t)t.r]++] = t;
  // sum repeated pairs (adjacent now) in order
  P->row_ptr.assign((size_t)rows + 1, 0);
  P->cols.clear();
  P->vals.clear();
  P->cols.reserve(m);
  P->vals.reserve(m);
  for (size_t i = 0; i < m;) {
    size_t j = i + 1;
    double s = sorted[i].v;
    while (j < m && sorted[j].r == sorted[i].r && sorted[j].c == sorted[i].c) s += sorted[j++].v;
    P->cols.push_back(sorted[i].c);
    P->vals.push_back(s);
    ++P->row_ptr[(size_t)sorted[i].r + 1];
    i = j;
  }
  for (size_t i = 1; i < P->row_ptr.size(); ++i) P->row_ptr[i] += P->row_ptr[i - 1];
  return FS_OK;
}

}  // namespace

extern "C" {

int fs_mtx_read(const char* path, fs_mtx_info* info, void** handle) {
  if (!path || !info || !handle) return err(FS_ERR_VALUE, "fs_mtx_read: null argument%s", "");
  *handle = nullptr;
  Parsed* P = new (std::nothrow) Parsed;
  if (!P) return err(FS_ERR_VALUE, "fs_mtx_read: out of memory%s", "");
  int rc;
  try {
    rc = parse(path, P);
  } catch (const std::bad_alloc&) {
    rc = err(FS_ERR_VALUE, "fs_mtx_read: out of memory reading %s", path);
  }
  if (rc) {
    delete P;
    return rc;
  }
  info->n_rows = P->n_rows;
  info->n_cols = P->n_cols;
  info->nnz = (int64_t)P->cols.size();
  info->field = P->field;
  info->symmetry = P->symmetry;
  *handle = P;
  return FS_OK;
}

int fs_mtx_fill(void* handle, int64_t* row_ptr, int64_t* col_idx, double* values) {
  Parsed* P = static_cast<Parsed*>(handle);
  if (!P) return err(FS_ERR_VALUE, "fs_mtx_fill: null handle%s", "");
  if (row_ptr) std::memcpy(row_ptr, P->row_ptr.data(), P->row_ptr.size() * sizeof(int64_t));
  if (col_idx && !P->cols.empty())
    std::memcpy(col_idx, P->cols.data(), P->cols.size() * sizeof(int64_t));
  if (values && !P->vals.empty())
    std::memcpy(values, P->vals.data(), P->vals.size() * sizeof(double));
  return FS_OK;
}

void fs_mtx_free(void* handle) { delete static_cast<Parsed*>(handle); }

const char* fs_mtx_last_error(void) { return g_err; }

}  // extern "C"
