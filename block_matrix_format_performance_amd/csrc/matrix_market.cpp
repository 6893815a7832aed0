// matrix_market.cpp -- Matrix Market I/O and BCSR block-size detection
// (host code of the C ABI).
//
// The reference reads and writes Matrix Market through the NIST mmio library
// (common/mmio.{h,c}); its csr_matrix_class::write_mm (matrix_converters.h:
// 119-144) writes "coordinate real general" with 1-based "%d %d %10.20g"
// entries.  bmfp_mm_write writes the same text.  bmfp_mm_read parses the
// coordinate (real, integer, pattern; general, symmetric, skew-symmetric) and
// array (real, integer; general) forms straight into CSR.
// bmfp_block_size_detect picks the block size for BCSR storage of a CSR
// pattern by the compulsory bytes of one BCSR SpMV.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

using namespace bmfp;

namespace {

struct Entry {
  int row, col;
  double value;
};

std::string lower(std::string s)
{
  for (char &c : s) c = (char) std::tolower((unsigned char) c);
  return s;
}

// Whole file in memory (Matrix Market files are read once, sequentially).
int slurp(const char *path, std::string *text)
{
  FILE *f = std::fopen(path, "rb");
  if (!f) return fail(BMFP_ERR_INVALID, "cannot open %s: %s", path, std::strerror(errno));
  std::string s;
  char buf[1 << 16];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, got);
  const bool bad = std::ferror(f) != 0;
  std::fclose(f);
  if (bad) return fail(BMFP_ERR_INVALID, "read error on %s", path);
  *text = std::move(s);
  return BMFP_OK;
}

class Lines
{
  const std::string &t_;
  size_t pos_ = 0;
  long line_ = 0;

public:
  explicit Lines(const std::string &t) : t_(t) {}
  long line() const { return line_; }
  // next line that is neither empty nor a comment (the banner is read first)
  bool next(const char **b, const char **e, bool keep_comments = false)
  {
    while (pos_ < t_.size()) {
      const size_t end = t_.find('\n', pos_);
      const size_t stop = end == std::string::npos ? t_.size() : end;
      const char *lb = t_.data() + pos_, *le = t_.data() + stop;
      pos_ = stop + 1;
      line_++;
      while (lb < le && (*lb == ' ' || *lb == '\t' || *lb == '\r')) lb++;
      while (le > lb && (le[-1] == ' ' || le[-1] == '\t' || le[-1] == '\r')) le--;
      if (lb == le || (!keep_comments && *lb == '%')) continue;
      *b = lb;
      *e = le;
      return true;
    }
    return false;
  }
};

bool parse_long(const char *&p, const char *e, long long *v)
{
  while (p < e && (*p == ' ' || *p == '\t')) p++;
  if (p == e) return false;
  char *end = nullptr;
  errno = 0;
  const long long x = std::strtoll(p, &end, 10);
  if (end == p || errno) return false;
  p = end;
  *v = x;
  return true;
}

bool parse_double(const char *&p, const char *e, double *v)
{
  while (p < e && (*p == ' ' || *p == '\t')) p++;
  if (p == e) return false;
  char *end = nullptr;
  const double x = std::strtod(p, &end);
  if (end == p) return false;
  p = end;
  *v = x;
  return true;
}

struct Parsed {
  int n_rows = 0, n_cols = 0;
  std::vector<int> row_ptr, cols;
  std::vector<double> vals;
};

int parse_mm(const std::string &text, const char *path, Parsed *out)
{
  Lines L(text);
  const char *b, *e;
  if (!L.next(&b, &e, true)) return fail(BMFP_ERR_INVALID, "%s: empty file", path);
  char obj[64] = {}, fmt[64] = {}, field[64] = {}, sym[64] = {};
  {
    const std::string banner(b, e);
    if (banner.rfind("%%MatrixMarket", 0) != 0 ||
        std::sscanf(banner.c_str() + 14, "%63s %63s %63s %63s", obj, fmt, field, sym) != 4)
      return fail(BMFP_ERR_INVALID, "%s: no %%%%MatrixMarket banner", path);
  }
  const std::string o = lower(obj), f = lower(fmt), fl = lower(field), s = lower(sym);
  if (o != "matrix") return fail(BMFP_ERR_UNSUPPORTED, "%s: object '%s' is not a matrix", path, obj);
  const bool coord = f == "coordinate";
  if (!coord && f != "array") return fail(BMFP_ERR_INVALID, "%s: format '%s'", path, fmt);
  const bool pattern = fl == "pattern";
  if (!pattern && fl != "real" && fl != "integer" && fl != "double")
    return fail(BMFP_ERR_UNSUPPORTED, "%s: field '%s' (real, integer and pattern are supported)", path, field);
  const bool symmetric = s == "symmetric", skew = s == "skew-symmetric";
  if (!symmetric && !skew && s != "general")
    return fail(BMFP_ERR_UNSUPPORTED, "%s: symmetry '%s' (general, symmetric, skew-symmetric)", path, sym);
  if (!coord && (symmetric || skew || pattern))
    return fail(BMFP_ERR_UNSUPPORTED, "%s: array matrices must be general real or integer", path);

  if (!L.next(&b, &e)) return fail(BMFP_ERR_INVALID, "%s: missing size line", path);
  long long M = 0, N = 0, NZ = 0;
  {
    const char *p = b;
    if (!parse_long(p, e, &M) || !parse_long(p, e, &N) || (coord && !parse_long(p, e, &NZ)))
      return fail(BMFP_ERR_INVALID, "%s:%ld: bad size line", path, L.line());
  }
  if (M < 0 || N < 0 || NZ < 0 || M > INT_MAX - 1 || N > INT_MAX)
    return fail(BMFP_ERR_INVALID, "%s: size %lld x %lld out of range", path, M, N);
  if ((symmetric || skew) && M != N) return fail(BMFP_ERR_INVALID, "%s: symmetric matrix is not square", path);

  std::vector<Entry> ent;
  if (coord) {
    // an entry line takes at least 4 bytes ("i j\n"), so the file bounds the
    // entries it can hold whatever its size line declares
    const size_t fit = std::min<size_t>((size_t) NZ, text.size() / 4 + 1);
    ent.reserve((symmetric || skew) ? 2 * fit : fit);
    for (long long k = 0; k < NZ; k++) {
      if (!L.next(&b, &e)) return fail(BMFP_ERR_INVALID, "%s: %lld of %lld entries present", path, k, NZ);
      const char *p = b;
      long long i = 0, j = 0;
      double v = 1.0;
      if (!parse_long(p, e, &i) || !parse_long(p, e, &j) || (!pattern && !parse_double(p, e, &v)))
        return fail(BMFP_ERR_INVALID, "%s:%ld: bad entry", path, L.line());
      if (i < 1 || i > M || j < 1 || j > N)
        return fail(BMFP_ERR_INVALID, "%s:%ld: index (%lld, %lld) outside %lld x %lld", path, L.line(), i, j, M, N);
      ent.push_back({(int) (i - 1), (int) (j - 1), v});
      if ((symmetric || skew) && i != j) ent.push_back({(int) (j - 1), (int) (i - 1), skew ? -v : v});
    }
  } else {
    // dense, column by column; explicit zeros are not stored
    for (long long j = 0; j < N; j++)
      for (long long i = 0; i < M; i++) {
        if (!L.next(&b, &e)) return fail(BMFP_ERR_INVALID, "%s: array ends early", path);
        const char *p = b;
        double v = 0;
        if (!parse_double(p, e, &v)) return fail(BMFP_ERR_INVALID, "%s:%ld: bad value", path, L.line());
        if (v != 0.0) ent.push_back({(int) i, (int) j, v});
      }
  }
  if (ent.size() > (size_t) INT_MAX) return fail(BMFP_ERR_INVALID, "%s: too many nonzeros for int32", path);

  // CSR: rows in order, columns ascending within a row, duplicates summed
  // (in file order, the mirrored half of a symmetric entry after it)
  std::stable_sort(ent.begin(), ent.end(),
                   [](const Entry &a, const Entry &c) { return a.row != c.row ? a.row < c.row : a.col < c.col; });
  out->n_rows = (int) M;
  out->n_cols = (int) N;
  out->row_ptr.assign((size_t) M + 1, 0);
  out->cols.clear();
  out->vals.clear();
  for (size_t k = 0; k < ent.size();) {
    size_t q = k + 1;
    double v = ent[k].value;
    while (q < ent.size() && ent[q].row == ent[k].row && ent[q].col == ent[k].col) v += ent[q++].value;
    out->cols.push_back(ent[k].col);
    out->vals.push_back(v);
    out->row_ptr[(size_t) ent[k].row + 1]++;
    k = q;
  }
  for (size_t r = 0; r < (size_t) M; r++) out->row_ptr[r + 1] += out->row_ptr[r];
  return BMFP_OK;
}

int mm_read(const char *path, int *n_rows, int *n_cols, int *nnz, int *row_ptr, int *columns, double *values)
{
  if (!path || !n_rows || !n_cols || !nnz) return fail(BMFP_ERR_INVALID, "NULL argument");
  std::string text;
  BMFP_TRY(slurp(path, &text));
  Parsed P;
  BMFP_TRY(parse_mm(text, path, &P));
  const int count = (int) P.cols.size();
  if (row_ptr) {
    if (*n_rows != P.n_rows || *nnz != count)
      return fail(BMFP_ERR_INVALID, "%s: arrays sized for %d rows / %d nonzeros, file has %d / %d", path, *n_rows,
                  *nnz, P.n_rows, count);
    if (count && (!columns || !values)) return fail(BMFP_ERR_INVALID, "NULL columns or values");
    std::copy(P.row_ptr.begin(), P.row_ptr.end(), row_ptr);
    std::copy(P.cols.begin(), P.cols.end(), columns);
    std::copy(P.vals.begin(), P.vals.end(), values);
  }
  *n_rows = P.n_rows;
  *n_cols = P.n_cols;
  *nnz = count;
  return BMFP_OK;
}

int mm_write(const char *path, int n_rows, int n_cols, const int *row_ptr, const int *columns, const double *values)
{
  if (!path || !row_ptr || n_rows < 0 || n_cols < 0) return fail(BMFP_ERR_INVALID, "bad argument");
  const int nnz = row_ptr[n_rows];
  if (nnz && (!columns || !values)) return fail(BMFP_ERR_INVALID, "NULL columns or values");
  FILE *f = std::fopen(path, "w");
  if (!f) return fail(BMFP_ERR_INVALID, "cannot open %s: %s", path, std::strerror(errno));
  std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%d %d %d\n", n_rows, n_cols, nnz);
  for (int r = 0; r < n_rows; r++)
    for (int k = row_ptr[r]; k < row_ptr[r + 1]; k++) std::fprintf(f, "%d %d %10.20g\n", r + 1, columns[k] + 1, values[k]);
  const bool bad = std::ferror(f) != 0;
  if (std::fclose(f) != 0 || bad) return fail(BMFP_ERR_INVALID, "write error on %s", path);
  return BMFP_OK;
}

int block_size_detect(int n_rows, int n_cols, const int *row_ptr, const int *columns, int value_bytes,
                      const int *candidates, int n_candidates, int *bs_out, int *nnzb_out, double *fill_out)
{
  if (n_rows < 0 || n_cols < 0 || !row_ptr || !candidates || n_candidates < 1 || !bs_out || value_bytes < 1)
    return fail(BMFP_ERR_INVALID, "bad argument");
  const int nnz = row_ptr[n_rows];
  if (nnz && !columns) return fail(BMFP_ERR_INVALID, "columns is NULL");
  int best = 0;
  long long best_nnzb = 0;
  double best_bytes = 0;
  std::vector<int> bc;
  for (int c = 0; c < n_candidates; c++) {
    const int bs = candidates[c];
    if (bs < 1) return fail(BMFP_ERR_INVALID, "candidate block size %d", bs);
    if (n_rows % bs || n_cols % bs) continue;
    long long nnzb = 0;
    for (int br = 0; br < n_rows / bs; br++) {
      bc.clear();
      for (int k = row_ptr[br * bs]; k < row_ptr[(br + 1) * bs]; k++) bc.push_back(columns[k] / bs);
      std::sort(bc.begin(), bc.end());
      nnzb += std::unique(bc.begin(), bc.end()) - bc.begin();
    }
    // compulsory bytes of one BCSR SpMV: values, block columns, row pointers, x and y
    const double bytes = (double) nnzb * bs * bs * value_bytes + 4.0 * nnzb + 4.0 * (n_rows / bs + 1) +
                         (double) value_bytes * ((double) n_rows + n_cols);
    if (!best || bytes < best_bytes) {
      best = bs;
      best_nnzb = nnzb;
      best_bytes = bytes;
    }
  }
  if (!best) return fail(BMFP_ERR_INVALID, "no candidate block size divides %d x %d", n_rows, n_cols);
  if (best_nnzb * best * best > INT_MAX) return fail(BMFP_ERR_INVALID, "blocked matrix too large for int32");
  *bs_out = best;
  if (nnzb_out) *nnzb_out = (int) best_nnzb;
  if (fill_out) *fill_out = nnz ? (double) best_nnzb * best * best / nnz : 1.0;
  return BMFP_OK;
}

}  // namespace

extern "C" int bmfp_mm_read(const char *path, int *n_rows, int *n_cols, int *nnz, int *row_ptr, int *columns,
                            double *values)
{
  return bmfp::guarded("bmfp_mm_read", [&] { return mm_read(path, n_rows, n_cols, nnz, row_ptr, columns, values); });
}

extern "C" int bmfp_mm_write(const char *path, int n_rows, int n_cols, const int *row_ptr, const int *columns,
                             const double *values)
{
  return bmfp::guarded("bmfp_mm_write", [&] { return mm_write(path, n_rows, n_cols, row_ptr, columns, values); });
}

extern "C" int bmfp_block_size_detect(int n_rows, int n_cols, const int *row_ptr, const int *columns, int value_bytes,
                                      const int *candidates, int n_candidates, int *bs_out, int *nnzb_out,
                                      double *fill_out)
{
  return bmfp::guarded("bmfp_block_size_detect", [&] {
    return block_size_detect(n_rows, n_cols, row_ptr, columns, value_bytes, candidates, n_candidates, bs_out,
                             nnzb_out, fill_out);
  });
}
