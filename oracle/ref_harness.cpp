// ref_harness.cpp -- drives the reference's own host code, compiled from its
// sources where they lie (common/matrix_converters.h, common/measurement_class.cpp),
// and prints JSON golden vectors to stdout.  TEST INFRASTRUCTURE ONLY: built
// by oracle/Makefile into oracle/_ref/ (never committed, never shipped) and
// run only by tests/golden/make_golden.py and the oracle tests.
//
// What it exercises:
//   gen_n_diag_bcsr            matrix_converters.h:157-196
//   csr_matrix_class(bcsr)     matrix_converters.h:88-117
//   transpose_blocks           matrix_converters.h:32-49
//   get_block_data_by_column   matrix_converters.h:61-70
//   measurement_class          measurement_class.h:11-58, measurement_class.cpp:7-16
//   compare_results            measurement_class.h:60-89 (its stderr text)
//   csr_matrix_class::write_mm matrix_converters.h:119-144 (through common/mmio.c)
#include <string>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <sstream>
#include <iostream>
#include <unistd.h>

#include "matrix_converters.h"
#include "measurement_class.h"

static void put_ints(const char *name, const int *a, long n, bool comma = true)
{
  std::printf("\"%s\": [", name);
  for (long i = 0; i < n; i++)
    std::printf("%s%d", i ? ", " : "", a[i]);
  std::printf("]%s\n", comma ? "," : "");
}

static void put_doubles(const char *name, const double *a, long n, bool comma = true)
{
  std::printf("\"%s\": [", name);
  for (long i = 0; i < n; i++)
    std::printf("%s%.17g", i ? ", " : "", a[i]);
  std::printf("]%s\n", comma ? "," : "");
}

// compare_results writes to std::cerr; capture it.
static std::string capture_compare(unsigned n, const double *a, const double *b)
{
  std::stringstream ss;
  std::streambuf *old = std::cerr.rdbuf(ss.rdbuf());
  compare_results<double>(n, a, b);
  std::cerr.rdbuf(old);
  std::string s = ss.str();
  std::string out;
  for (char ch : s)
    out += (ch == '\n') ? '|' : ch;
  return out;
}

int main()
{
  std::printf("{\n\"matrices\": [\n");
  const int cases[][3] = {{7, 3, 1}, {7, 3, 2}, {7, 3, 3}, {7, 3, 4}, {7, 3, 6}, {9, 6, 2}, {5, 5, 3}};
  const int ncases = sizeof(cases) / sizeof(cases[0]);
  for (int ci = 0; ci < ncases; ci++) {
    const int N = cases[ci][0], bpr = cases[ci][1], bs = cases[ci][2];
    auto m = gen_n_diag_bcsr<double, int>(N, bpr, bs);
    csr_matrix_class<double, int> csr(*m);
    std::vector<double> tr(m->size());
    m->transpose_blocks(tr.data());
    std::printf("{\"n_rows\": %d, \"blocks_per_row\": %d, \"bs\": %d, \"nnzb\": %d,\n", N, bpr, bs, m->nnzb);
    put_ints("bcsr_row_ptr", m->row_ptr.get(), N + 1);
    put_ints("bcsr_columns", m->columns.get(), m->nnzb);
    put_doubles("bcsr_values", m->values.get(), m->size());
    put_doubles("bcsr_values_transposed", tr.data(), m->size());
    put_ints("csr_row_ptr", csr.row_ptr.get(), csr.n_rows + 1);
    put_ints("csr_columns", csr.columns.get(), csr.nnz);
    // block lookup by column: first value of the block at column row (clamped band)
    std::vector<double> lookup;
    for (int row = 0; row < N; row++) {
      const int col = m->columns[m->row_ptr[row] + bpr / 2];
      lookup.push_back(*m->get_block_data_by_column(row, col));
    }
    put_doubles("block_lookup_mid", lookup.data(), (long) lookup.size());
    {
      // the Matrix Market text write_mm produces for this CSR matrix
      char path[] = "/tmp/ref_mm_XXXXXX";
      const int fd = mkstemp(path);
      if (fd < 0) return 1;
      close(fd);
      csr.write_mm(path);
      FILE *f = std::fopen(path, "rb");
      std::string text;
      int c;
      while (f && (c = std::fgetc(f)) != EOF) text += (char) c;
      if (f) std::fclose(f);
      std::remove(path);
      std::printf("\"mm_text\": \"");
      for (char ch : text) {
        if (ch == '\n') std::printf("\\n");
        else if (ch == '"' || ch == '\\') std::printf("\\%c", ch);
        else std::putchar(ch);
      }
      std::printf("\",\n");
    }
    put_doubles("csr_values", csr.values.get(), csr.nnz, false);
    std::printf("}%s\n", ci + 1 < ncases ? "," : "");
  }
  std::printf("],\n");

  // measurement_class arithmetic
  measurement_class acc;
  measurement_class a("fmt_a", 2.0, 8e9, 4e9);
  measurement_class b("fmt_b", 0.5, 1e9, 3e9);
  acc += a;
  acc += b;
  std::printf("\"measurement\": {\"a\": [%.17g, %.17g, %.17g], \"b\": [%.17g, %.17g, %.17g], ",
              a.get_elapsed(), a.get_effective_bandwidth(), a.get_computational_throughput(),
              b.get_elapsed(), b.get_effective_bandwidth(), b.get_computational_throughput());
  std::printf("\"sum\": [%.17g, %.17g, %.17g, \"%s\"], ", acc.get_elapsed(), acc.get_effective_bandwidth(),
              acc.get_computational_throughput(), acc.get_format().c_str());
  acc.finalize();
  std::printf("\"finalized\": [%.17g, %.17g, %.17g]},\n", acc.get_elapsed(), acc.get_effective_bandwidth(),
              acc.get_computational_throughput());

  // compare_results stderr text for a few vector pairs
  const double r1[] = {1.0, 2.0, 3.0, 4.0};
  const double y_same[] = {1.0, 2.0, 3.0, 4.0};
  const double y_bad[] = {1.0, 2.5, 3.0, 4.0};
  const double y_tiny[] = {1.0, 2.0 + 1e-6, 3.0, 4.0};
  std::printf("\"compare\": {\"same\": \"%s\", \"bad\": \"%s\", \"tiny\": \"%s\"}\n",
              capture_compare(4, y_same, r1).c_str(), capture_compare(4, y_bad, r1).c_str(),
              capture_compare(4, y_tiny, r1).c_str());
  std::printf("}\n");
  return 0;
}
