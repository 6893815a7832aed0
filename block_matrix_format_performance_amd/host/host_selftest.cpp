// host_selftest.cpp -- exercises the C++ drop-in host layer (no GPU) and
// prints JSON in the layout of tests/golden/ref_host.json, so the test suite
// can compare it field by field with the reference's own host code.
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include <algorithm>
#include <cstdlib>
#include <unistd.h>
#include "bmfp/matrix_converters.h"
#include "bmfp/matrix_market.h"
#include "bmfp/measurement_class.h"
#include "bmfp/truss_bridge.h"

template <class T>
static void put(const char *name, const T *a, long n, bool comma = true)
{
  std::printf("\"%s\": [", name);
  for (long i = 0; i < n; i++) {
    if (i) std::printf(", ");
    if constexpr (std::is_floating_point<T>::value) std::printf("%.17g", (double) a[i]);
    else std::printf("%lld", (long long) a[i]);
  }
  std::printf("]%s\n", comma ? "," : "");
}

static std::string capture_compare(unsigned n, const double *a, const double *b)
{
  std::stringstream ss;
  std::streambuf *old = std::cerr.rdbuf(ss.rdbuf());
  compare_results<double>(n, a, b);
  std::cerr.rdbuf(old);
  std::string out;
  for (char ch : ss.str()) out += ch == '\n' ? '|' : ch;
  return out;
}

// Text of a file as a JSON string body.
static std::string file_json(const std::string &path)
{
  std::string out;
  FILE *f = std::fopen(path.c_str(), "rb");
  int c;
  while (f && (c = std::fgetc(f)) != EOF) {
    if (c == '\n') out += "\\n";
    else if (c == '"' || c == '\\') (out += '\\') += (char) c;
    else out += (char) c;
  }
  if (f) std::fclose(f);
  return out;
}

// Matrix Market text of csr through matrix_converters.h's write_mm and through
// the C ABI, and whether reading it back restores the matrix exactly.
static void put_mm(const csr_matrix_class<double, int> &csr)
{
  char tmpl[] = "/tmp/bmfp_mm_XXXXXX";
  const int fd = mkstemp(tmpl);
  if (fd < 0) std::exit(1);
  close(fd);
  const std::string path = tmpl;
  csr.write_mm(path);
  std::printf("\"mm_text\": \"%s\",\n", file_json(path).c_str());
  write_matrix_market(csr, path);
  std::printf("\"mm_text_abi\": \"%s\",\n", file_json(path).c_str());
  auto back = read_matrix_market<double, int>(path);
  bool same = back->n_rows == csr.n_rows && back->n_cols == csr.n_cols && back->nnz == csr.nnz &&
              std::equal(csr.row_ptr.get(), csr.row_ptr.get() + csr.n_rows + 1, back->row_ptr.get()) &&
              std::equal(csr.columns.get(), csr.columns.get() + csr.nnz, back->columns.get()) &&
              std::equal(csr.values.get(), csr.values.get() + csr.nnz, back->values.get());
  std::printf("\"mm_readback_exact\": %s,\n", same ? "true" : "false");
  std::remove(path.c_str());
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
    auto back = csr_to_bcsr<double, int>(csr, bs);
    std::vector<double> lookup;
    for (int row = 0; row < N; row++) lookup.push_back(*m->get_block_data_by_column(row, m->columns[m->row_ptr[row] + bpr / 2]));
    std::printf("{\"n_rows\": %d, \"blocks_per_row\": %d, \"bs\": %d, \"nnzb\": %d,\n", N, bpr, bs, m->nnzb);
    put("bcsr_row_ptr", m->row_ptr.get(), N + 1);
    put("bcsr_columns", m->columns.get(), m->nnzb);
    put("bcsr_values", m->values.get(), m->size());
    put("bcsr_values_transposed", tr.data(), m->size());
    put("csr_row_ptr", csr.row_ptr.get(), csr.n_rows + 1);
    put("csr_columns", csr.columns.get(), csr.nnz);
    put("block_lookup_mid", lookup.data(), (long) lookup.size());
    put_mm(csr);
    put("roundtrip_row_ptr", back->row_ptr.get(), back->n_rows + 1);
    put("roundtrip_columns", back->columns.get(), back->nnzb);
    put("roundtrip_values", back->values.get(), back->size());
    put("csr_values", csr.values.get(), csr.nnz, false);
    std::printf("}%s\n", ci + 1 < ncases ? "," : "");
  }
  std::printf("],\n");
  measurement_class acc;
  measurement_class a("fmt_a", 2.0, 8e9, 4e9);
  measurement_class b("fmt_b", 0.5, 1e9, 3e9);
  acc += a;
  acc += b;
  std::printf("\"measurement\": {\"a\": [%.17g, %.17g, %.17g], \"b\": [%.17g, %.17g, %.17g], ", a.get_elapsed(),
              a.get_effective_bandwidth(), a.get_computational_throughput(), b.get_elapsed(),
              b.get_effective_bandwidth(), b.get_computational_throughput());
  std::printf("\"sum\": [%.17g, %.17g, %.17g, \"%s\"], ", acc.get_elapsed(), acc.get_effective_bandwidth(),
              acc.get_computational_throughput(), acc.get_format().c_str());
  acc.finalize();
  std::printf("\"finalized\": [%.17g, %.17g, %.17g]},\n", acc.get_elapsed(), acc.get_effective_bandwidth(),
              acc.get_computational_throughput());
  const double r1[] = {1.0, 2.0, 3.0, 4.0};
  const double y_same[] = {1.0, 2.0, 3.0, 4.0};
  const double y_bad[] = {1.0, 2.5, 3.0, 4.0};
  const double y_tiny[] = {1.0, 2.0 + 1e-6, 3.0, 4.0};
  std::printf("\"compare\": {\"same\": \"%s\", \"bad\": \"%s\", \"tiny\": \"%s\"},\n",
              capture_compare(4, y_same, r1).c_str(), capture_compare(4, y_bad, r1).c_str(),
              capture_compare(4, y_tiny, r1).c_str());
  // a small suspension-bridge truss: geometry, supports, K and f
  {
    auto load = [](double x) { return std::pair<double, double>(x > 80 && x < 100 ? 3e5 : 0.0, x > 80 && x < 100 ? -1e6 : 0.0); };
    golden_gate_bridge_2d<double, int> br(load, 100.0, 40.0, 100.0, 10.0);
    const int nn = br.nodes_count(), ne = br.elements_count();
    std::vector<double> xy, ea;
    std::vector<int> ends, fixed;
    for (int i = 0; i < nn; i++) {
      xy.push_back(br.node_x(i));
      xy.push_back(br.node_y(i));
    }
    for (int e = 0; e < ne; e++) {
      ends.push_back(br.element_nodes(e).first);
      ends.push_back(br.element_nodes(e).second);
      ea.push_back(br.element_ea(e));
    }
    for (int d = 0; d < 2 * nn; d++) fixed.push_back(br.dof_fixed(d) ? 1 : 0);
    const auto &K = *br.matrix;
    std::printf("\"bridge\": {\"nodes\": %d,\n", nn);
    put("xy", xy.data(), (long) xy.size());
    put("ends", ends.data(), (long) ends.size());
    put("ea", ea.data(), (long) ea.size());
    put("fixed", fixed.data(), (long) fixed.size());
    put("row_ptr", K.row_ptr.get(), K.n_rows + 1);
    put("columns", K.columns.get(), K.nnzb);
    put("values", K.values.get(), K.size());
    put("rhs", br.forces_rhs.get(), 2L * nn, false);
    std::printf("},\n");
  }
  // 9-point stencil generated through the library (for the bitwise check)
  auto st = gen_stencil9_bcsr<double, int>(5, 4, 3, 1, 7);
  std::printf("\"stencil\": {\"nx\": 5, \"ny\": 4, \"bs\": 3, \"kind\": 1, \"seed\": 7,\n");
  put("row_ptr", st->row_ptr.get(), st->n_rows + 1);
  put("columns", st->columns.get(), st->nnzb);
  put("values", st->values.get(), st->size(), false);
  std::printf("}\n}\n");
  return 0;
}
