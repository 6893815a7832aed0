// main.cpp -- benchmark driver, drop-in for the reference's main.cpp:
// generates a matrix, times the single-threaded CPU CSR SpMV and every GPU
// SpMV family, prints a speed-up table and writes result.json
// ({ "<bs>": { "<format>": mean_seconds } }).  Without arguments it runs the
// reference's published workload (FP32, N = 50,000 block rows, 6 blocks per
// row, bs in {2,4,8,16,32}, x = 1; main.cpp:386-398).
//
// Options (all optional):
//   --type float|double        value type (default float)
//   --bs 2,4,8,16,32           block sizes
//   --matrix ndiag|stencil     gen_n_diag_bcsr (default) or the seeded
//                              9-point block stencil (kind 0, seed 0)
//   --n-rows N --bpr K         ndiag shape (default 50000, 6)
//   --values V                 stencil: grid side round(sqrt(V/(9 bs^2)))
//   --x ones|random            x = 1 (reference) or uniform[0,1] (seed 0)
//   --reps R                   repetitions averaged (default 10)
//   --out FILE                 JSON output (default result.json)
//   --out-gbs FILE             also each row's effective bandwidth (GB/s of
//                              its compulsory bytes), same layout
//   --mm FILE                  a Matrix Market file instead of a generated
//                              matrix: CSR rows on it as read, BCSR rows at
//                              the detected block size (or each --bs that
//                              divides its dimensions)
//   --bridge                   the suspension-bridge truss (bs = 2) instead of
//                              a generated matrix: every SpMV row on it
//   --bridge-solve             build the bridge, write matrix.mtx and
//                              output_1.vtk, solve K u = f with the Jacobi-
//                              preconditioned FP64 BiCGStab and write
//                              output_2.vtk (displaced), bridge_rhs.txt and
//                              bridge_solution.txt (main.cpp:349-381)
//   --bridge-span M            main span in metres (default 1280; sides 345)
//   --bridge-eps E             relative residual target (default 1e-8: the
//                              slender truss is ill-conditioned, and its
//                              true residual stalls near 5e-9 in FP64)
//   --solve                    also run the FP64 BiCGStab of the 256x256
//                              bs=2 stencil (b = A*1, x0 = 0, eps 1e-8) and
//                              report its iterations, error and time
//   --solve-csr                solve through the reference caller's exact
//                              signature (main.cpp:372-373):
//                              gpu_bicgstab<T, C>(csr, precond) and
//                              solve(csr, b, eps, max_it), with T = --type and
//                              C = --index int|unsigned, on the side x side
//                              9-point stencil of --solve-bs (kind 0, seed 0;
//                              default 256, 2: config 0), b = A*1, x0 = 0;
//                              --solve-eps (default 1e-8), --precond 0|1,
//                              --x-out FILE writes x as raw doubles
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <optional>
#include <sstream>
#include <string>
#include <vector>

#include "bmfp/bicgstab.h"
#include "bmfp/gpu_matrix_multiplier.h"
#include "bmfp/matrix_converters.h"
#include "bmfp/matrix_market.h"
#include "bmfp/measurement_class.h"
#include "bmfp/truss_bridge.h"

namespace {

struct Options {
  std::string type = "float";
  std::vector<int> bs = {2, 4, 8, 16, 32};
  bool bs_given = false;
  std::string mm;
  bool bridge = false, bridge_solve = false;
  double bridge_span = 1280.0, bridge_eps = 1e-8;
  std::string matrix = "ndiag";
  int n_rows = 50000;
  int bpr = 6;
  double values = 4.4e7;
  std::string x = "ones";
  int reps = 10;
  std::string out = "result.json";
  std::string out_gbs;
  bool solve = false;
  bool solve_csr = false;
  std::string index = "int";
  int solve_side = 256, solve_bs = 2, precond = 0;
  double solve_eps = 1e-8;
  std::string x_out;
};

// main.cpp:21-63 with x as an input (the reference overwrites x with 1):
// y_row = sum over the row, in CSR order, of a*x[col]; bytes counted the
// reference's way (x per nonzero, main.cpp:50-56).
template <typename T, typename I>
measurement_class cpu_csr_spmv_single_thread_naive(const csr_matrix_class<T, I> &m, const T *x, T *y)
{
  std::fill_n(y, m.n_rows, T(0));
  const I *rp = m.row_ptr.get();
  const I *cols = m.columns.get();
  const T *vals = m.values.get();
  const auto begin = std::chrono::steady_clock::now();
  for (I row = 0; row < m.n_rows; row++) {
    T dot = 0;
    for (I e = rp[row]; e < rp[row + 1]; e++) dot += vals[e] * x[cols[e]];
    y[row] = dot;
  }
  const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - begin).count();
  const double bytes = (double) m.nnz * sizeof(T) * 2 + (double) m.nnz * sizeof(I) + 2.0 * m.n_rows * sizeof(I) +
                       (double) m.n_rows * sizeof(T);
  return measurement_class("CPU CSR", elapsed, bytes, 2.0 * m.nnz);
}

// main.cpp:70-113 without colours.
class time_printer
{
  double reference;

public:
  explicit time_printer(double reference_time) : reference(reference_time) {}
  void print_time(const measurement_class &m) const
  {
    std::cout << "\t" << std::left << std::setw(64) << m.get_format() << ":  " << std::setw(14)
              << std::setprecision(6) << m.get_elapsed() << std::setw(10) << reference / m.get_elapsed()
              << std::setw(10) << m.get_effective_bandwidth() << " GB/s" << std::endl;
  }
};

// Per format: mean seconds and effective bandwidth (GB/s).
struct Rows {
  std::map<std::string, double> seconds, gbs;
  void add(const measurement_class &m)
  {
    seconds[m.get_format()] = m.get_elapsed();
    gbs[m.get_format()] = m.get_effective_bandwidth();
  }
};

template <typename T, typename I>
Rows perform_measurements(csr_matrix_class<T, I> &matrix, bcsr_matrix_class<T, I> &block_matrix, const Options &o)
{
  Rows results;
  const size_t n = (size_t) matrix.n_rows;
  std::vector<T> x(n), reference(n);
  if (o.x == "random") {
    std::vector<double> xd(n);
    // uniform[0,1) from the library's counter-based generator, stream 4, seed 0
    void *d = nullptr;
    bmfp_malloc(&d, n * sizeof(double));
    bmfp_fill(d, BMFP_F64, n, 0, 0, 4, 0.0, nullptr);
    bmfp_memcpy_d2h(xd.data(), d, n * sizeof(double));
    bmfp_free(d);
    std::copy(xd.begin(), xd.end(), x.begin());
  } else {
    std::fill(x.begin(), x.end(), T(1));
  }
  const T *xp = o.x == "random" ? x.data() : nullptr;

  auto measure_multiple_times = [&](const std::function<measurement_class()> &action) {
    measurement_class acc;
    for (int i = 0; i < o.reps; i++) acc += action();
    acc.finalize();
    results.add(acc);
    return acc;
  };

  auto cpu = measure_multiple_times(
      [&] { return cpu_csr_spmv_single_thread_naive(matrix, x.data(), reference.data()); });
  time_printer printer(cpu.get_elapsed());
  printer.print_time(cpu);
  printer.print_time(measure_multiple_times(
      [&] { return gpu_csr_spmv<T, I>(matrix, xp, reference.data(), nullptr); }));
  printer.print_time(measure_multiple_times(
      [&] { return gpu_csr_vector_spmv<T, I>(matrix, xp, reference.data(), nullptr); }));
  printer.print_time(measure_multiple_times(
      [&] { return gpu_csr_merge_spmv<T, I>(matrix, xp, reference.data(), nullptr); }));

  std::vector<T> transposed((size_t) block_matrix.size());
  block_matrix.transpose_blocks(transposed.data());
  {
    // "jit" (main.cpp:170-294): one timed launch of the run-time-specialised kernel
    auto jit = gpu_bcsr_jit_spmv<T, I>(block_matrix, transposed.data(), xp, reference.data());
    results.add(jit);
    printer.print_time(jit);
  }
  // o.reps calls, averaged by a default-constructed accumulator per format
  // (the reference seeds the accumulators with an 11th call, main.cpp:296-307).
  std::vector<measurement_class> acc;
  for (int i = 0; i < o.reps; i++) {
    auto r = gpu_bcsr_spmv<T, I>(block_matrix, transposed.data(), xp, reference.data());
    if (acc.empty()) acc.resize(r.size());
    for (size_t k = 0; k < r.size(); k++) acc[k] += r[k];
  }
  for (auto &m : acc) {
    m.finalize();
    results.add(m);
    printer.print_time(m);
  }
  return results;
}

std::string json_escape(const std::string &s)
{
  std::string o;
  for (char c : s) {
    if (c == '"' || c == '\\') o += '\\';
    o += c;
  }
  return o;
}

std::string json_number(double v)
{
  if (!std::isfinite(v)) return "null";
  std::ostringstream ss;
  ss << std::setprecision(17) << v;
  return ss.str();
}

template <typename T>
Rows run_bs(int bs, const Options &o)
{
  std::unique_ptr<bcsr_matrix_class<T, int>> block_matrix;
  if (o.matrix == "stencil") {
    const int side = (int) std::lround(std::sqrt(o.values / (9.0 * bs * bs)));
    std::cout << "\nBS: " << bs << " (9-point stencil " << side << "x" << side << ")\n";
    block_matrix = gen_stencil9_bcsr<T, int>(side, side, bs, 0, 0);
  } else {
    std::cout << "\nBS: " << bs << "\n";
    block_matrix = gen_n_diag_bcsr<T, int>(o.n_rows, o.bpr, bs);
  }
  csr_matrix_class<T, int> matrix(*block_matrix);
  return perform_measurements(matrix, *block_matrix, o);
}

// Matrix Market input: result maps keyed by block size.
template <typename T>
std::map<int, Rows> run_mm(const Options &o)
{
  auto csr = read_matrix_market<T, int>(o.mm);
  std::cout << o.mm << ": " << csr->n_rows << " x " << csr->n_cols << ", " << csr->nnz << " nonzeros\n";
  std::vector<int> sizes;
  if (o.bs_given) {
    for (int bs : o.bs)
      if (bs >= 1 && csr->n_rows % bs == 0 && csr->n_cols % bs == 0) sizes.push_back(bs);
      else std::cout << "bs " << bs << " does not divide the matrix; skipped\n";
  } else {
    const block_size_choice c = detect_block_size(*csr);
    std::cout << "detected block size " << c.bs << " (" << c.nnzb << " blocks, fill " << c.fill << ")\n";
    sizes.push_back(c.bs);
  }
  std::map<int, Rows> out;
  for (int bs : sizes) {
    std::cout << "\nBS: " << bs << "\n";
    auto block_matrix = csr_to_bcsr<T, int>(*csr, bs);
    out[bs] = perform_measurements(*csr, *block_matrix, o);
  }
  return out;
}

// The bridge of main.cpp:349-381: towers 260 m high, 7.62 m panels, 2 MN
// downward on every deck node within 450 m of mid-span.
golden_gate_bridge_2d<double, int> make_bridge(const Options &o)
{
  const double side = 345.0, main_span = o.bridge_span;
  auto load = [=](double x) {
    const double mid = (main_span + 2 * side) / 2, window = 450;
    return std::pair<double, double>(0.0, x > mid - window && x < mid + window ? -2000000.0 : 0.0);
  };
  return golden_gate_bridge_2d<double, int>(load, main_span, side, 260.0, 7.62);
}

int run_bridge_solve(const Options &o)
{
  auto bridge = make_bridge(o);
  const bcsr_matrix_class<double, int> &K = *bridge.matrix;
  const size_t n = (size_t) K.n_rows * K.bs;
  std::cout << "bridge: " << bridge.nodes_count() << " nodes, " << bridge.elements_count() << " bars, n = " << n
            << ", " << K.nnzb << " blocks\n";
  csr_matrix_class<double, int>(K).write_mm("matrix.mtx");
  bridge.write_vtk("output_1.vtk");
  gpu_bicgstab<double, int> solver(K, true);
  const double *u = solver.solve(K, bridge.forces_rhs.get(), o.bridge_eps, 200000);
  const bmfp_solve_info &info = solver.last_info();
  bridge.write_vtk("output_2.vtk", u);
  std::ofstream rhs("bridge_rhs.txt"), sol("bridge_solution.txt");
  rhs << std::setprecision(17);
  sol << std::setprecision(17);
  double umax = 0;
  for (size_t i = 0; i < n; i++) {
    rhs << bridge.forces_rhs[i] << "\n";
    sol << u[i] << "\n";
    umax = std::max(umax, std::fabs(u[i]));
  }
  std::cout << "Jacobi BiCGStab: status " << info.status << ", iterations " << info.iterations << ", ||r||/||b|| "
            << info.rel_residual << ", true " << info.true_rel_residual << ", " << info.seconds * 1e3
            << " ms; max |u| " << umax << " m" << std::endl;
  return info.status == BMFP_OK ? 0 : 1;
}

int run_solve()
{
  // 256x256 node grid, bs = 2, diagonally dominant, b = A*1, x0 = 0, eps 1e-8
  auto A = gen_stencil9_bcsr<double, int>(256, 256, 2, 0, 0);
  const size_t n = (size_t) A->n_rows * A->bs;
  csr_matrix_class<double, int> csr(*A);
  std::vector<double> ones(n, 1.0), b(n);
  cpu_csr_spmv_single_thread_naive(csr, ones.data(), b.data());
  gpu_bicgstab<double, int> solver(*A, false);
  const double *x = solver.solve(*A, b.data(), 1e-8, 10000);
  double err = 0;
  for (size_t i = 0; i < n; i++) err = std::max(err, std::fabs(x[i] - 1.0));
  const bmfp_solve_info &info = solver.last_info();
  std::cout << "BiCGStab 256x256 bs=2: iterations " << info.iterations << ", ||r||/||b|| " << info.rel_residual
            << ", true " << info.true_rel_residual << ", max|x-1| " << err << ", " << info.seconds * 1e3 << " ms"
            << std::endl;
  return info.status == BMFP_OK ? 0 : 1;
}

// The reference caller's solver path: CSR constructor and CSR solve
// (main.cpp:372-373, bicgstab.h:17-20) for one (T, C) instantiation.
template <class T, class C>
int run_solve_csr(const Options &o)
{
  auto A = gen_stencil9_bcsr<T, C>(o.solve_side, o.solve_side, o.solve_bs, 0, 0);
  const csr_matrix_class<T, C> csr(*A);
  const size_t n = (size_t) csr.n_rows;
  std::vector<T> b(n);
  for (size_t row = 0; row < n; row++) {
    T dot = 0;  // b = A * 1 in CSR order
    for (C e = csr.row_ptr[row]; e < csr.row_ptr[row + 1]; e++) dot += csr.values[e];
    b[row] = dot;
  }
  gpu_bicgstab<T, C> solver(csr, o.precond != 0);
  const T *x = solver.solve(csr, b.data(), (T) o.solve_eps, 100000);
  const bmfp_solve_info &info = solver.last_info();
  double err = 0;
  for (size_t i = 0; i < n; i++) err = std::max(err, std::fabs((double) x[i] - 1.0));
  std::cout << "BiCGStab CSR<" << o.type << ", " << o.index << "> " << o.solve_side << "x" << o.solve_side
            << " bs=" << o.solve_bs << ": status " << info.status << ", iterations " << info.iterations
            << ", ||r||/||b|| " << info.rel_residual << ", true " << info.true_rel_residual << ", max|x-1| " << err
            << std::endl;
  if (!o.x_out.empty()) {
    std::vector<double> xd(x, x + n);
    std::ofstream f(o.x_out, std::ios::binary);
    f.write(reinterpret_cast<const char *>(xd.data()), (std::streamsize) (n * sizeof(double)));
  }
  return info.status == BMFP_OK ? 0 : 1;
}

std::vector<int> parse_list(const char *s)
{
  std::vector<int> v;
  std::stringstream ss(s);
  std::string tok;
  while (std::getline(ss, tok, ',')) v.push_back(std::atoi(tok.c_str()));
  return v;
}

}  // namespace

int main(int argc, char **argv)
{
  Options o;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    auto next = [&]() -> const char * {
      if (i + 1 >= argc) {
        std::cerr << "missing value for " << a << std::endl;
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "--type") o.type = next();
    else if (a == "--bs") {
      o.bs = parse_list(next());
      o.bs_given = true;
    }
    else if (a == "--mm") o.mm = next();
    else if (a == "--bridge") o.bridge = true;
    else if (a == "--bridge-solve") o.bridge_solve = true;
    else if (a == "--bridge-span") o.bridge_span = std::atof(next());
    else if (a == "--bridge-eps") o.bridge_eps = std::atof(next());
    else if (a == "--matrix") o.matrix = next();
    else if (a == "--n-rows") o.n_rows = std::atoi(next());
    else if (a == "--bpr") o.bpr = std::atoi(next());
    else if (a == "--values") o.values = std::atof(next());
    else if (a == "--x") o.x = next();
    else if (a == "--reps") o.reps = std::atoi(next());
    else if (a == "--out") o.out = next();
    else if (a == "--out-gbs") o.out_gbs = next();
    else if (a == "--solve") o.solve = true;
    else if (a == "--solve-csr") o.solve_csr = true;
    else if (a == "--index") o.index = next();
    else if (a == "--solve-side") o.solve_side = std::atoi(next());
    else if (a == "--solve-bs") o.solve_bs = std::atoi(next());
    else if (a == "--solve-eps") o.solve_eps = std::atof(next());
    else if (a == "--precond") o.precond = std::atoi(next());
    else if (a == "--x-out") o.x_out = next();
    else {
      std::cerr << "unknown option " << a << std::endl;
      return 2;
    }
  }
  if (bmfp_init(0) != BMFP_OK) {
    std::cerr << "bmfp_init: " << bmfp_last_error() << std::endl;
    return 1;
  }
  if (o.bridge_solve) return run_bridge_solve(o);
  if (o.solve_csr) {
    try {
      const bool dbl = o.type == "double", uns = o.index == "unsigned";
      if (dbl) return uns ? run_solve_csr<double, unsigned>(o) : run_solve_csr<double, int>(o);
      return uns ? run_solve_csr<float, unsigned>(o) : run_solve_csr<float, int>(o);
    } catch (const std::exception &e) {
      std::cerr << "--solve-csr: " << e.what() << std::endl;
      return 1;
    }
  }
  std::map<int, Rows> all;
  if (o.bridge) {
    auto bridge = make_bridge(o);
    std::cout << "\nBS: 2 (suspension bridge, " << bridge.nodes_count() << " nodes)\n";
    csr_matrix_class<double, int> csr(*bridge.matrix);
    if (o.type == "double") {
      all[2] = perform_measurements(csr, *bridge.matrix, o);
    } else {
      auto b32 = csr_to_bcsr<float, int>(
          csr_matrix_class<float, int>(csr.n_rows, csr.n_cols, csr.nnz, csr.row_ptr.get(), csr.columns.get(),
                                       std::vector<float>(csr.values.get(), csr.values.get() + csr.nnz).data()),
          2);
      csr_matrix_class<float, int> c32(*b32);
      all[2] = perform_measurements(c32, *b32, o);
    }
  } else if (!o.mm.empty()) {
    try {
      all = o.type == "double" ? run_mm<double>(o) : run_mm<float>(o);
    } catch (const std::exception &e) {
      std::cerr << "--mm " << o.mm << ": " << e.what() << std::endl;
      return 1;
    }
  } else {
    for (int bs : o.bs) all[bs] = o.type == "double" ? run_bs<double>(bs, o) : run_bs<float>(bs, o);
  }

  // result.json layout { "<bs>": { "<format>": value } }
  auto write_json = [&](const std::string &path, bool gbs) {
    std::ofstream os(path);
    os << "{\n";
    size_t i = 0;
    for (const auto &kv : all) {
      const auto &rows = gbs ? kv.second.gbs : kv.second.seconds;
      os << "  \"" << kv.first << "\": {\n";
      size_t j = 0;
      for (const auto &r : rows)
        os << "    \"" << json_escape(r.first) << "\": " << json_number(r.second) << (++j < rows.size() ? ",\n" : "\n");
      os << "  }" << (++i < all.size() ? ",\n" : "\n");
    }
    os << "}\n";
  };
  write_json(o.out, false);
  if (!o.out_gbs.empty()) write_json(o.out_gbs, true);
  std::cout << "wrote " << o.out << std::endl;
  return o.solve ? run_solve() : 0;
}
