// measurement_class.h -- timing record and result comparison, drop-in for the
// reference's common/measurement_class.{h,cpp} (measurement_class.h:11-89,
// measurement_class.cpp:7-16): same constructor, accessors, += / finalize
// averaging and compare_results report.
#ifndef BMFP_MEASUREMENT_CLASS_H
#define BMFP_MEASUREMENT_CLASS_H

#include <cmath>
#include <iostream>
#include <string>

class measurement_class
{
  const size_t giga = 1000000000;

public:
  measurement_class() = default;
  // effective bandwidth = load_store_bytes / (elapsed * 1e9) [GB/s],
  // throughput = operations_count / (elapsed * 1e9) [GFLOP/s].
  measurement_class(std::string format, double elapsed, double load_store_bytes, double operations_count);

  double get_elapsed() const { return elapsed; }
  double get_effective_bandwidth() const { return effective_bandwidth; }
  double get_computational_throughput() const { return computational_throughput; }
  const std::string &get_format() const { return matrix_format; }

  // Accumulates for averaging; the format name follows the latest addend.
  measurement_class &operator+=(const measurement_class &rhs)
  {
    elapsed += rhs.get_elapsed();
    effective_bandwidth += rhs.get_effective_bandwidth();
    computational_throughput += rhs.get_computational_throughput();
    matrix_format = rhs.get_format();
    measurements_count++;
    return *this;
  }

  // Divides the sums by the number of += calls (a default-constructed
  // accumulator; see main.cpp:296-307 for the reference's off-by-one use,
  // which this build's driver does not repeat).
  void finalize()
  {
    if (measurements_count) {
      elapsed /= measurements_count;
      effective_bandwidth /= measurements_count;
      computational_throughput /= measurements_count;
    }
  }

  measurement_class &operator=(const measurement_class &o)
  {
    elapsed = o.elapsed;
    effective_bandwidth = o.effective_bandwidth;
    computational_throughput = o.computational_throughput;
    matrix_format = o.matrix_format;
    measurements_count = o.measurements_count;
    return *this;
  }
  measurement_class(const measurement_class &o) { *this = o; }

private:
  double elapsed{};
  double effective_bandwidth{};
  double computational_throughput{};
  std::string matrix_format;
  unsigned int measurements_count{};
};

// Squared relative L2 error sum (a-b)^2 / sum b^2, accumulated in double.
template <typename data_type>
double compare_results_metric(size_t y_size, const data_type *a, const data_type *b)
{
  double num = 0.0, den = 0.0;
  for (size_t i = 0; i < y_size; i++) {
    const double d = (double) a[i] - (double) b[i];
    num += d * d;
    den += (double) b[i] * (double) b[i];
  }
  return den > 0 ? num / den : (num > 0 ? INFINITY : 0.0);
}

// measurement_class.h:60-89: prints "ERROR: <metric>" and the first entry
// differing by more than 1e-8 to stderr when the metric exceeds 1e-9.
// Returns true when the results agree (the reference returns nothing).
template <typename data_type>
bool compare_results(size_t y_size, const data_type *a, const data_type *b)
{
  const double error = compare_results_metric(y_size, a, b);
  if (error > 1e-9) {
    std::cerr << "ERROR: " << error << std::endl;
    for (size_t i = 0; i < y_size; i++)
      if (std::abs((double) a[i] - (double) b[i]) > 1e-8) {
        std::cerr << "a[" << i << "] = " << a[i] << "; b[" << i << "] = " << b[i] << std::endl;
        break;
      }
  }
  std::cerr.flush();
  return !(error > 1e-9);
}

#endif
