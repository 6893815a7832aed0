// matrix_market.h -- Matrix Market files and BCSR block-size detection for
// the C++ drop-in layer, over the C ABI (bmfp_mm_read, bmfp_mm_write,
// bmfp_block_size_detect).  The reference reads and writes Matrix Market
// through common/mmio.{h,c}; csr_matrix_class::write_mm (matrix_converters.h)
// writes the same text as write_matrix_market here.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "bmfp.h"
#include "matrix_converters.h"

// CSR from a coordinate (real, integer, pattern; general, symmetric,
// skew-symmetric) or array (real, integer; general) Matrix Market file:
// columns ascending within a row, duplicates summed, symmetric halves
// mirrored.  Throws std::runtime_error with the library's message.
template <typename data_type, typename index_type>
std::unique_ptr<csr_matrix_class<data_type, index_type>> read_matrix_market(const std::string &filename)
{
  int n_rows = 0, n_cols = 0, nnz = 0;
  if (bmfp_mm_read(filename.c_str(), &n_rows, &n_cols, &nnz, nullptr, nullptr, nullptr) != BMFP_OK)
    throw std::runtime_error(bmfp_last_error());
  std::vector<int> rp((size_t) n_rows + 1), cols((size_t) nnz);
  std::vector<double> vals((size_t) nnz);
  if (bmfp_mm_read(filename.c_str(), &n_rows, &n_cols, &nnz, rp.data(), cols.data(), vals.data()) != BMFP_OK)
    throw std::runtime_error(bmfp_last_error());
  std::vector<index_type> rpi(rp.begin(), rp.end()), ci(cols.begin(), cols.end());
  std::vector<data_type> v(vals.begin(), vals.end());
  return std::unique_ptr<csr_matrix_class<data_type, index_type>>(new csr_matrix_class<data_type, index_type>(
      (index_type) n_rows, (index_type) n_cols, (index_type) nnz, rpi.data(), ci.data(), v.data()));
}

// "coordinate real general" text, identical to csr_matrix_class::write_mm.
template <typename data_type, typename index_type>
void write_matrix_market(const csr_matrix_class<data_type, index_type> &m, const std::string &filename)
{
  std::vector<int> rp(m.row_ptr.get(), m.row_ptr.get() + m.n_rows + 1), cols(m.columns.get(), m.columns.get() + m.nnz);
  std::vector<double> vals(m.values.get(), m.values.get() + m.nnz);
  if (bmfp_mm_write(filename.c_str(), (int) m.n_rows, (int) m.n_cols, rp.data(), cols.data(), vals.data()) != BMFP_OK)
    throw std::runtime_error(bmfp_last_error());
}

struct block_size_choice {
  int bs;        // chosen block size
  int nnzb;      // nonzero blocks at that size
  double fill;   // stored values / nonzeros (1 = no explicit zeros added)
};

// The candidate (dividing both dimensions) whose BCSR SpMV moves the fewest
// compulsory bytes with values of sizeof(data_type) bytes.
template <typename data_type, typename index_type>
block_size_choice detect_block_size(const csr_matrix_class<data_type, index_type> &m,
                                    const std::vector<int> &candidates = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32})
{
  std::vector<int> rp(m.row_ptr.get(), m.row_ptr.get() + m.n_rows + 1), cols(m.columns.get(), m.columns.get() + m.nnz);
  block_size_choice c{};
  if (bmfp_block_size_detect((int) m.n_rows, (int) m.n_cols, rp.data(), cols.data(), (int) sizeof(data_type),
                             candidates.data(), (int) candidates.size(), &c.bs, &c.nnzb, &c.fill) != BMFP_OK)
    throw std::runtime_error(bmfp_last_error());
  return c;
}
