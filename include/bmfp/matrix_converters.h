// matrix_converters.h -- host BCSR / CSR containers and converters, drop-in for
// the reference's common/matrix_converters.h (same class names, members,
// constructors and methods; matrix_converters.h:13-196).
//
// Additions beyond the reference: csr_to_bcsr (the inverse of the CSR
// expansion), gen_stencil9_bcsr (the seeded 9-point block-stencil workloads,
// generated by libbmfp), and a Matrix Market writer of its own
// (csr_matrix_class::write_mm).
#ifndef BMFP_MATRIX_CONVERTERS_H
#define BMFP_MATRIX_CONVERTERS_H

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "bmfp.h"

template <typename data_type, typename index_type>
class bcsr_matrix_class
{
public:
  // matrix_converters.h:16-29: storage for nnzb blocks of bs x bs values.
  bcsr_matrix_class(index_type n_rows_arg, index_type n_cols_arg, index_type bs_arg, index_type nnzb_arg)
    : n_rows(n_rows_arg)
    , n_cols(n_cols_arg)
    , bs(bs_arg)
    , nnzb(nnzb_arg)
    , values(new data_type[(size_t) nnzb_arg * bs_arg * bs_arg])
    , columns(new index_type[nnzb_arg])
    , row_ptr(new index_type[(size_t) n_rows_arg + 1])
  {
  }

  // matrix_converters.h:32-49: per-block transpose into new_values (row-major
  // blocks -> column-major blocks and back).
  void transpose_blocks(data_type *new_values) const
  {
    const size_t bs2 = (size_t) bs * bs;
    for (index_type row = 0; row < n_rows; row++)
      for (index_type b = row_ptr[row]; b < row_ptr[row + 1]; b++) {
        const data_type *o = values.get() + bs2 * b;
        data_type *n = new_values + bs2 * b;
        for (index_type i = 0; i < bs; i++)
          for (index_type j = 0; j < bs; j++) n[j * bs + i] = o[i * bs + j];
      }
  }

  index_type size() const { return nnzb * bs * bs; }

  data_type *get_block_data(index_type row, index_type block_in_row)
  {
    return values.get() + (size_t) (row_ptr[row] + block_in_row) * bs * bs;
  }

  // matrix_converters.h:61-70: block of `row` at block column `column`
  // (lower_bound over the row's sorted columns).
  data_type *get_block_data_by_column(index_type row, index_type column)
  {
    const index_type *first = columns.get() + row_ptr[row];
    const index_type *last = columns.get() + row_ptr[row + 1];
    return get_block_data(row, (index_type) std::distance(first, std::lower_bound(first, last, column)));
  }

public:
  const index_type n_rows{};
  const index_type n_cols{};
  const index_type bs{};
  const index_type nnzb{};

  const std::unique_ptr<data_type[]> values;
  const std::unique_ptr<index_type[]> columns;
  const std::unique_ptr<index_type[]> row_ptr;
};

template <typename data_type, typename index_type>
class csr_matrix_class
{
public:
  // matrix_converters.h:88-117: scalar row br*bs+r lists, block by block, the
  // bs entries (col*bs+c, values[b*bs*bs + r*bs + c]).
  explicit csr_matrix_class(const bcsr_matrix_class<data_type, index_type> &m)
    : n_rows(m.n_rows * m.bs)
    , n_cols(m.n_cols * m.bs)
    , nnz(m.nnzb * m.bs * m.bs)
    , values(new data_type[(size_t) m.nnzb * m.bs * m.bs])
    , columns(new index_type[(size_t) m.nnzb * m.bs * m.bs])
    , row_ptr(new index_type[(size_t) m.n_rows * m.bs + 1])
  {
    size_t off = 0;
    const size_t bs2 = (size_t) m.bs * m.bs;
    for (index_type br = 0; br < m.n_rows; br++)
      for (index_type r = 0; r < m.bs; r++) {
        row_ptr[(size_t) br * m.bs + r] = (index_type) off;
        for (index_type b = m.row_ptr[br]; b < m.row_ptr[br + 1]; b++)
          for (index_type c = 0; c < m.bs; c++) {
            columns[off] = m.columns[b] * m.bs + c;
            values[off++] = m.values[bs2 * b + (size_t) r * m.bs + c];
          }
      }
    row_ptr[n_rows] = (index_type) off;
  }

  // Plain CSR from arrays (copied).
  csr_matrix_class(index_type n_rows_arg, index_type n_cols_arg, index_type nnz_arg, const index_type *rp,
                   const index_type *cols, const data_type *vals)
    : n_rows(n_rows_arg)
    , n_cols(n_cols_arg)
    , nnz(nnz_arg)
    , values(new data_type[nnz_arg])
    , columns(new index_type[nnz_arg])
    , row_ptr(new index_type[(size_t) n_rows_arg + 1])
  {
    std::copy(rp, rp + n_rows + 1, row_ptr.get());
    std::copy(cols, cols + nnz, columns.get());
    std::copy(vals, vals + nnz, values.get());
  }

  // Coordinate Matrix Market file, 1-based indices ("%d %d %.20g" per entry),
  // as matrix_converters.h:119-144 writes through mmio.
  void write_mm(const std::string &filename) const
  {
    FILE *fp = std::fopen(filename.c_str(), "w");
    if (!fp) throw std::runtime_error("write_mm: cannot open " + filename);
    std::fprintf(fp, "%%%%MatrixMarket matrix coordinate real general\n");
    std::fprintf(fp, "%lld %lld %lld\n", (long long) n_rows, (long long) n_cols, (long long) nnz);
    for (index_type row = 0; row < n_rows; row++)
      for (index_type e = row_ptr[row]; e < row_ptr[row + 1]; e++)
        std::fprintf(fp, "%lld %lld %10.20g\n", (long long) row + 1, (long long) columns[e] + 1,
                     (double) values[e]);
    std::fclose(fp);
  }

public:
  const index_type n_rows{};
  const index_type n_cols{};
  const index_type nnz{};

  const std::unique_ptr<data_type[]> values;
  const std::unique_ptr<index_type[]> columns;
  const std::unique_ptr<index_type[]> row_ptr;
};

// matrix_converters.h:157-196: banded synthetic BCSR; block row `row` holds
// blocks_per_row consecutive block columns starting at
// clamp(row - blocks_per_row/2, 0, n_rows - blocks_per_row), entry i of the
// block at column col equal to (col + i) / n_rows.
template <typename data_type, typename index_type>
std::unique_ptr<bcsr_matrix_class<data_type, index_type>> gen_n_diag_bcsr(index_type n_rows_arg,
                                                                          index_type blocks_per_row,
                                                                          index_type bs_arg)
{
  std::unique_ptr<bcsr_matrix_class<data_type, index_type>> m(
      new bcsr_matrix_class<data_type, index_type>(n_rows_arg, n_rows_arg, bs_arg, n_rows_arg * blocks_per_row));
  index_type block_id = 0;
  for (index_type row = 0; row < n_rows_arg; row++) {
    m->row_ptr[row] = row * blocks_per_row;
    index_type first = row - blocks_per_row / 2;
    if (row < blocks_per_row / 2) first = 0;
    else if (row > n_rows_arg - blocks_per_row) first = n_rows_arg - blocks_per_row;
    for (index_type e = 0; e < blocks_per_row; e++) {
      const index_type col = first + e;
      data_type *blk = m->values.get() + (size_t) block_id * bs_arg * bs_arg;
      for (unsigned i = 0; i < (unsigned) (bs_arg * bs_arg); i++)
        blk[i] = (static_cast<data_type>(col) + i) / n_rows_arg;
      m->columns[block_id++] = col;
    }
  }
  m->row_ptr[n_rows_arg] = n_rows_arg * blocks_per_row;
  return m;
}

// Inverse of csr_matrix_class(const bcsr_matrix_class&) (no reference
// counterpart): groups the CSR pattern into bs x bs blocks with sorted block
// columns, filling absent entries of a touched block with zeros.  Round trip
// BCSR -> CSR -> BCSR is the identity for sorted block columns.
template <typename data_type, typename index_type>
std::unique_ptr<bcsr_matrix_class<data_type, index_type>> csr_to_bcsr(
    const csr_matrix_class<data_type, index_type> &c, index_type bs)
{
  if (bs < 1 || c.n_rows % bs || c.n_cols % bs)
    throw std::invalid_argument("csr_to_bcsr: matrix dimensions must be multiples of bs");
  const index_type nb = c.n_rows / bs;
  std::vector<std::vector<index_type>> bcols((size_t) nb);
  size_t nnzb = 0;
  for (index_type br = 0; br < nb; br++) {
    auto &v = bcols[br];
    for (index_type r = 0; r < bs; r++)
      for (index_type e = c.row_ptr[br * bs + r]; e < c.row_ptr[br * bs + r + 1]; e++) v.push_back(c.columns[e] / bs);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    nnzb += v.size();
  }
  std::unique_ptr<bcsr_matrix_class<data_type, index_type>> m(
      new bcsr_matrix_class<data_type, index_type>(nb, c.n_cols / bs, bs, (index_type) nnzb));
  std::fill(m->values.get(), m->values.get() + nnzb * bs * bs, data_type(0));
  size_t k = 0;
  m->row_ptr[0] = 0;
  for (index_type br = 0; br < nb; br++) {
    const auto &v = bcols[br];
    std::copy(v.begin(), v.end(), m->columns.get() + k);
    for (index_type r = 0; r < bs; r++)
      for (index_type e = c.row_ptr[br * bs + r]; e < c.row_ptr[br * bs + r + 1]; e++) {
        const size_t pos = (size_t) (std::lower_bound(v.begin(), v.end(), c.columns[e] / bs) - v.begin());
        m->values[(k + pos) * bs * bs + (size_t) r * bs + c.columns[e] % bs] = c.values[e];
      }
    k += v.size();
    m->row_ptr[br + 1] = (index_type) k;
  }
  return m;
}

// Seeded 9-point block stencil on an nx x ny node grid (kind 0: diagonally
// dominant nonsymmetric; kind 1: convective -(K+0.3C) off-diagonal blocks),
// generated by libbmfp's host generator (identical bits on every host).
template <typename data_type, typename index_type>
std::unique_ptr<bcsr_matrix_class<data_type, index_type>> gen_stencil9_bcsr(int nx, int ny, int bs, int kind,
                                                                            uint64_t seed)
{
  const int64_t nnzb = bmfp_stencil_nnzb(nx, ny);
  const int n = nx * ny;
  std::vector<int> rp((size_t) n + 1), cols((size_t) nnzb);
  std::vector<double> vals((size_t) nnzb * bs * bs);
  if (bmfp_gen_stencil9(nx, ny, bs, kind, seed, rp.data(), cols.data(), vals.data()) != BMFP_OK)
    throw std::invalid_argument(std::string("gen_stencil9_bcsr: ") + bmfp_last_error());
  std::unique_ptr<bcsr_matrix_class<data_type, index_type>> m(
      new bcsr_matrix_class<data_type, index_type>(n, n, bs, (index_type) nnzb));
  std::copy(rp.begin(), rp.end(), m->row_ptr.get());
  std::copy(cols.begin(), cols.end(), m->columns.get());
  for (size_t i = 0; i < vals.size(); i++) m->values[i] = static_cast<data_type>(vals[i]);
  return m;
}

#endif
