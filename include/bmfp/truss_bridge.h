// truss_bridge.h -- 2-D suspension-bridge truss: a real bs = 2 system for the
// Jacobi-preconditioned BiCGStab (the role of the reference's
// fem_2d/golden_gate_bridge.h:48-909 and its caller main.cpp:349-381: build
// the bridge, write the matrix and a VTK file, solve K u = f, write the
// displaced bridge).  Header only, host code.
//
// The structure (this build's own, bars only):
//   deck    a Warren truss with verticals: top chord at deck_height, bottom
//           chord deck_depth lower, a panel point every `panel` metres over
//           side + main + side; pinned (x and y) at both ends;
//   towers  one bar each from a pinned base (y = 0) to tower_height, at the
//           panel points nearest x = side and x = side + main;
//   cables  a parabolic main cable between the tower tops, lowest point 1 m
//           above the deck at mid-span; straight side cables from each tower
//           top down to the deck's end nodes;
//   hangers vertical bars from every cable node to the deck's top chord.
// Each bar contributes E A / L * [n n^T, -n n^T; -n n^T, n n^T] (n its unit
// direction) to the 2x2 blocks of its end nodes.  A supported degree of
// freedom becomes an identity row and column with a zero right-hand side, so
// K stays symmetric and nonsingular.  Nodes are numbered along x (top, bottom,
// then cable or tower nodes of each panel point), which keeps the matrix banded.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "matrix_converters.h"

template <typename data_type, typename index_type, bool use_frames = false>
class golden_gate_bridge_2d
{
  static_assert(!use_frames, "frame (bending) elements are not built: use_frames must be false");

public:
  static constexpr index_type bs = 2;  // x and y displacement per node

  // load(x) -> (fx, fy) in newtons for the deck's top-chord node at x.
  template <typename load_type>
  golden_gate_bridge_2d(const load_type &load, data_type main_part_length_arg, data_type side_length_arg,
                        data_type tower_height_arg = 230.0, data_type segment_length_arg = 7.62)
    : main_length(main_part_length_arg), side_length(side_length_arg), tower_height(tower_height_arg),
      panel(segment_length_arg)
  {
    if (!(panel > 0) || !(main_length > 2 * panel) || !(side_length > 2 * panel) || !(tower_height > deck_height + 2))
      throw std::invalid_argument("golden_gate_bridge_2d: bad dimensions");
    build_geometry();
    assemble();
    forces_rhs.reset(new data_type[(size_t) bs * nodes.size()]);
    std::fill_n(forces_rhs.get(), (size_t) bs * nodes.size(), data_type(0));
    for (size_t i = 0; i < top.size(); i++) {
      const std::pair<data_type, data_type> f = load(nodes[top[i]].x);
      forces_rhs[(size_t) bs * top[i] + 0] = f.first;
      forces_rhs[(size_t) bs * top[i] + 1] = f.second;
    }
    for (size_t d = 0; d < fixed.size(); d++)
      if (fixed[d]) forces_rhs[d] = 0;
  }

  index_type nodes_count() const { return (index_type) nodes.size(); }
  index_type elements_count() const { return (index_type) elements.size(); }
  data_type node_x(index_type i) const { return nodes[i].x; }
  data_type node_y(index_type i) const { return nodes[i].y; }
  std::pair<index_type, index_type> element_nodes(index_type e) const { return {elements[e].a, elements[e].b}; }
  data_type element_ea(index_type e) const { return elements[e].ea; }
  bool dof_fixed(index_type dof) const { return fixed[dof] != 0; }

  // Legacy VTK unstructured grid: the nodes (displaced when `displacement`
  // is given), one line cell per bar, and the point scalar AvgDisplacement =
  // (|ux| + |uy|) / 2 (the layout fem_2d/golden_gate_bridge.h:239-283 writes).
  void write_vtk(const std::string &filename, const data_type *displacement = nullptr) const
  {
    std::ofstream vtk(filename);
    if (!vtk) throw std::runtime_error("write_vtk: cannot open " + filename);
    vtk << std::setprecision(17);
    vtk << "# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET UNSTRUCTURED_GRID\n";
    vtk << "POINTS " << nodes.size() << " double\n";
    for (size_t i = 0; i < nodes.size(); i++) {
      const double ux = displacement ? (double) displacement[bs * i] : 0.0;
      const double uy = displacement ? (double) displacement[bs * i + 1] : 0.0;
      vtk << (double) nodes[i].x + ux << " " << (double) nodes[i].y + uy << " 0\n";
    }
    vtk << "CELLS " << elements.size() << " " << 3 * elements.size() << "\n";
    for (const Bar &e : elements) vtk << "2 " << e.a << " " << e.b << "\n";
    vtk << "CELL_TYPES " << elements.size() << "\n";
    for (size_t e = 0; e < elements.size(); e++) vtk << "3\n";  // VTK_LINE
    vtk << "POINT_DATA " << nodes.size() << "\nSCALARS AvgDisplacement double 1\nLOOKUP_TABLE default\n";
    for (size_t i = 0; i < nodes.size(); i++)
      vtk << (displacement ? (std::fabs((double) displacement[bs * i]) + std::fabs((double) displacement[bs * i + 1])) / 2
                           : 0.0)
          << "\n";
  }

  std::unique_ptr<bcsr_matrix_class<data_type, index_type>> matrix;  // K, supports as identity rows
  std::unique_ptr<data_type[]> forces_rhs;                           // f, zero at supported dofs

  // physical constants (steel)
  static constexpr double E = 2e11;                                           // Pa
  static constexpr double deck_area = 0.45, tower_area = 0.62, cable_area = 0.36, hanger_area = 0.06;  // m^2
  static constexpr double deck_height = 78.0, deck_depth = 7.62;             // m

private:
  struct Node {
    data_type x, y;
  };
  struct Bar {
    index_type a, b;
    double ea;
  };

  const data_type main_length, side_length, tower_height, panel;
  std::vector<Node> nodes;
  std::vector<Bar> elements;
  std::vector<index_type> top, bottom;  // deck chord nodes per panel point
  std::vector<char> fixed;              // per dof

  index_type add_node(data_type x, data_type y)
  {
    nodes.push_back({x, y});
    return (index_type) nodes.size() - 1;
  }
  void add_bar(index_type a, index_type b, double area) { elements.push_back({a, b, E * area}); }

  void build_geometry()
  {
    const data_type total = side_length + main_length + side_length;
    const int np = std::max(4, (int) std::lround(total / panel));  // panels
    const data_type dx = total / np;
    const int il = (int) std::lround(side_length / dx), ir = (int) std::lround((side_length + main_length) / dx);
    // main cable: y = a (x - xm)^2 + ym through the tower tops, ym 1 m above the deck
    const data_type xl = il * dx, xr = ir * dx, xm = (xl + xr) / 2, ym = deck_height + 1;
    const data_type a = (tower_height - ym) / ((xl - xm) * (xl - xm));
    std::vector<index_type> cable(np + 1, -1);
    index_type tower_top_l = -1, tower_top_r = -1, base_l = -1, base_r = -1;
    for (int i = 0; i <= np; i++) {
      const data_type x = i * dx;
      top.push_back(add_node(x, deck_height));
      bottom.push_back(add_node(x, deck_height - deck_depth));
      if (i == il || i == ir) {
        const index_type g = add_node(x, 0), p = add_node(x, tower_height);
        (i == il ? base_l : base_r) = g;
        (i == il ? tower_top_l : tower_top_r) = p;
        cable[i] = p;
      } else if (i > 0 && i < np) {
        data_type y;
        if (i < il) y = deck_height + (tower_height - deck_height) * x / xl;                        // left side cable
        else if (i > ir) y = deck_height + (tower_height - deck_height) * (total - x) / (total - xr);  // right
        else y = a * (x - xm) * (x - xm) + ym;                                                      // main cable
        cable[i] = add_node(x, y);
      }
    }
    for (int i = 0; i < np; i++) {
      add_bar(top[i], top[i + 1], deck_area);
      add_bar(bottom[i], bottom[i + 1], deck_area);
      add_bar(top[i], bottom[i], deck_area);
      if (i % 2 == 0) add_bar(top[i], bottom[i + 1], deck_area);
      else add_bar(bottom[i], top[i + 1], deck_area);
    }
    add_bar(top[np], bottom[np], deck_area);
    add_bar(base_l, tower_top_l, tower_area);
    add_bar(base_r, tower_top_r, tower_area);
    // cables: deck end -> side cable -> tower top -> main cable -> tower top -> side cable -> deck end
    index_type prev = top[0];
    for (int i = 1; i < np; i++) {
      add_bar(prev, cable[i], cable_area);
      if (i != il && i != ir) add_bar(cable[i], top[i], hanger_area);
      prev = cable[i];
    }
    add_bar(prev, top[np], cable_area);
    fixed.assign((size_t) bs * nodes.size(), 0);
    for (index_type n : {top[0], bottom[0], top[np], bottom[np], base_l, base_r})
      fixed[(size_t) bs * n] = fixed[(size_t) bs * n + 1] = 1;
  }

  void assemble()
  {
    const size_t nn = nodes.size();
    std::vector<std::vector<index_type>> adj(nn);
    for (size_t i = 0; i < nn; i++) adj[i].push_back((index_type) i);
    for (const Bar &e : elements) {
      adj[e.a].push_back(e.b);
      adj[e.b].push_back(e.a);
    }
    size_t nnzb = 0;
    for (auto &v : adj) {
      std::sort(v.begin(), v.end());
      v.erase(std::unique(v.begin(), v.end()), v.end());
      nnzb += v.size();
    }
    matrix.reset(new bcsr_matrix_class<data_type, index_type>((index_type) nn, (index_type) nn, bs, (index_type) nnzb));
    index_type off = 0;
    for (size_t i = 0; i < nn; i++) {
      matrix->row_ptr[i] = off;
      for (index_type c : adj[i]) matrix->columns[off++] = c;
    }
    matrix->row_ptr[nn] = off;
    std::fill_n(matrix->values.get(), (size_t) nnzb * bs * bs, data_type(0));
    auto block = [&](index_type r, index_type c) {
      const index_type *b = matrix->columns.get() + matrix->row_ptr[r], *e = matrix->columns.get() + matrix->row_ptr[r + 1];
      return matrix->values.get() + (size_t) (std::lower_bound(b, e, c) - matrix->columns.get()) * bs * bs;
    };
    for (const Bar &e : elements) {
      const double dx = (double) nodes[e.b].x - nodes[e.a].x, dy = (double) nodes[e.b].y - nodes[e.a].y;
      const double len = std::sqrt(dx * dx + dy * dy), c = dx / len, s = dy / len, k = e.ea / len;
      const double kk[2][2] = {{k * c * c, k * c * s}, {k * c * s, k * s * s}};
      for (int p = 0; p < 2; p++)
        for (int q = 0; q < 2; q++) {
          const index_type r = p ? e.b : e.a, col = q ? e.b : e.a;
          data_type *blk = block(r, col);
          const double sign = p == q ? 1.0 : -1.0;
          for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) blk[i * bs + j] += (data_type) (sign * kk[i][j]);
        }
    }
    // supports: identity row and column
    for (size_t r = 0; r < nn; r++)
      for (index_type k = matrix->row_ptr[r]; k < matrix->row_ptr[r + 1]; k++) {
        const index_type c = matrix->columns[k];
        data_type *blk = matrix->values.get() + (size_t) k * bs * bs;
        for (int i = 0; i < bs; i++)
          for (int j = 0; j < bs; j++)
            if (fixed[bs * r + i] || fixed[(size_t) bs * c + j]) blk[i * bs + j] = (r == (size_t) c && i == j) ? 1 : 0;
      }
  }
};
