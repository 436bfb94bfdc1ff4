// Host symbolic analysis (C++): minimum-degree order, elimination tree, levels and the device
// plan of one SPD pattern. See symbolic.cpp for the reference functions it restates.
#pragma once
#include <cstdint>
#include <vector>

namespace cells {

struct Symbolic {
  int32_t n = 0;
  std::vector<int32_t> order;         // [n] k-th eliminated node
  std::vector<int32_t> rank;          // [n] inverse permutation
  std::vector<int64_t> column_start;  // [n+1]
  std::vector<int32_t> row_rank;      // [nnz] each entry's row, as a rank
  std::vector<int32_t> col_rank;      // [nnz] each entry's column, as a rank
  std::vector<int32_t> parent;        // [n] elimination tree by rank, -1 at a root
  std::vector<int32_t> level;         // [n] leaves 0, a parent above its children
  std::vector<int64_t> row_ptr;       // [n+1] strictly lower factor by rows
  std::vector<int32_t> row_col;       // columns m < r holding row r, ascending
  std::vector<int32_t> row_pos;       // their factor positions
  std::vector<int32_t> level_ptr;     // [levels+1]
  std::vector<int32_t> level_col;     // [n] column ranks grouped by level, ascending
  int64_t nnz() const { return column_start.empty() ? 0 : column_start.back(); }
  int32_t levels() const { return level_ptr.empty() ? 0 : static_cast<int32_t>(level_ptr.size()) - 1; }
};

std::vector<int32_t> minimum_degree(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols);
Symbolic bind(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols, const int32_t* order);
int64_t position(const Symbolic& s, int32_t i, int32_t j);

}  // namespace cells
