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

// Fundamental supernodes of a bound pattern and the multifrontal schedule over them. Supernode t
// holds the consecutive columns [first, first + width) whose structures nest (column j's rows are
// {j + 1} and column j + 1's rows, and column j + 1 has no other child), so its front is dense of
// order rows = width + |S| with S the rows below the last column. Fronts are listed level by
// level of the supernodal elimination tree (leaves first), children before parents.
struct Supernodes {
  std::vector<int32_t> first, width, rows;  // per supernode (schedule order)
  std::vector<int32_t> level;               // supernodal tree level
  std::vector<int32_t> parent;              // schedule index of the parent, -1 at a root
  std::vector<int64_t> update_offset;       // start of the packed |S| x |S| update matrix
  std::vector<int32_t> cls;                 // the front's class (the device's lane group)
  std::vector<int32_t> child_ptr, child;    // children (schedule indices, ascending first column)
  std::vector<int32_t> rel_ptr, rel;        // per supernode: its rows S as positions in the
                                            // parent's front rows
  int64_t update_total = 0;
  int32_t levels = 0;
};

// Supernodes of s, scheduled by level; within a level the fronts are ordered by their class, `cls`
// of the front order (the device groups fronts of one class), or, with `widen`, widen(cls(rows),
// the number of fronts at the front's level): a class may depend on how many fronts share its level.
Supernodes supernodes(const Symbolic& s, int32_t (*cls)(int32_t rows),
                      int32_t (*widen)(int32_t cls, int32_t level_fronts) = nullptr);

// Update matrices sharing storage: a supernode's update is written when its level runs and read
// when its parent's level runs, so it lives over the levels [level, parent's level]; updates whose
// lives do not meet take the same doubles (first fit, in schedule order, freed blocks coalesced).
// Rewrites update_offset and sets update_total to the peak. Throws if two live updates overlap.
void share_update_storage(Supernodes& sn);

// Relaxed supernodes: a fundamental supernode whose parent supernode starts at its next column
// and has no other child joins its parent's front when the joined front stays within max_rows
// and gives its first column at most max_zeros explicit zeros (chains joined from the top down).
// Each joined column's rows become the dense trapezoid's; the etree is unchanged.
// Returns those zeros as pattern entries (node pairs, flattened i0, j0, i1, j1, ...): binding the
// pattern plus them with s's order gives the relaxed symbolic, whose fundamental supernodes are
// the joined fronts.
std::vector<int32_t> relaxed_padding(const Symbolic& s, int32_t max_rows, int32_t max_zeros);

// A matrix-vector product's plan on a symmetric pattern with its whole diagonal: each row's
// entries (its neighbours and itself) in ascending column order, rows in order (CSR), and each
// row's diagonal entry.
struct ProductPlan {
  std::vector<int64_t> row_start;  // [n+1]
  std::vector<int32_t> column;     // [nnz]
  std::vector<int64_t> diagonal;   // [n]
};
ProductPlan product_plan(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols);

// The capacitance topology of a field layout (bind's host part): nodes are the n cells, then the o
// outside nodes; each active membrane joins its cell and its outside node into one capacitor
// component. A component holding a reservoir (a voltage clamp) is held; any other is floating and
// fixes one representative node: the gauge where it holds it, else its lowest outside node. Every
// floating component must reach a clamp or the gauge through conductive faces or return paths.
struct CapacitorTopology {
  std::vector<int32_t> component;  // [n+o] component index (first appearance by node), -1 inactive
  std::vector<int32_t> floating;   // [n+o] floating component index, -1 where held or inactive
  std::vector<uint8_t> fixed;      // [n+o] inactive, reservoir or representative
  std::vector<uint8_t> offset_free;  // [floating components] its representative is not the gauge
  int32_t components = 0;
};
// Throws std::invalid_argument with the reference's messages for the refused layouts.
CapacitorTopology capacitor_topology(int32_t n, int32_t o, int64_t m, const int32_t* membrane_cell,
                                     const int32_t* membrane_outside, const uint8_t* membrane_active,
                                     const uint8_t* cell_active, const uint8_t* outside_active,
                                     const uint8_t* reservoir, const uint8_t* gauge, int64_t f,
                                     const int32_t* face_sender, const int32_t* face_receiver,
                                     const uint8_t* face_conductive, int64_t r, const int32_t* return_paths);

std::vector<int32_t> minimum_degree(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols);
Symbolic bind(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols, const int32_t* order);
int64_t position(const Symbolic& s, int32_t i, int32_t j);

}  // namespace cells
