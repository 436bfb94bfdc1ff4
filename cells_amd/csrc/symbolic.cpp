// Host symbolic analysis of the stage matrix's pattern, once per lattice shape.
//
// Restates the reference's src/cells/numerics/sparse_cholesky.py host phase:
//   minimum_degree (:137-162)  heap of (current degree, original degree, index), stale entries
//                              skipped, the eliminated node's neighbours made a clique;
//   bind (:165-215)            rank = inverse permutation; below-sets by the elimination tree's
//                              union rule; parent = min(below); column-major entry positions,
//                              diagonal first then rows ascending by rank; levels.
// Their integer outputs are bit-identical to the reference's (order, rank, column_start, entry
// rows and columns, parent, level). From them this file also builds the device plan the batched
// factor and solves read (cells_hip.hip): the row structures (CSR of the strictly lower factor by
// rows, with factor positions), each entry's column, the level schedule, and each lattice entry's
// factor position.
#include "symbolic.h"

#include <algorithm>
#include <cstring>
#include <iterator>
#include <map>
#include <queue>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace cells {

namespace {

using Adjacency = std::vector<std::vector<int32_t>>;

Adjacency neighbours(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols) {
  Adjacency adj(n);
  for (int64_t q = 0; q < m; ++q) {
    int32_t i = rows[q], j = cols[q];
    if (i < 0 || j < 0 || i >= n || j >= n) throw std::invalid_argument("A sparse pattern's coordinates must lie within its size");
    if (i == j) continue;
    adj[i].push_back(j);
    adj[j].push_back(i);
  }
  for (auto& a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  return adj;
}

}  // namespace

std::vector<int32_t> minimum_degree(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols) {
  Adjacency adj = neighbours(n, m, rows, cols);
  std::vector<int32_t> original(n);
  for (int32_t i = 0; i < n; ++i) original[i] = static_cast<int32_t>(adj[i].size());
  using Key = std::tuple<int32_t, int32_t, int32_t>;
  std::priority_queue<Key, std::vector<Key>, std::greater<Key>> heap;
  for (int32_t i = 0; i < n; ++i) heap.emplace(original[i], original[i], i);
  std::vector<char> alive(n, 1);
  std::vector<int32_t> order;
  order.reserve(n);
  std::vector<int32_t> merged;
  while (!heap.empty()) {
    auto [degree, orig, node] = heap.top();
    (void)orig;
    heap.pop();
    if (!alive[node] || degree != static_cast<int32_t>(adj[node].size())) continue;
    alive[node] = 0;
    order.push_back(node);
    const std::vector<int32_t> clique = adj[node];
    for (int32_t a : clique) {
      // adjacency[a] = (adjacency[a] - {node}) | (clique - {a})
      std::vector<int32_t>& own = adj[a];
      merged.clear();
      merged.reserve(own.size() + clique.size());
      size_t x = 0, y = 0;
      while (x < own.size() || y < clique.size()) {
        int32_t v;
        if (y == clique.size() || (x < own.size() && own[x] < clique[y])) {
          v = own[x++];
        } else if (x == own.size() || clique[y] < own[x]) {
          v = clique[y++];
        } else {
          v = own[x++];
          ++y;
        }
        if (v != node && v != a) merged.push_back(v);
      }
      own.swap(merged);
      heap.emplace(static_cast<int32_t>(own.size()), original[a], a);
    }
    adj[node].clear();
  }
  return order;
}

Symbolic bind(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols, const int32_t* given) {
  if (n < 1) throw std::invalid_argument("a sparse pattern needs a positive size");
  Adjacency adj = neighbours(n, m, rows, cols);
  Symbolic s;
  s.n = n;
  if (given) {
    s.order.assign(given, given + n);
  } else {
    s.order = minimum_degree(n, m, rows, cols);
  }
  std::vector<char> seen(n, 0);
  if (static_cast<int32_t>(s.order.size()) != n) throw std::invalid_argument("order is not a permutation");
  for (int32_t v : s.order) {
    if (v < 0 || v >= n || seen[v]) throw std::invalid_argument("order is not a permutation");
    seen[v] = 1;
  }
  s.rank.assign(n, 0);
  for (int32_t k = 0; k < n; ++k) s.rank[s.order[k]] = k;
  // below sets by the union rule
  std::vector<std::vector<int32_t>> below(n);
  for (int32_t i = 0; i < n; ++i) {
    int32_t r = s.rank[i];
    for (int32_t a : adj[i])
      if (s.rank[a] > r) below[r].push_back(s.rank[a]);
  }
  for (auto& b : below) std::sort(b.begin(), b.end());
  s.parent.assign(n, -1);
  std::vector<int32_t> merged;
  for (int32_t k = 0; k < n; ++k) {
    if (below[k].empty()) continue;
    int32_t p = below[k].front();
    s.parent[k] = p;
    merged.clear();
    std::set_union(below[p].begin(), below[p].end(), below[k].begin() + 1, below[k].end(),
                   std::back_inserter(merged));
    below[p].swap(merged);
  }
  s.column_start.assign(n + 1, 0);
  for (int32_t k = 0; k < n; ++k) s.column_start[k + 1] = s.column_start[k] + 1 + static_cast<int64_t>(below[k].size());
  const int64_t nnz = s.column_start[n];
  s.row_rank.resize(nnz);
  s.col_rank.resize(nnz);
  for (int32_t k = 0; k < n; ++k) {
    int64_t p = s.column_start[k];
    s.row_rank[p] = k;
    s.col_rank[p] = k;
    for (size_t q = 0; q < below[k].size(); ++q) {
      s.row_rank[p + 1 + q] = below[k][q];
      s.col_rank[p + 1 + q] = k;
    }
  }
  s.level.assign(n, 0);
  for (int32_t k = 0; k < n; ++k)
    if (s.parent[k] >= 0) s.level[s.parent[k]] = std::max(s.level[s.parent[k]], s.level[k] + 1);
  // row structures: for row r, the columns m < r holding it, ascending, with positions
  s.row_ptr.assign(n + 1, 0);
  for (int64_t p = 0; p < nnz; ++p)
    if (s.row_rank[p] != s.col_rank[p]) s.row_ptr[s.row_rank[p] + 1]++;
  for (int32_t r = 0; r < n; ++r) s.row_ptr[r + 1] += s.row_ptr[r];
  s.row_col.resize(s.row_ptr[n]);
  s.row_pos.resize(s.row_ptr[n]);
  std::vector<int64_t> fill(s.row_ptr.begin(), s.row_ptr.end() - 1);
  for (int32_t k = 0; k < n; ++k)
    for (int64_t p = s.column_start[k] + 1; p < s.column_start[k + 1]; ++p) {
      int32_t r = s.row_rank[p];
      s.row_col[fill[r]] = k;
      s.row_pos[fill[r]] = static_cast<int32_t>(p);
      fill[r]++;
    }
  // level schedule
  int32_t levels = 0;
  for (int32_t k = 0; k < n; ++k) levels = std::max(levels, s.level[k] + 1);
  s.level_ptr.assign(levels + 1, 0);
  for (int32_t k = 0; k < n; ++k) s.level_ptr[s.level[k] + 1]++;
  for (int32_t v = 0; v < levels; ++v) s.level_ptr[v + 1] += s.level_ptr[v];
  s.level_col.resize(n);
  std::vector<int32_t> lf(s.level_ptr.begin(), s.level_ptr.end() - 1);
  for (int32_t k = 0; k < n; ++k) s.level_col[lf[s.level[k]]++] = k;
  return s;
}

ProductPlan product_plan(int32_t n, int64_t m, const int32_t* rows, const int32_t* cols) {
  // numerics/sparse_cholesky.py:444-464: row i's entries are its neighbours and itself, ascending
  Adjacency adj = neighbours(n, m, rows, cols);
  ProductPlan p;
  p.row_start.assign(1, 0);
  p.diagonal.resize(n);
  for (int32_t i = 0; i < n; ++i) {
    const auto& a = adj[i];
    auto at = std::lower_bound(a.begin(), a.end(), i);
    p.column.insert(p.column.end(), a.begin(), at);
    p.diagonal[i] = static_cast<int64_t>(p.column.size());
    p.column.push_back(i);
    p.column.insert(p.column.end(), at, a.end());
    p.row_start.push_back(static_cast<int64_t>(p.column.size()));
  }
  return p;
}

CapacitorTopology capacitor_topology(int32_t n, int32_t o, int64_t m, const int32_t* membrane_cell,
                                     const int32_t* membrane_outside, const uint8_t* membrane_active,
                                     const uint8_t* cell_active, const uint8_t* outside_active,
                                     const uint8_t* reservoir, const uint8_t* gauge, int64_t f,
                                     const int32_t* face_sender, const int32_t* face_receiver,
                                     const uint8_t* face_conductive, int64_t r, const int32_t* return_paths) {
  // numerics/surface_field.py:253-340 (bind) with representative (:238-250); the layout checks of
  // numerics/surface_transport.py:194-195 that bind relies on
  if (n < 0 || o < 1) throw std::invalid_argument("Surface geometry requires active cell and outside owners");
  const int32_t size = n + o;
  bool any_reservoir = false;
  int32_t gauges = 0;
  for (int32_t k = 0; k < o; ++k) {
    if (reservoir[k] && !outside_active[k]) throw std::invalid_argument("Reservoir must be an active outside owner");
    any_reservoir = any_reservoir || reservoir[k];
    gauges += gauge[k] ? 1 : 0;
    if (gauge[k] && !outside_active[k])
      throw std::invalid_argument("Use reservoir voltage clamps or exactly one active outside gauge");
  }
  if (gauges != (any_reservoir ? 0 : 1))
    throw std::invalid_argument("Use reservoir voltage clamps or exactly one active outside gauge");
  std::vector<int32_t> parent(size);
  for (int32_t i = 0; i < size; ++i) parent[i] = i;
  auto root = [&](int32_t i) {
    while (parent[i] != i) i = parent[i] = parent[parent[i]];
    return i;
  };
  auto join = [&](int32_t i, int32_t j) { parent[root(j)] = root(i); };
  auto node_of = [&](int64_t k, const int32_t* idx, int32_t count, int32_t base, const char* key) {
    if (idx[k] < 0 || idx[k] >= count) throw std::invalid_argument(std::string("Invalid resolved-surface owner: ") + key);
    return base + idx[k];
  };
  auto active = [&](int32_t i) { return i < n ? cell_active[i] != 0 : outside_active[i - n] != 0; };
  for (int64_t k = 0; k < m; ++k) {
    const int32_t i = node_of(k, membrane_cell, n, 0, "membrane_cell");
    const int32_t j = node_of(k, membrane_outside, o, n, "membrane_outside");
    if (!membrane_active[k]) continue;
    if (!active(i) || !active(j)) throw std::invalid_argument("Every route must join active owners");
    join(i, j);
  }
  CapacitorTopology t;
  t.component.assign(size, -1);
  t.floating.assign(size, -1);
  t.fixed.assign(size, 0);
  // components in order of their lowest active node (bind's groups, in its insertion order)
  std::vector<int32_t> comp_of_root(size, -1), first_member;
  std::vector<std::vector<int32_t>> members;
  for (int32_t i = 0; i < size; ++i) {
    if (!active(i)) {
      t.fixed[i] = 1;
      continue;
    }
    const int32_t rt = root(i);
    if (comp_of_root[rt] < 0) {
      comp_of_root[rt] = (int32_t)members.size();
      members.emplace_back();
    }
    t.component[i] = comp_of_root[rt];
    members[comp_of_root[rt]].push_back(i);
  }
  for (int32_t k = 0; k < o; ++k)
    if (reservoir[k]) t.fixed[n + k] = 1;
  t.components = (int32_t)members.size();
  std::vector<int32_t> floating_of(members.size(), -1);
  for (size_t c = 0; c < members.size(); ++c) {
    int32_t chosen = -1, lowest_outside = -1;
    bool held = false;
    for (int32_t i : members[c]) {
      if (i >= n && reservoir[i - n]) held = true;
      if (i >= n && gauge[i - n]) chosen = i;
      if (i >= n && lowest_outside < 0) lowest_outside = i;
    }
    if (held) continue;  // the held bath's imposed potential fixes the component
    if (chosen < 0) chosen = lowest_outside;
    if (chosen < 0)  // a component of cells alone: no outside node to fix (bind's min() has none)
      throw std::invalid_argument("Floating capacitor components require an outside return path to a clamp/gauge");
    t.fixed[chosen] = 1;
    floating_of[c] = (int32_t)t.offset_free.size();
    t.offset_free.push_back(gauge[chosen - n] ? 0 : 1);
    for (int32_t i : members[c]) t.floating[i] = floating_of[c];
  }
  // connectivity of the reduced offsets through guaranteed outside conductance: a component is
  // reached when it holds a clamp or the gauge, or shares a conductive face or a return path with a
  // reached one (to the fixed point)
  std::vector<uint8_t> reached(members.size(), 0);
  for (int32_t k = 0; k < o; ++k)
    if ((reservoir[k] || gauge[k]) && t.component[n + k] >= 0) reached[t.component[n + k]] = 1;
  std::vector<std::pair<int32_t, int32_t>> links;
  auto comp_at = [&](int32_t node) { return t.component[node]; };
  for (int64_t k = 0; k < f; ++k) {
    const int32_t a = node_of(k, face_sender, o, n, "face_sender"), b = node_of(k, face_receiver, o, n, "face_receiver");
    if (!face_conductive[k]) continue;
    if (!active(a) || !active(b)) throw std::invalid_argument("Every route must join active owners");
    links.emplace_back(comp_at(a), comp_at(b));
  }
  for (int64_t k = 0; k < r; ++k)
    links.emplace_back(comp_at(node_of(2 * k, return_paths, o, n, "return path")),
                       comp_at(node_of(2 * k + 1, return_paths, o, n, "return path")));
  for (bool grew = true; grew;) {
    grew = false;
    for (const auto& [a, b] : links) {
      if (a < 0 || b < 0) continue;
      if (reached[a] != reached[b]) {
        reached[a] = reached[b] = 1;
        grew = true;
      }
    }
  }
  for (size_t c = 0; c < members.size(); ++c)
    if (!reached[c])
      throw std::invalid_argument("Floating capacitor components require an outside return path to a clamp/gauge");
  return t;
}

int64_t position(const Symbolic& s, int32_t i, int32_t j) {
  int32_t a = s.rank[i], b = s.rank[j];
  int32_t hi = std::max(a, b), lo = std::min(a, b);
  int64_t p0 = s.column_start[lo], p1 = s.column_start[lo + 1];
  auto it = std::lower_bound(s.row_rank.begin() + p0, s.row_rank.begin() + p1, hi);
  if (it == s.row_rank.begin() + p1 || *it != hi) throw std::out_of_range("entry not in the bound pattern");
  return it - s.row_rank.begin();
}

Supernodes supernodes(const Symbolic& s, int32_t (*cls)(int32_t rows),
                      int32_t (*widen)(int32_t cls, int32_t level_fronts)) {
  const int32_t n = s.n;
  auto col_len = [&](int32_t j) { return static_cast<int32_t>(s.column_start[j + 1] - s.column_start[j]); };
  std::vector<int32_t> nchild(n, 0);
  for (int32_t k = 0; k < n; ++k)
    if (s.parent[k] >= 0) nchild[s.parent[k]]++;
  // fundamental supernodes in column order
  std::vector<int32_t> f0, fw, of(n);
  for (int32_t k = 0; k < n;) {
    int32_t j = k;
    while (j + 1 < n && s.parent[j] == j + 1 && nchild[j + 1] == 1 && col_len(j) == col_len(j + 1) + 1) ++j;
    for (int32_t q = k; q <= j; ++q) of[q] = static_cast<int32_t>(f0.size());
    f0.push_back(k);
    fw.push_back(j - k + 1);
    k = j + 1;
  }
  const int32_t ns = static_cast<int32_t>(f0.size());
  std::vector<int32_t> fpar(ns, -1), flev(ns, 0);
  for (int32_t t = 0; t < ns; ++t) {
    int32_t last = f0[t] + fw[t] - 1;
    if (s.parent[last] >= 0) fpar[t] = of[s.parent[last]];
  }
  for (int32_t t = 0; t < ns; ++t)  // parents follow children in column order
    if (fpar[t] >= 0) flev[fpar[t]] = std::max(flev[fpar[t]], flev[t] + 1);
  // schedule: by level, then class, then first column
  std::vector<int32_t> sched(ns);
  for (int32_t t = 0; t < ns; ++t) sched[t] = t;
  std::vector<int32_t> fcls(ns), per_level;
  for (int32_t t = 0; t < ns; ++t) {
    if (flev[t] >= (int32_t)per_level.size()) per_level.resize(flev[t] + 1, 0);
    per_level[flev[t]]++;
  }
  for (int32_t t = 0; t < ns; ++t) {
    fcls[t] = cls(col_len(f0[t]));
    if (widen) fcls[t] = widen(fcls[t], per_level[flev[t]]);
  }
  std::sort(sched.begin(), sched.end(), [&](int32_t a, int32_t b) {
    if (flev[a] != flev[b]) return flev[a] < flev[b];
    if (fcls[a] != fcls[b]) return fcls[a] < fcls[b];
    return f0[a] < f0[b];
  });
  std::vector<int32_t> pos(ns);
  for (int32_t q = 0; q < ns; ++q) pos[sched[q]] = q;
  Supernodes out;
  out.first.resize(ns);
  out.width.resize(ns);
  out.rows.resize(ns);
  out.level.resize(ns);
  out.parent.resize(ns);
  out.update_offset.resize(ns);
  out.cls.resize(ns);
  for (int32_t q = 0; q < ns; ++q) {
    int32_t t = sched[q];
    out.cls[q] = fcls[t];
    out.first[q] = f0[t];
    out.width[q] = fw[t];
    out.rows[q] = col_len(f0[t]);
    out.level[q] = flev[t];
    out.parent[q] = fpar[t] >= 0 ? pos[fpar[t]] : -1;
    int64_t S = out.rows[q] - out.width[q];
    out.update_offset[q] = out.update_total;
    out.update_total += S * (S + 1) / 2;
    out.levels = std::max(out.levels, flev[t] + 1);
  }
  // children, ascending by first column
  std::vector<std::vector<int32_t>> kids(ns);
  for (int32_t q = 0; q < ns; ++q)
    if (out.parent[q] >= 0) kids[out.parent[q]].push_back(q);
  out.child_ptr.assign(ns + 1, 0);
  for (int32_t q = 0; q < ns; ++q) {
    std::sort(kids[q].begin(), kids[q].end(), [&](int32_t a, int32_t b) { return out.first[a] < out.first[b]; });
    out.child_ptr[q + 1] = out.child_ptr[q] + static_cast<int32_t>(kids[q].size());
    out.child.insert(out.child.end(), kids[q].begin(), kids[q].end());
  }
  // each supernode's rows below its last column, as positions in its parent's front
  out.rel_ptr.assign(ns + 1, 0);
  std::vector<int32_t> where(n, -1);
  for (int32_t q = 0; q < ns; ++q) {
    int32_t last = out.first[q] + out.width[q] - 1;
    int64_t b0 = s.column_start[last] + 1, b1 = s.column_start[last + 1];
    int32_t p = out.parent[q];
    if (p >= 0) {
      int64_t c0 = s.column_start[out.first[p]], c1 = s.column_start[out.first[p] + 1];
      for (int64_t x = c0; x < c1; ++x) where[s.row_rank[x]] = static_cast<int32_t>(x - c0);
      for (int64_t x = b0; x < b1; ++x) {
        int32_t r = where[s.row_rank[x]];
        if (r < 0) throw std::runtime_error("a supernode's rows are not in its parent's front");
        out.rel.push_back(r);
      }
      for (int64_t x = c0; x < c1; ++x) where[s.row_rank[x]] = -1;
    } else if (b1 > b0) {
      throw std::runtime_error("a root supernode has rows below it");
    }
    out.rel_ptr[q + 1] = static_cast<int32_t>(out.rel.size());
  }
  return out;
}

void share_update_storage(Supernodes& sn) {
  const int32_t ns = static_cast<int32_t>(sn.first.size());
  auto size = [&](int32_t t) {
    const int64_t S = sn.rows[t] - sn.width[t];
    return sn.parent[t] >= 0 ? S * (S + 1) / 2 : 0;
  };
  std::map<int64_t, int64_t> free_blocks;  // start -> end of each free block below the top
  std::vector<std::vector<int32_t>> release(sn.levels + 1);
  int64_t top = 0;
  int32_t L = -1;
  sn.update_total = 0;
  for (int32_t t = 0; t < ns; ++t) {
    if (sn.level[t] < L) throw std::runtime_error("supernodes not in level order");
    for (; L < sn.level[t];) {  // a level's start frees the updates read at the level before it
      ++L;
      for (int32_t c : release[L]) {
        int64_t a = sn.update_offset[c], b = a + size(c);
        auto next = free_blocks.lower_bound(a);
        if (next != free_blocks.begin()) {
          auto prev = std::prev(next);
          if (prev->second == a) {
            a = prev->first;
            free_blocks.erase(prev);
          }
        }
        if (next != free_blocks.end() && next->first == b) {
          b = next->second;
          free_blocks.erase(next);
        }
        if (b == top) top = a;
        else free_blocks[a] = b;
      }
    }
    const int64_t need = size(t);
    if (need == 0) {
      sn.update_offset[t] = 0;
      continue;
    }
    auto it = free_blocks.begin();
    while (it != free_blocks.end() && it->second - it->first < need) ++it;
    if (it != free_blocks.end()) {
      sn.update_offset[t] = it->first;
      const int64_t rest = it->first + need, end = it->second;
      free_blocks.erase(it);
      if (rest < end) free_blocks[rest] = end;
    } else {
      sn.update_offset[t] = top;
      top += need;
    }
    sn.update_total = std::max(sn.update_total, top);
    release[sn.level[sn.parent[t]] + 1].push_back(t);
  }
  // check: at every level the live updates are disjoint
  std::vector<std::vector<std::pair<int64_t, int64_t>>> live(sn.levels);
  for (int32_t t = 0; t < ns; ++t)
    if (size(t) > 0)
      for (int32_t v = sn.level[t]; v <= sn.level[sn.parent[t]]; ++v)
        live[v].emplace_back(sn.update_offset[t], sn.update_offset[t] + size(t));
  for (auto& blocks : live) {
    std::sort(blocks.begin(), blocks.end());
    for (size_t x = 1; x < blocks.size(); ++x)
      if (blocks[x].first < blocks[x - 1].second) throw std::runtime_error("live update matrices overlap");
  }
}

std::vector<int32_t> relaxed_padding(const Symbolic& s, int32_t max_rows, int32_t max_zeros) {
  const int32_t n = s.n;
  auto col_len = [&](int32_t j) { return static_cast<int32_t>(s.column_start[j + 1] - s.column_start[j]); };
  std::vector<int32_t> nchild(n, 0);
  for (int32_t k = 0; k < n; ++k)
    if (s.parent[k] >= 0) nchild[s.parent[k]]++;
  std::vector<int32_t> f0, fw, of(n);
  for (int32_t k = 0; k < n;) {
    int32_t j = k;
    while (j + 1 < n && s.parent[j] == j + 1 && nchild[j + 1] == 1 && col_len(j) == col_len(j + 1) + 1) ++j;
    for (int32_t q = k; q <= j; ++q) of[q] = static_cast<int32_t>(f0.size());
    f0.push_back(k);
    fw.push_back(j - k + 1);
    k = j + 1;
  }
  const int32_t ns = static_cast<int32_t>(f0.size());
  std::vector<int32_t> fpar(ns, -1), kids(ns, 0);
  for (int32_t t = 0; t < ns; ++t) {
    const int32_t last = f0[t] + fw[t] - 1;
    if (s.parent[last] >= 0) fpar[t] = of[s.parent[last]];
    if (fpar[t] >= 0) kids[fpar[t]]++;
  }
  // chains from the top down: top[t] is the top supernode of t's front, order[top] its order
  std::vector<int32_t> top(ns), order(ns);
  for (int32_t t = ns - 1; t >= 0; --t) {
    top[t] = t;
    order[t] = col_len(f0[t]);
    const int32_t p = fpar[t];
    if (p >= 0 && kids[p] == 1 && f0[t] + fw[t] == f0[p] && s.parent[f0[p] - 1] == f0[p] && order[top[p]] + fw[t] <= max_rows) {
      // zeros of t's first column in the joined front (its later columns add no more): the
      // front's order less the column's own count
      if (order[top[p]] + fw[t] - col_len(f0[t]) <= max_zeros) {
        top[t] = top[p];
        order[top[p]] += fw[t];
      }
    }
  }
  std::vector<int32_t> pad;
  std::vector<char> in(n, 0);
  for (int32_t t = 0; t < ns; ++t) {
    if (top[t] == t) continue;
    const int32_t T = top[t], end = f0[T] + fw[T];  // the front's columns end at the top's end
    const int32_t last = end - 1;
    for (int32_t j = f0[t]; j < f0[t] + fw[t]; ++j) {
      // the trapezoid's rows below column j: columns j+1 .. last, then the top's rows below
      for (int64_t p = s.column_start[j] + 1; p < s.column_start[j + 1]; ++p) in[s.row_rank[p]] = 1;
      auto want = [&](int32_t r) {
        if (!in[r]) {
          pad.push_back(s.order[r]);
          pad.push_back(s.order[j]);
        }
      };
      for (int32_t r = j + 1; r <= last; ++r) want(r);
      for (int64_t p = s.column_start[last] + 1; p < s.column_start[last + 1]; ++p) want(s.row_rank[p]);
      for (int64_t p = s.column_start[j] + 1; p < s.column_start[j + 1]; ++p) in[s.row_rank[p]] = 0;
    }
  }
  return pad;
}

}  // namespace cells
