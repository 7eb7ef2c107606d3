// Device layout of the inverse factor: postorder renumbering and supernode
// blocks.
//
// The factor S = D^{-1/2} L^{-1} P comes from factor.cpp with rows in
// elimination order and original column indices (the restatement of
// DP/sparse.py:291-320). For the products on the device it is renumbered:
// with the structural elimination tree of P A P^T (Liu's algorithm) and a
// postorder of that tree, (L^{-1})_{ij} can be nonzero only when j is a
// descendant of i (every path in the graph of L climbs the tree), and the
// descendants of i are the postorder range ending at i. So row r of S is the
// dense range [r - subtree(r) + 1, r]. Where a row's parent has it as its
// only child, the parent's range is the child's plus one column, so a chain
// of such rows (a supernode) is one dense trapezoid, stored as a block; a
// small subtree is stored whole as one block. The values are the factor's
// own numbers, only placed differently.
#include <algorithm>
#include <functional>
#include <queue>
#include <stdexcept>
#include <utility>

#include "setup.h"

namespace dphd {

namespace {

// Largest subtree stored as one block (a leaf of the dissection has at most
// 24 vertices).
constexpr int kAmalgamate = 32;

// Structural elimination tree of the symmetric pattern of `a` in the order
// `perm` (perm[position] = original index). parent[i] = -1 for roots.
std::vector<int> elimination_tree(const Csr& a, const std::vector<int>& perm, const std::vector<int>& pos) {
  const int n = a.n;
  std::vector<int> parent(n, -1), ancestor(n, -1);
  for (int i = 0; i < n; ++i) {
    const int o = perm[i];
    for (int p = a.off[o]; p < a.off[o + 1]; ++p) {
      int j = pos[a.col[p]];
      while (j != -1 && j < i) {
        const int next = ancestor[j];
        ancestor[j] = i;
        if (next == -1) parent[j] = i;
        j = next;
      }
    }
  }
  return parent;
}

// Postorder of a forest given by parent pointers: children visited in
// ascending order, roots in ascending order. Returns order[k] = node.
std::vector<int> postorder(const std::vector<int>& parent) {
  const int n = static_cast<int>(parent.size());
  std::vector<int> head(n, -1), next(n, -1), roots;
  for (int j = n - 1; j >= 0; --j) {
    if (parent[j] >= 0) {
      next[j] = head[parent[j]];
      head[parent[j]] = j;
    }
  }
  for (int j = 0; j < n; ++j)
    if (parent[j] < 0) roots.push_back(j);
  std::vector<int> order;
  order.reserve(n);
  std::vector<int> stack, cursor(n, -2);
  for (const int root : roots) {
    stack.push_back(root);
    cursor[root] = head[root];
    while (!stack.empty()) {
      const int v = stack.back();
      const int c = cursor[v];
      if (c >= 0) {
        cursor[v] = next[c];
        stack.push_back(c);
        cursor[c] = head[c];
      } else {
        order.push_back(v);
        stack.pop_back();
      }
    }
  }
  if (static_cast<int>(order.size()) != n) throw std::logic_error("elimination tree is not a forest");
  return order;
}

}  // namespace

FactorLayout factor_layout(const Csr& a, const Factor& f) {
  const int n = f.n;
  if (a.n != n) throw std::logic_error("factor and matrix sizes differ");
  std::vector<int> pos(n);
  for (int i = 0; i < n; ++i) pos[f.permutation[i]] = i;
  const std::vector<int> parent = elimination_tree(a, f.permutation, pos);
  const std::vector<int> order = postorder(parent);  // order[k] = elimination index
  std::vector<int> size(n, 1), children(n, 0);
  for (int i = 0; i < n; ++i) {
    if (parent[i] >= 0) {
      size[parent[i]] += size[i];  // children precede parents
      ++children[parent[i]];
    }
  }

  // Slots: postorder positions, each tree starting at a multiple of 4.
  FactorLayout out;
  std::vector<int> slot(n);  // postorder position -> slot
  int next_slot = 0;
  for (int k = 0; k < n;) {
    // The tree holding position k spans [k, k + size(root)); its root is the
    // first position whose node has no parent.
    int end = k;
    while (parent[order[end]] >= 0) ++end;
    next_slot = (next_slot + 3) & ~3;
    for (int j = k; j <= end; ++j) slot[j] = next_slot + (j - k);
    next_slot += end - k + 1;
    k = end + 1;
  }
  out.slots = (next_slot + 3) & ~3;
  std::vector<int> post_of(n);  // elimination index -> postorder position
  for (int k = 0; k < n; ++k) post_of[order[k]] = k;
  out.post.resize(n);
  for (int o = 0; o < n; ++o) out.post[o] = slot[post_of[pos[o]]];

  // Blocks: a whole subtree of at most kAmalgamate vertices (the leaves of
  // the dissection, whose many short chains would otherwise be tiny blocks)
  // is one block over its own column range, zero above the diagonal and left
  // of each row's range; elsewhere a block is one supernode chain.
  for (int k = 0; k < n;) {
    const int k0 = k;
    int k1 = k0 + 1;  // rows [k0, k1)
    int lo_pos = k0 - size[order[k0]] + 1;
    int top = -1;     // the largest subtree starting at k0 that is small enough
    for (int e = order[k0]; post_of[e] - size[e] + 1 == k0 && size[e] <= kAmalgamate; e = parent[e]) {
      top = e;
      if (parent[e] < 0) break;
    }
    if (top >= 0 && size[top] > 1) {
      k1 = k0 + size[top];
      lo_pos = k0;
    } else {
      while (k + 1 < n && parent[order[k]] == order[k + 1] && children[order[k + 1]] == 1) ++k;
      k1 = k + 1;
    }
    k = k1 - 1;
    FactorBlock blk;
    blk.row0 = slot[k0];
    blk.rows = k1 - k0;
    blk.lo = slot[lo_pos];
    blk.col0 = blk.lo & ~3;
    const int end_col = blk.row0 + blk.rows;  // exclusive; rows end at themselves
    blk.quads = (end_col - blk.col0 + 3) / 4;
    blk.offset = static_cast<int64_t>(out.val.size());
    const size_t width = 4 * static_cast<size_t>(blk.quads);
    out.val.resize(out.val.size() + width * blk.rows, 0.0);
    for (int j = 0; j < blk.rows; ++j) {
      const int i = order[k0 + j];
      const int row = blk.row0 + j;
      double* dst = out.val.data() + blk.offset + width * j;
      for (int p = f.s.off[i]; p < f.s.off[i + 1]; ++p) {
        const int c = out.post[f.s.col[p]];
        if (c < blk.lo || c > row) throw std::logic_error("inverse factor entry outside its descendant range");
        dst[c - blk.col0] = f.s.val[p];
      }
      out.entries += size[i];  // the row's own dense range
    }
    out.blocks.push_back(blk);
    k = k1;
  }
  return out;
}

// Longest-processing-time-first assignment (Graham's LPT rule): items by
// decreasing cost, each to the currently least loaded of `bins`; ties go to
// the lower bin and to the lower item.
namespace {
void lpt(const std::vector<int>& items, const std::vector<double>& cost, const std::vector<int>& bins,
         std::vector<int>& owner) {
  std::vector<int> order(items);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  using Load = std::pair<double, int>;
  std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
  for (const int w : bins) heap.push(Load{0.0, w});
  for (const int i : order) {
    Load t = heap.top();
    heap.pop();
    owner[i] = t.second;
    t.first += cost[i];
    heap.push(t);
  }
}
}  // namespace

namespace {
// XCD mode: the bodies onto the 8 XCDs by LPT (workgroup g runs on XCD
// g % 8, so a body's items share one L2), then each XCD's items onto its
// workers by LPT; with the bodies more than 5% from even over the XCDs,
// every item onto any worker.
void xcd_owners(const std::vector<double>& cost, const std::vector<int>& body, int bodies, int workers,
                int per_block, std::vector<int>& owner) {
  const int n = static_cast<int>(cost.size());
  const int blocks = workers / per_block;
  int xcds = blocks % 8 == 0 ? 8 : 1;
  // Bodies onto XCDs; when that leaves the XCDs more than 5% apart (few or
  // very unequal bodies), every item may go to any worker instead.
  std::vector<double> bcost(bodies, 0.0);
  double total = 0.0;
  for (int i = 0; i < n; ++i) {
    bcost[body[i]] += cost[i];
    total += cost[i];
  }
  std::vector<int> all_bodies(bodies), xcd_bins(8), bxcd(bodies, 0);
  for (int b = 0; b < bodies; ++b) all_bodies[b] = b;
  for (int x = 0; x < 8; ++x) xcd_bins[x] = x;
  if (xcds == 8) {
    lpt(all_bodies, bcost, xcd_bins, bxcd);
    std::vector<double> load(8, 0.0);
    for (int b = 0; b < bodies; ++b) load[bxcd[b]] += bcost[b];
    if (*std::max_element(load.begin(), load.end()) > 1.05 * total / 8) xcds = 1;
  }
  if (xcds == 1) std::fill(bxcd.begin(), bxcd.end(), 0);
  // Each XCD's items onto its workers.
  std::vector<std::vector<int>> xitems(xcds), xworkers(xcds);
  for (int i = 0; i < n; ++i) xitems[bxcd[body[i]]].push_back(i);
  for (int w = 0; w < workers; ++w) xworkers[(w / per_block) % xcds].push_back(w);
  for (int x = 0; x < xcds; ++x) lpt(xitems[x], cost, xworkers[x], owner);
}

// Spread mode: each body's items, costliest first, each to a distinct worker
// (rounds of `workers` when a body has more items): the least loaded workers
// so far, ties in an order that cycles over the workgroups (and with them
// the XCDs) before it reuses one. Any subset of the bodies then keeps every
// XCD busy, at the price of a body's inputs being cached in every L2.
void spread_owners(const std::vector<double>& cost, const std::vector<int>& body, int bodies, int workers,
                   int per_block, std::vector<int>& owner) {
  const int blocks = workers / per_block;
  std::vector<int> rank(workers);
  for (int p = 0; p < workers; ++p) rank[(p % blocks) * per_block + p / blocks] = p;
  std::vector<std::vector<int>> items(bodies);
  for (int i = 0; i < static_cast<int>(body.size()); ++i) items[body[i]].push_back(i);
  // Workers by (load, rank), least first: a body's round takes the first
  // ones off the heap and puts them back with their new loads afterwards.
  using Slot = std::pair<double, int>;  // (load, rank)
  std::priority_queue<Slot, std::vector<Slot>, std::greater<Slot>> heap;
  std::vector<int> by_rank(workers);
  for (int w = 0; w < workers; ++w) {
    by_rank[rank[w]] = w;
    heap.push(Slot{0.0, rank[w]});
  }
  std::vector<Slot> taken;
  for (int b = 0; b < bodies; ++b) {
    std::vector<int>& it = items[b];
    std::stable_sort(it.begin(), it.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    for (size_t k0 = 0; k0 < it.size(); k0 += static_cast<size_t>(workers)) {
      const size_t k1 = std::min(it.size(), k0 + static_cast<size_t>(workers));
      taken.clear();
      for (size_t k = k0; k < k1; ++k) {
        Slot t = heap.top();
        heap.pop();
        owner[it[k]] = by_rank[t.second];
        t.first += cost[it[k]];
        taken.push_back(t);
      }
      for (const Slot& t : taken) heap.push(t);
    }
  }
}
}  // namespace

void balanced_lists(const std::vector<double>& cost, const std::vector<int>& body, int bodies, int workers,
                    int per_block, std::vector<int>& off, std::vector<int>& list, bool spread) {
  const int n = static_cast<int>(cost.size());
  if (static_cast<int>(body.size()) != n || workers < 1 || per_block < 1 || workers % per_block != 0)
    throw std::invalid_argument("balanced_lists: inconsistent sizes");
  for (int i = 0; i < n; ++i)
    if (body[i] < 0 || body[i] >= bodies) throw std::invalid_argument("balanced_lists: body out of range");
  std::vector<int> owner(n, 0);
  if (spread) {
    spread_owners(cost, body, bodies, workers, per_block, owner);
  } else {
    xcd_owners(cost, body, bodies, workers, per_block, owner);
  }
  off.assign(workers + 1, 0);
  for (int i = 0; i < n; ++i) ++off[owner[i] + 1];
  for (int w = 0; w < workers; ++w) off[w + 1] += off[w];
  list.assign(n, 0);
  std::vector<int> fill(off.begin(), off.end() - 1);
  for (int i = 0; i < n; ++i) list[fill[owner[i]]++] = i;  // ascending within a worker
}

}  // namespace dphd
