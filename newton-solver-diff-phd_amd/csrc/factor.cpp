// Exact sparse inverse factor, constant-matrix assembly and element geometry.
//
// Reference: DP/sparse.py:18-191 (ordering, LDL^T, inverse rows, CSR),
// DP/solver.py:214-223 and 491-552 (Tet4 geometry, assembly), DP/hex.py:27-92
// (Hex8 geometry), DP/kernels.py:34-72 (incidence). The integer outputs
// (permutation, CSR structure, incidence) are bit-exact with the reference
// rules; values follow the reference's operation order with no contraction.
#include "setup.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace dphd {

// ---------------------------------------------------------------------------
// Validation (DP/sparse.py:73-95).

Csr validate_rows(const Csr& a) {
  const int n = a.n;
  Csr out;
  out.n = n;
  out.off.assign(n + 1, 0);
  out.col.reserve(a.col.size());
  out.val.reserve(a.val.size());
  for (int i = 0; i < n; ++i) {
    double diag = 0.0;
    for (int p = a.off[i]; p < a.off[i + 1]; ++p) {
      const int c = a.col[p];
      const double v = a.val[p];
      if (c < 0 || c >= n) throw std::runtime_error("matrix column indices must be integers in [0, size)");
      if (p > a.off[i] && a.col[p - 1] >= c) throw std::runtime_error("matrix columns must be strictly ascending");
      if (!std::isfinite(v)) throw std::runtime_error("matrix coefficients must be finite");
      if (v != 0.0) {
        out.col.push_back(c);
        out.val.push_back(v);
        if (c == i) diag = v;
      }
    }
    if (!(diag > 0.0)) throw std::runtime_error("matrix must be positive definite with positive diagonal entries");
    out.off[i + 1] = static_cast<int>(out.col.size());
  }
  // Exact symmetry: every (i, j) has an identical (j, i).
  for (int i = 0; i < n; ++i) {
    for (int p = out.off[i]; p < out.off[i + 1]; ++p) {
      const int j = out.col[p];
      const int* b = out.col.data() + out.off[j];
      const int* e = out.col.data() + out.off[j + 1];
      const int* f = std::lower_bound(b, e, i);
      if (f == e || *f != i || out.val[f - out.col.data()] != out.val[p])
        throw std::runtime_error("matrix_rows must contain both halves of an exactly symmetric matrix");
    }
  }
  return out;
}

// ---------------------------------------------------------------------------
// Ordering (DP/sparse.py:18-70).

namespace {

class Dissection {
 public:
  explicit Dissection(const Csr& a) : n_(a.n), mark_(a.n, 0), dist_(a.n, -1) {
    // Adjacency without self loops (DP/sparse.py:295).
    off_.assign(n_ + 1, 0);
    for (int i = 0; i < n_; ++i) {
      int count = 0;
      for (int p = a.off[i]; p < a.off[i + 1]; ++p) count += (a.col[p] != i);
      off_[i + 1] = off_[i] + count;
    }
    adj_.resize(off_[n_]);
    for (int i = 0; i < n_; ++i) {
      int q = off_[i];
      for (int p = a.off[i]; p < a.off[i + 1]; ++p)
        if (a.col[p] != i) adj_[q++] = a.col[p];
    }
  }

  std::vector<int> run() {
    std::vector<int> all(n_);
    std::iota(all.begin(), all.end(), 0);
    std::vector<int> order;
    order.reserve(n_);
    visit(all, order);
    return order;
  }

 private:
  int n_;
  std::vector<int> off_, adj_;
  std::vector<int> mark_;
  int stamp_ = 0;
  std::vector<int> dist_;

  int new_stamp(const std::vector<int>& vertices) {
    const int s = ++stamp_;
    for (int v : vertices) mark_[v] = s;
    return s;
  }

  // Breadth-first distances inside the marked set. Shortest-path lengths do
  // not depend on the visiting order, so these equal _distances exactly.
  void bfs(int start, int s, std::vector<int>& reached) {
    for (int v : reached) dist_[v] = -1;
    reached.clear();
    reached.push_back(start);
    dist_[start] = 0;
    for (size_t head = 0; head < reached.size(); ++head) {
      const int node = reached[head];
      for (int p = off_[node]; p < off_[node + 1]; ++p) {
        const int other = adj_[p];
        if (mark_[other] == s && dist_[other] < 0) {
          dist_[other] = dist_[node] + 1;
          reached.push_back(other);
        }
      }
    }
  }

  void clear_dist(const std::vector<int>& reached) {
    for (int v : reached) dist_[v] = -1;
  }

  // _minimum_degree: eliminate the vertex of least remaining degree (ties by
  // index), connecting its neighbours into a clique.
  void minimum_degree(const std::vector<int>& vertices, std::vector<int>& order) {
    const int m = static_cast<int>(vertices.size());
    if (m == 0) return;
    const int s = new_stamp(vertices);
    // Local ids follow ascending global ids (vertices are sorted), so ties
    // by local id are ties by global id.
    std::vector<int> local(m);
    std::vector<std::vector<int>> graph(m);
    for (int a = 0; a < m; ++a) local[a] = vertices[a];
    auto to_local = [&](int v) {
      return static_cast<int>(std::lower_bound(vertices.begin(), vertices.end(), v) - vertices.begin());
    };
    for (int a = 0; a < m; ++a) {
      const int v = vertices[a];
      for (int p = off_[v]; p < off_[v + 1]; ++p)
        if (mark_[adj_[p]] == s) graph[a].push_back(to_local(adj_[p]));
      std::sort(graph[a].begin(), graph[a].end());
    }
    std::vector<char> alive(m, 1);
    std::vector<int> merged;
    for (int step = 0; step < m; ++step) {
      int node = -1;
      size_t best = SIZE_MAX;
      for (int a = 0; a < m; ++a)
        if (alive[a] && graph[a].size() < best) {
          best = graph[a].size();
          node = a;
        }
      alive[node] = 0;
      order.push_back(local[node]);
      std::vector<int> neighbors;
      neighbors.swap(graph[node]);
      for (int other : neighbors) {
        std::vector<int>& g = graph[other];
        auto it = std::lower_bound(g.begin(), g.end(), node);
        if (it != g.end() && *it == node) g.erase(it);
        merged.clear();
        // g |= neighbors - {other}
        size_t i = 0, j = 0;
        while (i < g.size() || j < neighbors.size()) {
          int x;
          if (j == neighbors.size() || (i < g.size() && g[i] < neighbors[j])) {
            x = g[i++];
          } else if (i == g.size() || neighbors[j] < g[i]) {
            x = neighbors[j++];
            if (x == other) continue;
          } else {
            x = g[i++];
            ++j;
          }
          merged.push_back(x);
        }
        g.swap(merged);
      }
    }
  }

  void visit(const std::vector<int>& vertices, std::vector<int>& order) {
    if (vertices.size() <= 24) {
      minimum_degree(vertices, order);
      return;
    }
    int s = new_stamp(vertices);
    // start = argmin (degree inside the set, index).
    int start = -1;
    int best = INT_MAX;
    for (int v : vertices) {
      int degree = 0;
      for (int p = off_[v]; p < off_[v + 1]; ++p) degree += (mark_[adj_[p]] == s);
      if (degree < best) {
        best = degree;
        start = v;
      }
    }
    std::vector<int> reached;
    bfs(start, s, reached);
    if (reached.size() != vertices.size()) {
      // Disconnected: components in order of their least vertex, the first
      // being the one reached from start.
      std::vector<std::vector<int>> components;
      std::vector<int> first = reached;
      std::sort(first.begin(), first.end());
      clear_dist(reached);
      components.push_back(first);
      std::vector<int> remaining;
      {
        std::vector<char> in_first(0);
        size_t k = 0;
        for (int v : vertices) {
          while (k < first.size() && first[k] < v) ++k;
          if (k < first.size() && first[k] == v) continue;
          remaining.push_back(v);
        }
      }
      while (!remaining.empty()) {
        const int r = new_stamp(remaining);
        std::vector<int> comp;
        bfs(remaining.front(), r, comp);
        clear_dist(comp);
        std::sort(comp.begin(), comp.end());
        std::vector<int> rest;
        size_t k = 0;
        for (int v : remaining) {
          while (k < comp.size() && comp[k] < v) ++k;
          if (k < comp.size() && comp[k] == v) continue;
          rest.push_back(v);
        }
        components.push_back(std::move(comp));
        remaining.swap(rest);
      }
      for (const auto& c : components) visit(c, order);
      return;
    }
    // farthest = argmax (distance, -index).
    int farthest = -1;
    int far_dist = -1;
    for (int v : vertices)
      if (dist_[v] > far_dist) {
        far_dist = dist_[v];
        farthest = v;
      }
    bfs(farthest, s, reached);
    int max_level = 0;
    for (int v : vertices) max_level = std::max(max_level, dist_[v]);
    std::vector<long long> levels(max_level + 1, 0);
    for (int v : vertices) ++levels[dist_[v]];
    // cut = searchsorted(cumsum(levels), size / 2), side "left".
    const double half = static_cast<double>(vertices.size()) / 2.0;
    long long running = 0;
    int cut = max_level + 1;
    for (int l = 0; l <= max_level; ++l) {
      running += levels[l];
      if (static_cast<double>(running) >= half) {
        cut = l;
        break;
      }
    }
    std::vector<int> lower, upper, separator;
    for (int v : vertices) {
      const int d = dist_[v];
      if (d < cut)
        lower.push_back(v);
      else if (d > cut)
        upper.push_back(v);
      else
        separator.push_back(v);
    }
    clear_dist(reached);
    if (lower.empty() || upper.empty()) {
      minimum_degree(vertices, order);
      return;
    }
    visit(lower, order);
    visit(upper, order);
    minimum_degree(separator, order);
  }
};

}  // namespace

std::vector<int> nested_dissection(const Csr& a) {
  Dissection d(a);
  return d.run();
}

// ---------------------------------------------------------------------------
// LDL^T and inverse rows (DP/sparse.py:98-191).

Factor build_factor(const Csr& input) {
  const Csr a = validate_rows(input);
  const int n = a.n;
  Factor f;
  f.n = n;
  f.matrix_nnz = static_cast<int64_t>(a.col.size());
  f.permutation = nested_dissection(a);
  std::vector<int> inverse(n);
  for (int i = 0; i < n; ++i) inverse[f.permutation[i]] = i;

  // Permuted lower triangle of A: row i holds columns <= i.
  std::vector<int> loff(n + 1, 0), lcol;
  std::vector<double> lval;
  for (int i = 0; i < n; ++i) {
    const int orig = f.permutation[i];
    std::vector<std::pair<int, double>> row;
    for (int p = a.off[orig]; p < a.off[orig + 1]; ++p) {
      const int j = inverse[a.col[p]];
      if (j <= i) row.emplace_back(j, a.val[p]);
    }
    std::sort(row.begin(), row.end());
    for (auto& e : row) {
      lcol.push_back(e.first);
      lval.push_back(e.second);
    }
    loff[i + 1] = static_cast<int>(lcol.size());
  }

  // Symbolic: elimination tree and row patterns of L (the fill of
  // right-looking elimination, DP/sparse.py:106-123).
  std::vector<int> parent(n, -1), ancestor(n, -1);
  for (int i = 0; i < n; ++i) {
    for (int p = loff[i]; p < loff[i + 1]; ++p) {
      int r = lcol[p];
      if (r >= i) continue;
      while (ancestor[r] != -1 && ancestor[r] != i) {
        const int t = ancestor[r];
        ancestor[r] = i;
        r = t;
      }
      if (ancestor[r] == -1) {
        ancestor[r] = i;
        parent[r] = i;
      }
    }
  }
  std::vector<int> mark(n, -1);
  std::vector<int> poff(n + 1, 0), pcol;  // strictly lower pattern of L, ascending
  {
    std::vector<int> stack;
    for (int i = 0; i < n; ++i) {
      stack.clear();
      mark[i] = i;
      for (int p = loff[i]; p < loff[i + 1]; ++p) {
        int r = lcol[p];
        while (r < i && mark[r] != i) {
          mark[r] = i;
          stack.push_back(r);
          r = parent[r];
          if (r < 0) break;
        }
      }
      std::sort(stack.begin(), stack.end());
      pcol.insert(pcol.end(), stack.begin(), stack.end());
      poff[i + 1] = static_cast<int>(pcol.size());
    }
  }
  const int64_t lnnz = static_cast<int64_t>(pcol.size());
  f.factorization_nnz = n + lnnz;

  // Column lists of the pattern with the unscaled entry of each (j, k):
  // the value the right-looking sweep holds at pivot k.
  std::vector<int> ccount(n + 1, 0);
  for (int64_t p = 0; p < lnnz; ++p) ++ccount[pcol[p] + 1];
  std::vector<int> coff(n + 1, 0);
  for (int k = 0; k < n; ++k) coff[k + 1] = coff[k] + ccount[k + 1];
  std::vector<int> crow(lnnz);
  std::vector<double> cunscaled(lnnz);
  std::vector<int> cfill(coff.begin(), coff.end() - 1);

  std::vector<double> multiplier(lnnz);
  std::vector<double> diagonal(n);
  std::vector<double> work(n, 0.0);
  for (int i = 0; i < n; ++i) {
    for (int p = poff[i]; p < poff[i + 1]; ++p) work[pcol[p]] = 0.0;
    work[i] = 0.0;
    for (int p = loff[i]; p < loff[i + 1]; ++p) work[lcol[p]] = lval[p];
    // Updates reach each entry in ascending pivot order with the reference's
    // multiply-then-subtract rounding (DP/sparse.py:145-151).
    for (int p = poff[i]; p < poff[i + 1]; ++p) {
      const int k = pcol[p];
      const double entry_ik = work[k];
      const double m = entry_ik / diagonal[k];
      for (int q = coff[k]; q < cfill[k]; ++q) {
        const int j = crow[q];  // j in (k, i)
        work[j] -= m * cunscaled[q];
      }
      work[i] -= m * entry_ik;
      multiplier[p] = m;
    }
    const double pivot = work[i];
    if (!std::isfinite(pivot) || pivot <= 0.0)
      throw std::runtime_error("matrix is not positive definite: LDLT pivot " + std::to_string(i) + " is " +
                               std::to_string(pivot));
    diagonal[i] = pivot;
    for (int p = poff[i]; p < poff[i + 1]; ++p) {
      const int k = pcol[p];
      crow[cfill[k]] = i;
      cunscaled[cfill[k]] = work[k];
      ++cfill[k];
    }
  }

  // Rows of L^{-1} (unscaled), then S = D^{-1/2} L^{-1} P (DP/sparse.py:159-177).
  std::vector<int64_t> roff(n + 1, 0);
  std::vector<int> rcol;
  std::vector<double> rval;
  std::vector<double> current(n, 0.0);
  std::vector<std::vector<std::pair<int, double>>> srows(n);
  for (int i = 0; i < n; ++i) {
    std::fill(current.begin(), current.begin() + i + 1, 0.0);
    current[i] = 1.0;
    for (int p = poff[i]; p < poff[i + 1]; ++p) {
      const int k = pcol[p];
      const double coefficient = multiplier[p];
      for (int64_t q = roff[k]; q < roff[k + 1]; ++q) current[rcol[q]] -= coefficient * rval[q];
    }
    const double scale = 1.0 / std::sqrt(diagonal[i]);
    auto& row = srows[i];
    for (int c = 0; c <= i; ++c) {
      if (current[c] != 0.0) {
        rcol.push_back(c);
        rval.push_back(current[c]);
        row.emplace_back(f.permutation[c], current[c] * scale);
      }
    }
    roff[i + 1] = static_cast<int64_t>(rcol.size());
    std::sort(row.begin(), row.end());
  }
  const int64_t snnz = static_cast<int64_t>(rcol.size());
  if (snnz > INT_MAX) throw std::runtime_error("inverse factor fill exceeds int32 sparse indexing");

  f.s.n = n;
  f.s.off.assign(n + 1, 0);
  f.s.col.reserve(snnz);
  f.s.val.reserve(snnz);
  for (int i = 0; i < n; ++i) {
    for (auto& e : srows[i]) {
      f.s.col.push_back(e.first);
      f.s.val.push_back(e.second);
    }
    f.s.off[i + 1] = static_cast<int>(f.s.col.size());
    std::vector<std::pair<int, double>>().swap(srows[i]);
  }
  for (double v : f.s.val)
    if (!std::isfinite(v)) throw std::runtime_error("inverse factor overflowed float64");

  // Transpose: row = original column, entries ascending in S row index.
  f.st.n = n;
  f.st.off.assign(n + 1, 0);
  for (int c : f.s.col) ++f.st.off[c + 1];
  for (int c = 0; c < n; ++c) f.st.off[c + 1] += f.st.off[c];
  f.st.col.resize(snnz);
  f.st.val.resize(snnz);
  std::vector<int> fill(f.st.off.begin(), f.st.off.end() - 1);
  for (int i = 0; i < n; ++i)
    for (int p = f.s.off[i]; p < f.s.off[i + 1]; ++p) {
      const int c = f.s.col[p];
      f.st.col[fill[c]] = i;
      f.st.val[fill[c]] = f.s.val[p];
      ++fill[c];
    }
  return f;
}

// ---------------------------------------------------------------------------
// Geometry.

void tet_geometry(int elements, const float* poses, double* grads, double* volumes) {
  for (int e = 0; e < elements; ++e) {
    double p[9];
    for (int t = 0; t < 9; ++t) p[t] = static_cast<double>(poses[9 * e + t]);
    double* g = grads + 12 * e;
    for (int c = 0; c < 3; ++c) {
      g[c] = -((p[c] + p[3 + c]) + p[6 + c]);
      g[3 + c] = p[c];
      g[6 + c] = p[3 + c];
      g[9 + c] = p[6 + c];
    }
    const double det = p[0] * (p[4] * p[8] - p[5] * p[7]) - p[1] * (p[3] * p[8] - p[5] * p[6]) +
                       p[2] * (p[3] * p[7] - p[4] * p[6]);
    volumes[e] = std::fabs(1.0 / det) / 6.0;
  }
}

void hex_gradients(double dx, double* out) {
  const double inv_dx = 1.0 / dx;
  const double r = 1.0 / std::sqrt(3.0);
  for (int s = 0; s < 8; ++s) {
    const int bits[3] = {(s >> 2) & 1, (s >> 1) & 1, s & 1};
    double point[3];
    for (int a = 0; a < 3; ++a) point[a] = ((static_cast<double>(bits[a]) - 0.5) * r + 0.5) * dx;
    const double nx = point[0] * inv_dx, ny = point[1] * inv_dx, nz = point[2] * inv_dx;
    const double cnx = 1.0 - nx, cny = 1.0 - ny, cnz = 1.0 - nz;
    const double g[8][3] = {
        {-inv_dx * cny * cnz, cnx * -inv_dx * cnz, cnx * cny * -inv_dx},
        {-inv_dx * cny * nz, cnx * -inv_dx * nz, cnx * cny * inv_dx},
        {-inv_dx * ny * cnz, cnx * inv_dx * cnz, cnx * ny * -inv_dx},
        {-inv_dx * ny * nz, cnx * inv_dx * nz, cnx * ny * inv_dx},
        {inv_dx * cny * cnz, nx * -inv_dx * cnz, nx * cny * -inv_dx},
        {inv_dx * cny * nz, nx * -inv_dx * nz, nx * cny * inv_dx},
        {inv_dx * ny * cnz, nx * inv_dx * cnz, nx * ny * -inv_dx},
        {inv_dx * ny * nz, nx * inv_dx * nz, nx * ny * inv_dx},
    };
    for (int j = 0; j < 8; ++j)
      for (int a = 0; a < 3; ++a) out[(s * 8 + j) * 3 + a] = g[j][a];
  }
}

double hex_mean_volume(int elements, const int* indices, const double* rest, std::string* error) {
  if (elements <= 0) {
    if (error) *error = "hex_mesh must contain elements";
    return 0.0;
  }
  std::vector<double> volumes(elements);
  double total = 0.0;
  for (int e = 0; e < elements; ++e) {
    const int* nodes = indices + 8 * e;
    const double* o = rest + 3 * nodes[0];
    double ed[3][3];
    const int pick[3] = {1, 2, 4};
    for (int k = 0; k < 3; ++k)
      for (int a = 0; a < 3; ++a) ed[k][a] = rest[3 * nodes[pick[k]] + a] - o[a];
    const double cx = ed[1][1] * ed[2][2] - ed[1][2] * ed[2][1];
    const double cy = ed[1][2] * ed[2][0] - ed[1][0] * ed[2][2];
    const double cz = ed[1][0] * ed[2][1] - ed[1][1] * ed[2][0];
    volumes[e] = std::fabs(ed[0][0] * cx + ed[0][1] * cy + ed[0][2] * cz);
    total += volumes[e];
  }
  const double mean = total / elements;
  if (!std::isfinite(mean) || mean <= 0.0) {
    if (error) *error = "Hex8 cells must have positive, finite rest volume";
    return 0.0;
  }
  for (double v : volumes)
    if (!(v > 0.0)) {
      if (error) *error = "Hex8 cells must have positive, finite rest volume";
      return 0.0;
    }
  const double dx = std::pow(mean, 1.0 / 3.0);
  const double tol = 1.0e-9 * dx;
  for (int e = 0; e < elements; ++e) {
    const int* nodes = indices + 8 * e;
    const double* o = rest + 3 * nodes[0];
    for (int c = 0; c < 8; ++c) {
      const int bits[3] = {(c >> 2) & 1, (c >> 1) & 1, c & 1};
      for (int a = 0; a < 3; ++a) {
        const double expected = o[a] + dx * static_cast<double>(bits[a]);
        if (!(std::fabs(rest[3 * nodes[c] + a] - expected) <= tol)) {
          if (error) *error = "Reference Hex8 requires equally sized axis-aligned cubes in binary corner order";
          return 0.0;
        }
      }
    }
  }
  return mean;
}

void build_incidence(int vertices, int elements, int corners, const int* indices, std::vector<int>& offsets,
                     std::vector<int>& entries) {
  offsets.assign(vertices + 1, 0);
  const int64_t total = static_cast<int64_t>(elements) * corners;
  for (int64_t t = 0; t < total; ++t) ++offsets[indices[t] + 1];
  for (int v = 0; v < vertices; ++v) offsets[v + 1] += offsets[v];
  entries.assign(total, 0);
  std::vector<int> cursor(offsets.begin(), offsets.end() - 1);
  // Ascending entry order per vertex: the stable grouping of the reference.
  for (int64_t t = 0; t < total; ++t) entries[cursor[indices[t]]++] = static_cast<int>(t);
}

// ---------------------------------------------------------------------------
// Assembly (DP/solver.py:521-545).

Csr assemble_matrix(const AssemblyInput& in) {
  const int n = in.vertices;
  const int K = in.corners;
  std::vector<std::vector<std::pair<int, double>>> rows(n);
  for (int i = 0; i < n; ++i) rows[i].emplace_back(i, in.free[i] ? in.inertia : 1.0);
  auto add = [&](int i, int j, double value) {
    auto& r = rows[i];
    for (auto& e : r)
      if (e.first == j) {
        e.second = e.second + value;
        return;
      }
    r.emplace_back(j, 0.0 + value);
  };
  double block[8][8];
  for (int e = 0; e < in.elements; ++e) {
    const int* nodes = in.indices + static_cast<int64_t>(e) * K;
    for (int s = 0; s < in.samples; ++s) {
      const double* g = in.shared_grads ? in.grads + static_cast<int64_t>(s) * K * 3
                                        : in.grads + static_cast<int64_t>(e) * K * 3;
      const double w = in.weights[e];
      for (int a = 0; a < K; ++a)
        for (int b = 0; b < K; ++b) {
          const double dot = (g[3 * a] * g[3 * b] + g[3 * a + 1] * g[3 * b + 1]) + g[3 * a + 2] * g[3 * b + 2];
          block[a][b] = w * dot;
        }
      if (in.fibre_w && in.fibre_w[e] > 0.0) {
        const double* m = in.fibre_dir + 3 * static_cast<int64_t>(e);
        double fibre[8];
        for (int a = 0; a < K; ++a) fibre[a] = (g[3 * a] * m[0] + g[3 * a + 1] * m[1]) + g[3 * a + 2] * m[2];
        for (int a = 0; a < K; ++a)
          for (int b = 0; b < K; ++b) block[a][b] = block[a][b] + in.fibre_w[e] * (fibre[a] * fibre[b]);
      }
      for (int a = 0; a < K; ++a) {
        const int i = nodes[a];
        if (!in.free[i]) continue;
        for (int b = a; b < K; ++b) {
          const int j = nodes[b];
          if (!in.free[j]) continue;
          add(i, j, block[a][b]);
          if (i != j) add(j, i, block[a][b]);
        }
      }
    }
  }
  for (int p = 0; p < in.plane_count; ++p) {
    const int i = in.plane_vertices[p];
    if (in.free[i]) add(i, i, in.plane_stiffness[p]);
  }
  Csr a;
  a.n = n;
  a.off.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    auto& r = rows[i];
    std::sort(r.begin(), r.end());
    for (auto& e : r) {
      a.col.push_back(e.first);
      a.val.push_back(e.second);
    }
    a.off[i + 1] = static_cast<int>(a.col.size());
  }
  return a;
}

}  // namespace dphd
