// Host-side setup of the DiffPhD projective-dynamics path: element geometry,
// constant-matrix assembly and the exact sparse inverse factor A^{-1} = S^T S.
//
// Everything here runs once per (timestep, material set) on the host, as in
// the reference (DP/solver.py:491-552, DP/sparse.py:18-191, DP/hex.py:27-92).
// All floating-point code in this translation unit is compiled with
// -ffp-contract=off so every product and sum is rounded exactly where the
// reference's NumPy expressions round it.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace dphd {

// Symmetric matrix in full CSR form (both halves), columns ascending.
struct Csr {
  int n = 0;
  std::vector<int> off;     // n + 1
  std::vector<int> col;
  std::vector<double> val;
};

// Exact inverse factor of one SPD matrix: S = D^{-1/2} L^{-1} P.
// Rows of S are in elimination (permuted) order, columns are original
// indices; S^T is the transpose with rows in original order.
struct Factor {
  int n = 0;
  std::vector<int> permutation;   // permutation[position] = original index
  Csr s;                          // n x n, row = position, col = original
  Csr st;                         // transpose of s
  int64_t factorization_nnz = 0;  // n + strictly lower nnz(L)
  int64_t matrix_nnz = 0;         // nnz(A) after dropping exact zeros
};

// Device layout of one body's inverse factor. Rows and columns are
// renumbered in a postorder of the structural elimination tree of A (taken
// in the factor's elimination order), each tree of the forest starting at a
// multiple of 4 (the unused slots in between are gaps). Row i of L^{-1} is
// nonzero only at the descendants of i, which a postorder numbers
// contiguously, so row r of S is the dense range [lo(r), r]. A chain of rows
// whose parents have a single child (a supernode: the rows of one
// separator, typically) shares lo, so the chain is one dense block: rows
// [row0, row0 + rows) by columns [lo, row0 + rows); a small subtree is one
// block as a whole. A block is stored row-major from the 4-aligned column
// lo & ~3 with `quads` quads (4 columns) per row. Entries the factor does not
// hold (above the diagonal, left of a row's range, alignment, exact zeros of
// the factor) are zero.
struct FactorBlock {
  int row0 = 0, rows = 0;  // postorder rows
  int lo = 0;              // first column of every row (postorder)
  int col0 = 0;            // lo & ~3: column of the first stored value
  int quads = 0;           // stored quads per row
  int64_t offset = 0;      // first value
};

struct FactorLayout {
  std::vector<int> post;  // original local index -> postorder slot
  int slots = 0;          // postorder slots including gaps (a multiple of 4)
  std::vector<FactorBlock> blocks;  // ascending row0
  std::vector<double> val;
  int64_t entries = 0;    // dense-range entries (sum of row lengths)
};

// Throws std::logic_error if a factor entry falls outside the structure.
FactorLayout factor_layout(const Csr& a, const Factor& f);

// Static work lists of the factor products. Item i (cost[i], of body[i])
// goes to one of `workers` workers, `per_block` workers to a workgroup.
// Workgroups are dispatched round-robin over the 8 XCDs (workgroup g runs on
// XCD g % 8), so when the workgroup count is a multiple of 8 the bodies are
// first spread over the XCDs (costliest first, each onto the least loaded),
// and then each XCD's items over that XCD's workers the same way: a body's
// items then share one L2. If the bodies do not split over the XCDs within
// 5% of even, all items are spread over all workers instead. With `spread`
// (the lists that run while few bodies are active), each body's items go to
// distinct workers across all XCDs instead, costliest first onto the least
// loaded, so that a lone body still streams through every XCD. Worker w's
// items, ascending, are list[off[w] .. off[w + 1]).
void balanced_lists(const std::vector<double>& cost, const std::vector<int>& body, int bodies, int workers,
                    int per_block, std::vector<int>& off, std::vector<int>& list, bool spread = false);

// Validates like DP/sparse.py:73-95 (finite, positive diagonal, exactly
// symmetric) and drops exact zeros. Throws std::runtime_error.
Csr validate_rows(const Csr& a);

// BFS-level nested dissection with minimum-degree leaves (DP/sparse.py:18-70).
std::vector<int> nested_dissection(const Csr& a);

// Full factor build (DP/sparse.py:291-320). Throws std::runtime_error.
Factor build_factor(const Csr& a);

// ---------------------------------------------------------------------------
// Element geometry.

// Tet4 gradients from inverse rest edge matrices (DP/solver.py:214-223):
// g0 = -(P0 + P1 + P2), gk = P(k-1) (rows), volume = |1/det P| / 6.
// poses: E x 3 x 3 row-major float32. grads: E x 4 x 3. volumes: E.
void tet_geometry(int elements, const float* poses, double* grads, double* volumes);

// Hex8 shared Gauss gradients for a regular cube of width dx (DP/hex.py:63-80),
// out: 8 samples x 8 corners x 3.
void hex_gradients(double dx, double* out);

// Validates binary-order regular cubes and returns the mean volume
// (DP/hex.py:30-60). rest: vertices x 3, indices: E x 8 (global ids into rest).
double hex_mean_volume(int elements, const int* indices, const double* rest, std::string* error);

// Vertex incidence sorted by (element, corner): offsets[n+1], entries[E*K]
// with entry = element * K + corner (DP/kernels.py:34-72, DP/hex.py:83-92).
void build_incidence(int vertices, int elements, int corners, const int* indices,
                     std::vector<int>& offsets, std::vector<int>& entries);

// ---------------------------------------------------------------------------
// Constant matrix of one body (DP/solver.py:521-545). Indices are local to
// the body. sample_grads: per element (Tet4: 1 sample x 4 corners, element
// specific) or shared (Hex8: 8 samples x 8 corners). weights: per element,
// the summed energy weight (corotated + volume + neo-Hookean). Optional
// fibres: direction (E x 3) and weight (E) per element; zero weight skips.
struct AssemblyInput {
  int vertices = 0;
  int elements = 0;
  int corners = 4;                   // 4 or 8
  int samples = 1;                   // 1 or 8
  const int* indices = nullptr;      // E x corners, local ids
  const double* grads = nullptr;     // Tet4: E x 4 x 3; Hex8: 8 x 8 x 3 shared
  bool shared_grads = false;
  const double* weights = nullptr;   // E
  const double* fibre_dir = nullptr; // E x 3 or null
  const double* fibre_w = nullptr;   // E or null
  const uint8_t* free = nullptr;     // vertices
  double inertia = 0.0;              // m / h^2
  const int* plane_vertices = nullptr;  // planar vertex energies (Hex8 only)
  const double* plane_stiffness = nullptr;
  int plane_count = 0;
};

Csr assemble_matrix(const AssemblyInput& in);

}  // namespace dphd
