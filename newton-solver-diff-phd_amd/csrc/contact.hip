// gfx950 kernels of the mesh contact state force (reference:
// src/newton_solver_diff_phd/_src/solvers/diff_phd/mesh_contact.py:34-179).
//
//   k_mc_grid      per direction, the grid over the other side's faces (their
//                  boxes widened by the contact radius) and whether every
//                  position is finite
//   k_mc_bin / k_mc_scan / k_mc_rank
//                  faces and vertices binned into the cells (counting sort),
//                  each cell's lists ascending
//   k_mc_gather    _gather_forces (:140-170) + add_state_force (:173-177):
//                  a particle's own pairs (_pair_forces, :90-137) over its
//                  cell's faces in face order, then the reactions on the faces
//                  it is a corner of over their cells' vertices in vertex order
//
// Pair forces are computed with floating-point contraction off, so the hit
// test and the force are those of the reference's arithmetic and every sum
// has the reference's order (pairs outside the radius contribute exact zeros
// there and are skipped).
#include "contact.h"
#include "pd_math.h"

#include <climits>

namespace dphd {
namespace {

constexpr double kEps = 2.220446049250313e-16;

struct Vec3Src {
  const float* f32;
  const double* f64;
  __device__ __forceinline__ V3 operator()(int i) const {
    if (f64) return V3{f64[3 * i], f64[3 * i + 1], f64[3 * i + 2]};
    return V3{static_cast<double>(f32[3 * i]), static_cast<double>(f32[3 * i + 1]),
              static_cast<double>(f32[3 * i + 2])};
  }
};

struct PairForce {
  V3 force;
  V3 bary;
  bool hit;
};

// ClosestPointTriangle (mesh_contact.py:34-87): closest point, barycentric
// coordinates and squared distance.
__device__ __forceinline__ void closest_point_triangle(V3 p, V3 a, V3 b, V3 c, V3& point, V3& bary, double& dist2) {
#pragma clang fp contract(off)
  const V3 ab = b - a, ac = c - a, ap = p - a;
  const double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) {
    point = a;
    bary = V3{1.0, 0.0, 0.0};
    dist2 = dot(p - a, p - a);
    return;
  }
  const V3 bp = p - b;
  const double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) {
    point = b;
    bary = V3{0.0, 1.0, 0.0};
    dist2 = dot(p - b, p - b);
    return;
  }
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
    const double v = d1 / (d1 - d3);
    point = a + v * ab;
    bary = V3{1.0 - v, v, 0.0};
    dist2 = dot(p - point, p - point);
    return;
  }
  const V3 cp = p - c;
  const double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) {
    point = c;
    bary = V3{0.0, 0.0, 1.0};
    dist2 = dot(p - c, p - c);
    return;
  }
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
    const double w = d2 / (d2 - d6);
    point = a + w * ac;
    bary = V3{1.0 - w, 0.0, w};
    dist2 = dot(p - point, p - point);
    return;
  }
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    point = b + w * (c - b);
    bary = V3{0.0, 1.0 - w, w};
    dist2 = dot(p - point, p - point);
    return;
  }
  const double denom = 1.0 / (va + vb + vc);
  const double v = vb * denom, w = vc * denom;
  point = (a + v * ab) + w * ac;
  bary = V3{(1.0 - v) - w, v, w};
  dist2 = dot(p - point, p - point);
}

__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
#pragma clang fp contract(off)
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// _pair_forces (mesh_contact.py:90-137) for vertex `vid` against one face.
__device__ __forceinline__ PairForce pair_force(const Vec3Src& pos, const Vec3Src& vel, int vid, const int* face,
                                                double radius, double kn, double kf, double mu) {
#pragma clang fp contract(off)
  PairForce out{V3{0, 0, 0}, V3{0, 0, 0}, false};
  const V3 p = pos(vid), a = pos(face[0]), b = pos(face[1]), c = pos(face[2]);
  V3 closest, w;
  double dist2;
  closest_point_triangle(p, a, b, c, closest, w, dist2);
  if (dist2 >= radius * radius) return out;
  const double dist = sqrt(dist2);
  V3 n;
  if (dist > kEps) {
    const V3 d = p - closest;
    n = V3{d.x / dist, d.y / dist, d.z / dist};
  } else {
    n = cross3(b - a, c - a);
    const double n_norm = sqrt(dot(n, n));
    if (n_norm <= kEps) return out;
    n = V3{n.x / n_norm, n.y / n_norm, n.z / n_norm};
  }
  if (dot(n, p - closest) < 0.0) n = -1.0 * n;
  const double fn_mag = kn * (radius - dist);
  const V3 fn = fn_mag * n;
  const V3 vtri = (w.x * vel(face[0]) + w.y * vel(face[1])) + w.z * vel(face[2]);
  const V3 vrel = vel(vid) - vtri;
  const V3 vrel_t = vrel - dot(vrel, n) * n;
  const double vrel_t_mag = sqrt(dot(vrel_t, vrel_t));
  V3 ff{0, 0, 0};
  if (vrel_t_mag > kEps) ff = (-fmin(kf, mu * fn_mag / (vrel_t_mag + kEps))) * vrel_t;
  out.force = fn + ff;
  out.bary = w;
  out.hit = true;
  return out;
}

__global__ void k_mc_zero(int* pairs) {
  pairs[0] = 0;
  pairs[1] = 0;
}

__device__ __forceinline__ bool finite3(V3 v) { return isfinite(v.x) && isfinite(v.y) && isfinite(v.z); }

// Cell range of one axis for the interval [lo, hi] (clamped to the grid).
__device__ __forceinline__ int cell_of(double x, double origin, double h, int dims) {
  const double t = floor((x - origin) / h);
  return t < 0.0 ? 0 : (t >= dims ? dims - 1 : static_cast<int>(t));
}

// The grid of one direction (one workgroup): the bounds of the other side's
// face boxes widened by `reach`, the largest widened box edge, and whether
// every face corner and vertex is finite. Cell size: the largest box edge,
// or the bounds' largest edge over kGridDim when that is larger.
__global__ __launch_bounds__(256) void k_mc_grid(MeshContactSide side, Vec3Src pos, double reach) {
  __shared__ double red[7][256];
  __shared__ int bad[256];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY}, edge = 0.0;
  int nonfinite = 0;
  for (int f = threadIdx.x; f < side.nf; f += blockDim.x) {
    double blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < 3; ++k) {
      const V3 v = pos(side.faces[3 * f + k]);
      nonfinite |= !finite3(v);
      const double c[3] = {v.x, v.y, v.z};
      for (int d = 0; d < 3; ++d) {
        blo[d] = fmin(blo[d], c[d]);
        bhi[d] = fmax(bhi[d], c[d]);
      }
    }
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], blo[d] - reach);
      hi[d] = fmax(hi[d], bhi[d] + reach);
      edge = fmax(edge, (bhi[d] + reach) - (blo[d] - reach));
    }
  }
  for (int s = threadIdx.x; s < side.nv; s += blockDim.x) nonfinite |= !finite3(pos(side.verts[s]));
  for (int d = 0; d < 3; ++d) {
    red[d][threadIdx.x] = lo[d];
    red[3 + d][threadIdx.x] = hi[d];
  }
  red[6][threadIdx.x] = edge;
  bad[threadIdx.x] = nonfinite;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (static_cast<int>(threadIdx.x) < off) {
      for (int d = 0; d < 3; ++d) {
        red[d][threadIdx.x] = fmin(red[d][threadIdx.x], red[d][threadIdx.x + off]);
        red[3 + d][threadIdx.x] = fmax(red[3 + d][threadIdx.x], red[3 + d][threadIdx.x + off]);
      }
      red[6][threadIdx.x] = fmax(red[6][threadIdx.x], red[6][threadIdx.x + off]);
      bad[threadIdx.x] |= bad[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* g = side.grid;
    const bool on = !bad[0] && side.nf > 0;
    double h = red[6][0];
    for (int d = 0; d < 3; ++d) h = fmax(h, (red[3 + d][0] - red[d][0]) / kGridDim);
    // The margin keeps a widened box within two cells per axis whatever the
    // rounding of the cell coordinates (they are at most kGridDim + 1).
    h *= 1.0 + 1.0e-9;
    if (!(h > 0.0)) h = 1.0;
    for (int d = 0; d < 3; ++d) {
      g[d] = on ? red[d][0] : 0.0;
      g[8 + d] = on ? red[3 + d][0] : 0.0;
      const double cells = on ? floor((red[3 + d][0] - red[d][0]) / h) + 1.0 : 1.0;
      g[4 + d] = fmin(cells, static_cast<double>(kGridDim));
    }
    g[3] = h;
    g[7] = on ? 1.0 : 0.0;
  }
}

__global__ __launch_bounds__(256) void k_mc_clear(MeshContactSide side) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c <= kGridCells; c += gridDim.x * blockDim.x) {
    side.fcell_off[c] = 0;
    side.vcell_off[c] = 0;
  }
}

struct CellBox {
  int lo[3], hi[3];
};

// The cells a face's widened box overlaps.
__device__ __forceinline__ CellBox face_cells(const MeshContactSide& side, const Vec3Src& pos, int f, double reach) {
  const double* g = side.grid;
  CellBox cb;
  for (int d = 0; d < 3; ++d) {
    double blo = INFINITY, bhi = -INFINITY;
    for (int k = 0; k < 3; ++k) {
      const V3 v = pos(side.faces[3 * f + k]);
      const double c = d == 0 ? v.x : (d == 1 ? v.y : v.z);
      blo = fmin(blo, c);
      bhi = fmax(bhi, c);
    }
    const int dims = static_cast<int>(g[4 + d]);
    cb.lo[d] = cell_of(blo - reach, g[d], g[3], dims);
    cb.hi[d] = min(cell_of(bhi + reach, g[d], g[3], dims), cb.lo[d] + 1);
  }
  return cb;
}

__device__ __forceinline__ int cell_index(const double* g, int x, int y, int z) {
  return (x * static_cast<int>(g[5]) + y) * static_cast<int>(g[6]) + z;
}

// The cell of a vertex, or -1 when it lies outside the widened boxes' bounds
// (then no face is within the radius).
__device__ __forceinline__ int vertex_cell(const MeshContactSide& side, V3 p) {
  const double* g = side.grid;
  const double c[3] = {p.x, p.y, p.z};
  int idx[3];
  for (int d = 0; d < 3; ++d) {
    if (c[d] < g[d] || c[d] > g[8 + d]) return -1;
    idx[d] = cell_of(c[d], g[d], g[3], static_cast<int>(g[4 + d]));
  }
  return cell_index(g, idx[0], idx[1], idx[2]);
}

// Counts per cell (into *_off[cell + 1]); pass 1 writes the entries at the
// fill positions of the arrival-order lists instead.
__global__ __launch_bounds__(256) void k_mc_bin(MeshContactSide side, Vec3Src pos, double reach, int pass) {
  if (side.grid[7] == 0.0) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < side.nf) {
    const CellBox cb = face_cells(side, pos, t, reach);
    for (int x = cb.lo[0]; x <= cb.hi[0]; ++x)
      for (int y = cb.lo[1]; y <= cb.hi[1]; ++y)
        for (int z = cb.lo[2]; z <= cb.hi[2]; ++z) {
          const int c = cell_index(side.grid, x, y, z);
          if (pass == 0)
            atomicAdd(side.fcell_off + c + 1, 1);
          else
            side.unsorted[atomicAdd(side.cursor + c, 1)] = t;
        }
  } else if (t < side.nf + side.nv) {
    const int s = t - side.nf;
    const int c = vertex_cell(side, pos(side.verts[s]));
    if (pass == 0) {
      side.vcell_of[s] = c;
      if (c >= 0) atomicAdd(side.vcell_off + c + 1, 1);
    } else if (c >= 0) {
      side.unsorted[8 * side.nf + atomicAdd(side.cursor + kGridCells + c, 1)] = s;
    }
  }
}

// Prefix sums of both count arrays (one workgroup; each thread a contiguous
// run of cells, then a scan of the run totals) and the fill positions.
__global__ __launch_bounds__(256) void k_mc_scan(MeshContactSide side) {
  if (side.grid[7] == 0.0) return;
  __shared__ int run[2][256];
  constexpr int per = kGridCells / 256;
  const int c0 = threadIdx.x * per;
  for (int a = 0; a < 2; ++a) {
    int* off = a == 0 ? side.fcell_off : side.vcell_off;
    int total = 0;
    for (int c = c0; c < c0 + per; ++c) total += off[c + 1];
    run[a][threadIdx.x] = total;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    int acc = 0;
    for (int t = 0; t < 256; ++t) {
      const int v = run[threadIdx.x][t];
      run[threadIdx.x][t] = acc;
      acc += v;
    }
  }
  __syncthreads();
  for (int a = 0; a < 2; ++a) {
    int* off = a == 0 ? side.fcell_off : side.vcell_off;
    int acc = run[a][threadIdx.x];
    for (int c = c0; c < c0 + per; ++c) {
      acc += off[c + 1];
      off[c + 1] = acc;
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < kGridCells; c += blockDim.x) {
    side.cursor[c] = side.fcell_off[c];
    side.cursor[kGridCells + c] = side.vcell_off[c];
  }
}

// Each cell's lists ascending: an entry's place is the number of smaller
// entries of its cell (entries of a cell are distinct), one thread per entry.
__global__ __launch_bounds__(256) void k_mc_rank(MeshContactSide side) {
  if (side.grid[7] == 0.0) return;
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  const bool face = j < 8 * side.nf;
  if (!face) j -= 8 * side.nf;
  const int* off = face ? side.fcell_off : side.vcell_off;
  if (j >= off[kGridCells]) return;
  const int* in = side.unsorted + (face ? 0 : 8 * side.nf);
  int lo = 0, hi = kGridCells;  // the cell c with off[c] <= j < off[c + 1]
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (off[mid] <= j)
      lo = mid;
    else
      hi = mid;
  }
  const int v = in[j];
  int rank = 0;
  for (int i = off[lo]; i < off[lo + 1]; ++i) rank += in[i] < v;
  (face ? side.fcell : side.vcell)[off[lo] + rank] = v;
}

// The particle's own pairs (as vertex s of `side`) in face order; counts the hits.
__device__ __forceinline__ V3 own_pairs(const MeshContactSide& side, int s, const Vec3Src& pos, const Vec3Src& vel,
                                        double radius, double kf, double mu, V3 total, int& hits) {
  const bool on = side.grid[7] != 0.0;
  int b = 0, e = side.nf;
  if (on) {
    const int c = side.vcell_of[s];
    if (c < 0) return total;
    b = side.fcell_off[c];
    e = side.fcell_off[c + 1];
  }
  for (int k = b; k < e; ++k) {
    const int f = on ? side.fcell[k] : k;
    const PairForce h = pair_force(pos, vel, side.verts[s], side.faces + 3 * f, radius, side.kn, kf, mu);
    if (h.hit) {
      total = total + h.force;
      ++hits;
    }
  }
  return total;
}

// Reactions on the particle as corner `corner` of face f of `side`'s faces,
// over the side's vertices in order: the vertex lists of the face's cells
// (disjoint, each ascending) merged.
__device__ __forceinline__ V3 face_reactions(const MeshContactSide& side, int f, int corner, const Vec3Src& pos,
                                             const Vec3Src& vel, double radius, double kf, double mu, double reach,
                                             V3 total) {
  if (side.grid[7] == 0.0) {
    for (int s = 0; s < side.nv; ++s) {
      const PairForce h = pair_force(pos, vel, side.verts[s], side.faces + 3 * f, radius, side.kn, kf, mu);
      if (h.hit) total = total - vget(h.bary, corner) * h.force;
    }
    return total;
  }
  const CellBox cb = face_cells(side, pos, f, reach);
  int at[8], end[8], lists = 0;
  for (int x = cb.lo[0]; x <= cb.hi[0]; ++x)
    for (int y = cb.lo[1]; y <= cb.hi[1]; ++y)
      for (int z = cb.lo[2]; z <= cb.hi[2]; ++z) {
        const int c = cell_index(side.grid, x, y, z);
        if (lists < 8 && side.vcell_off[c] < side.vcell_off[c + 1]) {
          at[lists] = side.vcell_off[c];
          end[lists] = side.vcell_off[c + 1];
          ++lists;
        }
      }
  while (true) {
    int pick = -1, best = INT_MAX;
    for (int l = 0; l < lists; ++l)
      if (at[l] < end[l] && side.vcell[at[l]] < best) {
        best = side.vcell[at[l]];
        pick = l;
      }
    if (pick < 0) break;
    ++at[pick];
    const PairForce h = pair_force(pos, vel, side.verts[best], side.faces + 3 * f, radius, side.kn, kf, mu);
    if (h.hit) total = total - vget(h.bary, corner) * h.force;
  }
  return total;
}

__global__ __launch_bounds__(256) void k_mc_gather(MeshContactDev mc, Vec3Src pos, Vec3Src vel, double* target,
                                                   const int* vbody, const double* mass, double dt, double reach) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= mc.n) return;
  V3 total{0, 0, 0};
  int hits_ab = 0, hits_ba = 0;
  if (mc.slot_a[i] >= 0) total = own_pairs(mc.ab, mc.slot_a[i], pos, vel, mc.radius, mc.kf, mc.mu, total, hits_ab);
  if (mc.slot_b[i] >= 0) total = own_pairs(mc.ba, mc.slot_b[i], pos, vel, mc.radius, mc.kf, mc.mu, total, hits_ba);
  if (hits_ab) atomicAdd(mc.pairs, hits_ab);
  if (hits_ba) atomicAdd(mc.pairs + 1, hits_ba);
  // As a corner of a face of B pushed by A's vertices, then of A by B's.
  for (int k = mc.inc_b_off[i]; k < mc.inc_b_off[i + 1]; ++k)
    total = face_reactions(mc.ab, mc.inc_b[2 * k], mc.inc_b[2 * k + 1], pos, vel, mc.radius, mc.kf, mc.mu, reach,
                           total);
  for (int k = mc.inc_a_off[i]; k < mc.inc_a_off[i + 1]; ++k)
    total = face_reactions(mc.ba, mc.inc_a[2 * k], mc.inc_a[2 * k + 1], pos, vel, mc.radius, mc.kf, mc.mu, reach,
                           total);
  mc.force[3 * i] = total.x;
  mc.force[3 * i + 1] = total.y;
  mc.force[3 * i + 2] = total.z;
  if (target) {
    const double c = dt * dt / mass[vbody[i]];
    target[3 * i] = target[3 * i] + c * total.x;
    target[3 * i + 1] = target[3 * i + 1] + c * total.y;
    target[3 * i + 2] = target[3 * i + 2] + c * total.z;
  }
}

int blocks_for(long long threads) { return static_cast<int>((threads + 255) / 256); }

// ---------------------------------------------------------------------------
// Frictional boundary (contact.py).

__device__ __forceinline__ V3 ld3(const double* p, int i) { return V3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ void st3(double* p, int i, V3 v) {
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}

// boundary_gap / boundary_normal (contact.py:72-95).
__device__ __forceinline__ double boundary_gap(const FrictionDev& fc, V3 x) {
#pragma clang fp contract(off)
  const double* q = fc.params;
  if (fc.kind == 0) {
    const V3 axis{q[0], q[1], q[2]};
    const double norm = sqrt(dot(axis, axis));
    return dot(x, V3{axis.x / norm, axis.y / norm, axis.z / norm}) + q[3] / norm;
  }
  const V3 d = x - V3{q[0], q[1], q[2]};
  return sqrt(dot(d, d)) - q[3];
}

__device__ __forceinline__ V3 boundary_normal(const FrictionDev& fc, V3 x) {
#pragma clang fp contract(off)
  const double* q = fc.params;
  V3 v = fc.kind == 0 ? V3{q[0], q[1], q[2]} : x - V3{q[0], q[1], q[2]};
  const double length = sqrt(dot(v, v));
  return V3{v.x / length, v.y / length, v.z / length};
}

// The first face of minimal distance on the surface opposite `node`
// The opposing side of a mesh-boundary vertex: faces [first, last) in words
// [w0, w1). Vertices below `split` touch the right faces.
struct FaceRange {
  int first, last, w0, w1;
};

__device__ __forceinline__ FaceRange opposing_faces(const FrictionDev& fc, int node) {
  if (node >= fc.split) return FaceRange{0, fc.left_count, 0, fc.left_words};
  return FaceRange{fc.left_count, fc.face_count, fc.left_words, fc.words};
}

__device__ __forceinline__ int word_first_face(const FrictionDev& fc, int w) {
  return w < fc.left_words ? w * 64 : fc.left_count + (w - fc.left_words) * 64;
}

// Squared distance from p to word w's box: a lower bound of dist2 for every
// face of the word (each triangle lies in the box of its corners). An
// unbounded box gives 0.
__device__ __forceinline__ double box_bound2(const FrictionDev& fc, int w, V3 p) {
#pragma clang fp contract(off)
  const double* box = fc.face_box + 6 * static_cast<size_t>(w);
  const double dx = fmax(fmax(box[0] - p.x, p.x - box[3]), 0.0);
  const double dy = fmax(fmax(box[1] - p.y, p.y - box[4]), 0.0);
  const double dz = fmax(fmax(box[2] - p.z, p.z - box[5]), 0.0);
  return dx * dx + dy * dy + dz * dz;
}

// A word can be left out when its box lies farther from p than a face
// already found, with a margin (1e-6 relative, 1e-9 of the coordinates'
// scale) far above the rounding of the closest-point arithmetic, so no face
// the search leaves out could tie or beat the one it keeps.
__device__ __forceinline__ bool beyond(double bound2, double best2, V3 p) {
  const double scale = 1.0 + fmax(fabs(p.x), fmax(fabs(p.y), fabs(p.z)));
  return sqrt(bound2) > sqrt(best2) * (1.0 + 1.0e-6) + 1.0e-9 * scale;
}

// Builds face_box from x, one wavefront per word. A non-finite corner leaves
// the word's box unbounded, so no search leaves it out. Skipped when *active
// is zero (active may be null: always run).
__global__ __launch_bounds__(64) void k_fc_face_boxes(FrictionDev fc, const double* x, const int* active) {
  if (active && !*active) return;
  const int w = blockIdx.x, lane = threadIdx.x;
  const int f = word_first_face(fc, w) + lane;
  const int last = w < fc.left_words ? fc.left_count : fc.face_count;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  bool finite = true;
  if (f < last) {
    for (int k = 0; k < 3; ++k) {
      const V3 v = ld3(x, fc.faces[3 * f + k]);
      for (int d = 0; d < 3; ++d) {
        const double c = vget(v, d);
        finite = finite && isfinite(c);
        lo[d] = fmin(lo[d], c);
        hi[d] = fmax(hi[d], c);
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], __shfl_xor(lo[d], off, 64));
      hi[d] = fmax(hi[d], __shfl_xor(hi[d], off, 64));
    }
  const bool unbounded = __any(!finite);
  if (lane == 0) {
    double* box = fc.face_box + 6 * static_cast<size_t>(w);
    for (int d = 0; d < 3; ++d) {
      box[d] = unbounded ? -INFINITY : lo[d];
      box[3 + d] = unbounded ? INFINITY : hi[d];
    }
  }
}

// (mesh_closest, contact.py:98-137): closest point, unit face normal, dist2
// of the first face of minimal distance. The word nearest by box is searched
// first; the others only where their box can hold a face as near.
__device__ void mesh_closest(const FrictionDev& fc, V3 p, int node, const double* x, V3& closest, V3& normal,
                             double& best) {
#pragma clang fp contract(off)
  const FaceRange r = opposing_faces(fc, node);
  best = INFINITY;
  closest = V3{0, 0, 0};
  normal = V3{0, 0, 1};
  int best_face = r.last;
  int nearest = r.w1;
  double nearest2 = INFINITY;
  for (int w = r.w0; w < r.w1; ++w) {
    const double b2 = box_bound2(fc, w, p);
    if (nearest == r.w1 || b2 < nearest2) {
      nearest = w;
      nearest2 = b2;
    }
  }
  for (int k = -1; k < r.w1 - r.w0; ++k) {
    const int w = k < 0 ? nearest : r.w0 + k;
    if (w == r.w1 || (k >= 0 && (w == nearest || beyond(box_bound2(fc, w, p), best, p)))) continue;
    const int f0 = word_first_face(fc, w), f1 = min(f0 + 64, r.last);
    for (int f = f0; f < f1; ++f) {
      const int* face = fc.faces + 3 * f;
      const V3 a = ld3(x, face[0]), b = ld3(x, face[1]), c = ld3(x, face[2]);
      V3 point, wts;
      double dist2;
      closest_point_triangle(p, a, b, c, point, wts, dist2);
      if (dist2 < best || (dist2 == best && f < best_face)) {
        best = dist2;
        best_face = f;
        closest = point;
        const V3 n = cross3(b - a, c - a);
        const double length = sqrt(dot(n, n));
        normal = length > kEps ? V3{n.x / length, n.y / length, n.z / length} : V3{0, 0, 1};
      }
    }
  }
}

__device__ __forceinline__ V3 mesh_direction(V3 p, V3 closest, V3 face_normal) {
#pragma clang fp contract(off)
  const V3 offset = p - closest;
  const double length = sqrt(dot(offset, offset));
  if (length > 1.0e-12) return V3{offset.x / length, offset.y / length, offset.z / length};
  return face_normal;
}

__global__ __launch_bounds__(256) void k_fc_units(FrictionDev fc, int j0, double* rhs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= fc.n) return;
  for (int l = 0; l < 3; ++l) {
    const int j = j0 + l;
    rhs[3 * i + l] = (j < fc.K && fc.nodes[j] == i) ? 1.0 : 0.0;
  }
}

__global__ __launch_bounds__(256) void k_fc_store(FrictionDev fc, int j0, const double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= fc.n) return;
  for (int l = 0; l < 3 && j0 + l < fc.K; ++l) fc.panel[static_cast<size_t>(j0 + l) * fc.n + i] = out[3 * i + l];
}

__global__ __launch_bounds__(256) void k_fc_block(FrictionDev fc) {
  const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<long long>(fc.K) * fc.K) return;
  const int j = static_cast<int>(t / fc.K), i = static_cast<int>(t % fc.K);
  fc.blockT[t] = fc.panel[static_cast<size_t>(j) * fc.n + fc.nodes[i]];
  if (i == j) fc.mult[i] = 0.0;  // a new factor restarts the multipliers
}

// begin_contact_step / begin_mesh_step (contact.py:171-244).
__global__ __launch_bounds__(256) void k_fc_begin(FrictionDev fc, const double* position) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= fc.K) return;
  const int node = fc.nodes[c];
  const V3 p = ld3(position, node);
  V3 n, t1;
  if (fc.kind == 2) {
    V3 closest, face_normal;
    double dist2;
    mesh_closest(fc, p, node, position, closest, face_normal, dist2);
    const bool active = sqrt(dist2) - fc.radius <= 5.0 * fc.radius;
    n = mesh_direction(p, closest, face_normal);
    int axis = 0;
    for (int d = 1; d < 3; ++d)
      if (fabs(vget(n, d)) < fabs(vget(n, axis))) axis = d;
    V3 unit{0, 0, 0};
    vset(unit, axis, 1.0);
    t1 = unit - dot(n, unit) * n;
    const double l1 = sqrt(dot(t1, t1));
    t1 = V3{t1.x / l1, t1.y / l1, t1.z / l1};
    fc.candidate[c] = active ? 1 : 0;
    if (!active) fc.mult[c] = 0.0;
  } else {
    n = boundary_normal(fc, p);
    V3 best{0, 0, 0};
    for (int i = 0; i < 3; ++i) {
      V3 axis{0, 0, 0};
      vset(axis, i, 1.0);
      const V3 cand = cross3(n, axis);
      if (dot(cand, cand) > dot(best, best)) best = cand;
    }
    const double lb = sqrt(dot(best, best));
    t1 = V3{best.x / lb, best.y / lb, best.z / lb};
    fc.candidate[c] = 1;
  }
  if (fc.defect) st3(fc.fprev, c, V3{0, 0, 0});
  const V3 u = cross3(n, t1);
  const double l2 = sqrt(dot(u, u));
  st3(fc.normal, c, n);
  st3(fc.t1, c, t1);
  st3(fc.t2, c, V3{u.x / l2, u.y / l2, u.z / l2});
  fc.fl[2 * c] = fc.fl[2 * c + 1] = 0.0;
}

// contact_delassus (contact.py:247-263): W_ij = (n_lo . n_hi) B(hi, lo).
__global__ __launch_bounds__(256) void k_fc_delassus(FrictionDev fc, double dt) {
#pragma clang fp contract(off)
  const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<long long>(fc.K) * fc.K) return;
  const int i = static_cast<int>(t / fc.K), j = static_cast<int>(t % fc.K);
  const int lo = i < j ? i : j, hi = i < j ? j : i;
  const double w = dot(ld3(fc.normal, lo), ld3(fc.normal, hi)) * fc.blockT[static_cast<size_t>(lo) * fc.K + hi];
  fc.delassus[t] = w;
  if (i == j) fc.reg[i] = fmax(dt * dt * w, 1.0e-10);
}

__global__ __launch_bounds__(256) void k_fc_surface(FrictionDev fc, const double* x, const int* active) {
  if (!*active) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 3 * fc.n) fc.surface[t] = x[t];
}

__device__ __forceinline__ void complementarity(double delta, double lam, double r, double& omega, double& e,
                                                double& phi) {
#pragma clang fp contract(off)
  double s = sqrt(delta * delta + r * r * lam * lam);
  if (s < 1.0e-14) s = 1.0e-14;
  omega = 1.0 - delta / s;
  e = (1.0 - r * lam / s) * r;
  phi = delta + r * lam - s;
}

// contact_scalars / mesh_contact_scalars(_tiled) (contact.py:266-370): one
// wavefront per candidate; for the mesh boundary its lanes split the faces
// and the first face of minimal distance wins.
__global__ __launch_bounds__(64) void k_fc_scalars(FrictionDev fc, const double* x, const int* active) {
#pragma clang fp contract(off)
  if (!*active) return;
  const int c = blockIdx.x, lane = threadIdx.x;
  const int node = fc.nodes[c];
  const V3 p = ld3(x, node);
  double delta;
  if (fc.kind == 2) {
    if (!fc.candidate[c]) {
      if (lane == 0) {
        fc.gap[c] = 0.0;
        fc.omega[c] = 0.0;
        fc.diag[c] = 1.0;
        fc.resid[c] = 0.0;
      }
      return;
    }
    // The word nearest by box first, then every other word whose box can
    // hold a face as near as the nearest word's best; the first face of
    // minimal distance wins over all faces searched.
    const FaceRange r = opposing_faces(fc, node);
    const int last = r.last;
    double nearest2 = INFINITY;
    int nearest = r.w1;
    for (int w = r.w0 + lane; w < r.w1; w += 64) {
      const double b2 = box_bound2(fc, w, p);
      if (nearest == r.w1 || b2 < nearest2) {
        nearest = w;
        nearest2 = b2;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ob = __shfl_xor(nearest2, off, 64);
      const int ow = __shfl_xor(nearest, off, 64);
      if (ow < r.w1 && (nearest == r.w1 || ob < nearest2 || (ob == nearest2 && ow < nearest))) {
        nearest2 = ob;
        nearest = ow;
      }
    }
    double best = INFINITY;
    int best_face = last;
    auto search = [&](int w) {
      const int f = word_first_face(fc, w) + lane;
      if (f >= last) return;
      const int* face = fc.faces + 3 * f;
      V3 point, wts;
      double dist2;
      closest_point_triangle(p, ld3(x, face[0]), ld3(x, face[1]), ld3(x, face[2]), point, wts, dist2);
      if (dist2 < best || (dist2 == best && f < best_face)) {
        best = dist2;
        best_face = f;
      }
    };
    if (nearest < r.w1) search(nearest);
    double bound = best;
    for (int off = 32; off > 0; off >>= 1) bound = fmin(bound, __shfl_xor(bound, off, 64));
    for (int base = r.w0; base < r.w1; base += 64) {
      const int w = base + lane;
      const bool open = w < r.w1 && w != nearest && !beyond(box_bound2(fc, w, p), bound, p);
      unsigned long long mask = __ballot(open);
      while (mask) {
        search(base + __builtin_ctzll(mask));
        mask &= mask - 1;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ob = __shfl_xor(best, off, 64);
      const int of = __shfl_xor(best_face, off, 64);
      if (ob < best || (ob == best && of < best_face)) {
        best = ob;
        best_face = of;
      }
    }
    if (lane != 0) return;
    V3 closest{0, 0, 0};
    if (best_face < last) {
      const int* face = fc.faces + 3 * best_face;
      V3 w;
      double dist2;
      closest_point_triangle(p, ld3(x, face[0]), ld3(x, face[1]), ld3(x, face[2]), closest, w, dist2);
    }
    delta = dot(p - closest, ld3(fc.normal, c)) - fc.radius;
  } else {
    if (lane != 0) return;
    delta = boundary_gap(fc, p);
  }
  fc.gap[c] = delta;
  complementarity(delta, fc.mult[c], fc.reg[c], fc.omega[c], fc.diag[c], fc.resid[c]);
}

__device__ void fc_forces(const FrictionDev& fc, bool tangential) {
#pragma clang fp contract(off)
  for (int c = threadIdx.x; c < fc.K; c += blockDim.x) {
    V3 v = (fc.omega[c] * fc.mult[c]) * ld3(fc.normal, c);
    if (tangential) v = v + (fc.fl[2 * c] * ld3(fc.t1, c) + fc.fl[2 * c + 1] * ld3(fc.t2, c));
    st3(fc.force, c, v);
  }
}

// contact_response: B f over the coupled candidates, in their order.
__device__ void fc_response(const FrictionDev& fc, int A) {
#pragma clang fp contract(off)
  for (int c = threadIdx.x; c < fc.K; c += blockDim.x) {
    V3 v{0, 0, 0};
    for (int k = 0; k < A; ++k) {
      const int j = fc.order[k];
      v = v + fc.blockT[static_cast<size_t>(j) * fc.K + c] * ld3(fc.force, j);
    }
    st3(fc.response, c, v);
  }
}

// Cholesky solve of the coupled block (up to 32 rows), one lane.
__device__ void cholesky_solve(double* m, int n, double* b) {
#pragma clang fp contract(off)
  for (int j = 0; j < n; ++j) {
    double d = m[j * n + j];
    for (int k = 0; k < j; ++k) d -= m[j * n + k] * m[j * n + k];
    d = sqrt(d);
    m[j * n + j] = d;
    for (int i = j + 1; i < n; ++i) {
      double v = m[i * n + j];
      for (int k = 0; k < j; ++k) v -= m[i * n + k] * m[j * n + k];
      m[i * n + j] = v / d;
    }
  }
  for (int i = 0; i < n; ++i) {
    double v = b[i];
    for (int k = 0; k < i; ++k) v -= m[i * n + k] * b[k];
    b[i] = v / m[i * n + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    double v = b[i];
    for (int k = i + 1; k < n; ++k) v -= m[k * n + i] * b[k];
    b[i] = v / m[i * n + i];
  }
}

// Eigen's LDLT with diagonal pivoting on the lower triangle (dense.py:40-106),
// row updates spread over the workgroup.
__device__ void ldlt_solve_block(double* m, int n, double* b, int* perm, double* work) {
#pragma clang fp contract(off)
  __shared__ int s_big;
  for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = i;
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    if (threadIdx.x == 0) {
      int big = k;
      for (int i = k + 1; i < n; ++i)
        if (fabs(m[i * n + i]) > fabs(m[big * n + big])) big = i;
      s_big = big;
    }
    __syncthreads();
    const int big = s_big;
    if (big != k) {
      for (int j = threadIdx.x; j < k; j += blockDim.x) {
        const double t = m[k * n + j];
        m[k * n + j] = m[big * n + j];
        m[big * n + j] = t;
      }
      for (int i = big + 1 + threadIdx.x; i < n; i += blockDim.x) {
        const double t = m[i * n + k];
        m[i * n + k] = m[i * n + big];
        m[i * n + big] = t;
      }
      for (int i = k + 1 + threadIdx.x; i < big; i += blockDim.x) {
        const double t = m[i * n + k];
        m[i * n + k] = m[big * n + i];
        m[big * n + i] = t;
      }
      if (threadIdx.x == 0) {
        const double t = m[k * n + k];
        m[k * n + k] = m[big * n + big];
        m[big * n + big] = t;
        const int pi = perm[k];
        perm[k] = perm[big];
        perm[big] = pi;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int j = 0; j < k; ++j) m[k * n + k] = m[k * n + k] - m[k * n + j] * m[k * n + j] * m[j * n + j];
    __syncthreads();
    for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) {
      for (int j = 0; j < k; ++j) m[i * n + k] = m[i * n + k] - m[i * n + j] * m[j * n + j] * m[k * n + j];
      if (m[k * n + k] != 0.0) m[i * n + k] = m[i * n + k] / m[k * n + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int i = 0; i < n; ++i) work[i] = b[perm[i]];
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < i; ++j) work[i] = work[i] - m[i * n + j] * work[j];
    for (int i = 0; i < n; ++i) work[i] = fabs(m[i * n + i]) > 2.2250738585072014e-308 ? work[i] / m[i * n + i] : 0.0;
    for (int i = n - 1; i >= 0; --i)
      for (int j = i + 1; j < n; ++j) work[i] = work[i] - m[j * n + i] * work[j];
    for (int i = 0; i < n; ++i) b[perm[i]] = work[i];
  }
  __syncthreads();
}

// One Newton step on the multipliers and the friction update
// (FrictionalContact._newton / iterate, contact.py:865-987), one workgroup.
__global__ __launch_bounds__(256) void k_fc_newton(FrictionDev fc, const double* x, const double* position,
                                                   const int* active) {
#pragma clang fp contract(off)
  if (!*active) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ int s_wave[4];
  __shared__ int s_base;
  // The previous iteration's coupled set, for the FP32-factor mode's solution.
  if (fc.defect) {
    const int prev = *fc.nactive;
    for (int k = tid; k < prev; k += 256) fc.prev_order[k] = fc.order[k];
    if (tid == 0) *fc.nprev = prev;
  }
  // contact_active(_scan): coupled candidates (omega != 0) in candidate order.
  if (tid == 0) s_base = 0;
  __syncthreads();
  for (int start = 0; start < fc.K; start += 256) {
    const int c = start + tid;
    const bool flag = c < fc.K && fc.omega[c] != 0.0;
    const uint64_t mask = __ballot(flag);
    if (lane == 0) s_wave[wave] = __popcll(mask);
    __syncthreads();
    int offset = s_base;
    for (int w = 0; w < wave; ++w) offset += s_wave[w];
    const int k = offset + __popcll(mask & ((1ull << lane) - 1ull));
    if (c < fc.K) {
      fc.slot[c] = flag ? k : -1;
      if (flag) fc.order[k] = c;
    }
    __syncthreads();
    if (tid == 0) s_base += s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    __syncthreads();
  }
  const int A = s_base;
  if (tid == 0) *fc.nactive = A;
  fc_forces(fc, true);
  __syncthreads();
  fc_response(fc, A);
  __syncthreads();
  // contact_newton_rhs: -phi + omega n.x - omega n.(A^{-1} g).
  for (int c = tid; c < fc.K; c += 256) {
    const int node = fc.nodes[c];
    const V3 n = ld3(fc.normal, c);
    const double w = fc.omega[c];
    fc.rhs[c] = -fc.resid[c] + w * dot(n, ld3(x, node)) - w * dot(n, ld3(fc.freec, c) + ld3(fc.response, c));
  }
  __syncthreads();
  // contact_system: Omega W Omega + E over the coupled block.
  for (long long t = tid; t < static_cast<long long>(A) * A; t += 256) {
    const int i = static_cast<int>(t / A), j = static_cast<int>(t % A);
    const int a = fc.order[i], b = fc.order[j];
    double v = fc.omega[a] * fc.delassus[static_cast<size_t>(a) * fc.K + b] * fc.omega[b];
    if (i == j) v += fc.diag[a];
    fc.matrix[t] = v;
    if (j == 0) fc.compact[i] = fc.rhs[a];
  }
  __syncthreads();
  if (A > 32) {
    ldlt_solve_block(fc.matrix, A, fc.compact, fc.perm, fc.work);
  } else if (A > 0) {
    if (tid == 0) cholesky_solve(fc.matrix, A, fc.compact);
    __syncthreads();
  }
  // contact_update
  for (int c = tid; c < fc.K; c += 256) {
    const int k = fc.slot[c];
    const double step = k >= 0 ? fc.compact[k] : fc.rhs[c] / fc.diag[c];
    fc.mult[c] = fmin(fmax(fc.mult[c] + step, 0.0), 1.0e8);
  }
  __syncthreads();
  if (fc.friction > 0.0) {
    fc_forces(fc, false);
    __syncthreads();
    fc_response(fc, A);
    __syncthreads();
    // contact_friction (contact.py:546-584)
    for (int c = tid; c < fc.K; c += 256) {
      if (fc.mult[c] <= 0.0) {
        fc.fl[2 * c] = fc.fl[2 * c + 1] = 0.0;
        continue;
      }
      const int node = fc.nodes[c];
      const V3 step = ld3(fc.freec, c) + ld3(fc.response, c) - ld3(position, node);
      const V3 n = ld3(fc.normal, c), t1 = ld3(fc.t1, c), t2 = ld3(fc.t2, c);
      const V3 tangential = step - dot(step, n) * n;
      const double u1 = dot(tangential, t1), u2 = dot(tangential, t2);
      const double bd = fc.blockT[static_cast<size_t>(c) * fc.K + c];
      const double w11 = dot(t1, t1) * bd, w12 = dot(t1, t2) * bd, w22 = dot(t2, t2) * bd;
      const double det = w11 * w22 - w12 * w12;
      double s1 = 0.0, s2 = 0.0;
      if (fabs(det) > 1.0e-20) {
        s1 = -(w22 * u1 - w12 * u2) / det;
        s2 = -(-w12 * u1 + w11 * u2) / det;
      }
      const double radius = fc.friction * fc.mult[c];
      if (sqrt(s1 * s1 + s2 * s2) <= radius + 1.0e-14) {
        fc.fl[2 * c] = s1;
        fc.fl[2 * c + 1] = s2;
      } else {
        const double mag = sqrt(u1 * u1 + u2 * u2);
        fc.fl[2 * c] = mag > 1.0e-14 ? -radius * u1 / mag : 0.0;
        fc.fl[2 * c + 1] = mag > 1.0e-14 ? -radius * u2 / mag : 0.0;
      }
    }
    __syncthreads();
  }
  fc_forces(fc, true);
}

// The free solution A^{-1} b at the candidates: the solve's value there, less
// (FP32-factor mode) the response B f_prev of the forces already in b.
__global__ __launch_bounds__(256) void k_fc_free(FrictionDev fc, const double* solved, const int* active) {
#pragma clang fp contract(off)
  if (!*active) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= fc.K) return;
  V3 v = ld3(solved, fc.nodes[c]);
  if (fc.defect) {
    // f_prev is zero outside the previous iteration's coupled set (still in
    // `order`: the Newton kernel selects this iteration's after this one), so
    // the sum over that set, ascending, is the sum over every candidate.
    const int prev = *fc.nactive;
    for (int k = 0; k < prev; ++k) {
      const int j = fc.order[k];
      v = v - fc.blockT[static_cast<size_t>(j) * fc.K + c] * ld3(fc.fprev, j);
    }
  }
  st3(fc.freec, c, v);
}

// contact_solution: x = A^{-1} b + sum over coupled candidates of A^{-1} e_j f_j.
// In the FP32-factor mode b already holds the previous coupled forces, so
// only their change is added: x = solve + sum_j A^{-1} e_j (f_j - f_prev_j).
__global__ __launch_bounds__(256) void k_fc_solution(FrictionDev fc, double* xr, const int* active) {
#pragma clang fp contract(off)
  if (!*active) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= fc.n) return;
  V3 v = ld3(xr, i);
  if (fc.defect) {
    // f - f_prev can be nonzero only for candidates coupled now or in the
    // previous iteration: the two ascending sets merged, so the candidates
    // are visited in the order of a loop over all of them.
    const int na = *fc.nactive, np = *fc.nprev;
    int a = 0, b = 0;
    while (a < na || b < np) {
      const int ja = a < na ? fc.order[a] : INT_MAX, jb = b < np ? fc.prev_order[b] : INT_MAX;
      const int j = min(ja, jb);
      a += ja == j;
      b += jb == j;
      const V3 f = fc.slot[j] >= 0 ? ld3(fc.force, j) : V3{0, 0, 0};
      const V3 delta = f - ld3(fc.fprev, j);
      if (delta.x != 0.0 || delta.y != 0.0 || delta.z != 0.0)
        v = v + fc.panel[static_cast<size_t>(j) * fc.n + i] * delta;
    }
  } else {
    const int A = *fc.nactive;
    for (int k = 0; k < A; ++k) {
      const int j = fc.order[k];
      v = v + fc.panel[static_cast<size_t>(j) * fc.n + i] * ld3(fc.force, j);
    }
  }
  st3(xr, i, v);
}

// FP32-factor mode: the coupled forces of this iteration, for the next
// iteration's right-hand side.
__global__ __launch_bounds__(256) void k_fc_keep(FrictionDev fc, const int* active) {
  if (!*active) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= fc.K) return;
  st3(fc.fprev, c, fc.slot[c] >= 0 ? ld3(fc.force, c) : V3{0, 0, 0});
}

// apply_restitution / apply_mesh_restitution (contact.py:587-660).
__global__ __launch_bounds__(256) void k_fc_restitution(FrictionDev fc, const double* position, const float* velocity,
                                                        double dt, double* x) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= fc.K || fc.mult[c] <= 0.0) return;
  const int node = fc.nodes[c];
  const V3 point = ld3(x, node);
  V3 n;
  double distance;
  if (fc.kind == 2) {
    V3 closest, face_normal;
    double dist2;
    mesh_closest(fc, point, node, fc.surface, closest, face_normal, dist2);
    const V3 reference = ld3(fc.normal, c);
    distance = dot(point - closest, reference) - fc.radius;
    n = mesh_direction(point, closest, face_normal);
    if (dot(n, reference) < 0.0) n = -1.0 * n;
  } else {
    distance = boundary_gap(fc, point);
    n = boundary_normal(fc, point);
  }
  const double norm = sqrt(dot(n, n));
  if (norm <= 1.0e-12) return;
  n = V3{n.x / norm, n.y / norm, n.z / norm};
  if (distance > (fc.kind == 2 ? 0.1 * fc.radius : 1.0e-8)) return;
  const V3 v{static_cast<double>(velocity[3 * node]), static_cast<double>(velocity[3 * node + 1]),
             static_cast<double>(velocity[3 * node + 2])};
  const double incoming = dot(v, n);
  if (incoming < 0.0) {
    const V3 d = point - ld3(position, node);
    const double outgoing = dot(V3{d.x / dt, d.y / dt, d.z / dt}, n);
    const double change = -fc.restitution * incoming - outgoing;
    if (change > 0.0) st3(x, node, point + (dt * change) * n);
  }
}


// ---------------------------------------------------------------------------
// Woodbury rows and projection (adjoint.py:324-471, 1272-1354).

// _backward_tangents (adjoint.py:324-345).
__device__ void backward_tangents(V3 n, V3& t1, V3& t2) {
#pragma clang fp contract(off)
  int axis = 0;
  for (int d = 1; d < 3; ++d)
    if (fabs(vget(n, d)) < fabs(vget(n, axis))) axis = d;
  V3 unit{0, 0, 0};
  vset(unit, axis, 1.0);
  t1 = unit - dot(n, unit) * n;
  double length = sqrt(dot(t1, t1));
  t1 = length > 1.0e-12 ? V3{t1.x / length, t1.y / length, t1.z / length} : V3{1, 0, 0};
  t2 = cross3(n, t1);
  length = sqrt(dot(t2, t2));
  t2 = length > 1.0e-12 ? V3{t2.x / length, t2.y / length, t2.z / length} : V3{0, 1, 0};
}

// _basis, _project_rows, _select_rows, _compact_rows in one workgroup.
__global__ __launch_bounds__(256) void k_wb_rows(WoodburyDev w, const double* mu, const double* route) {
#pragma clang fp contract(off)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool diverged = route[1] != 0.0;
  for (int c = tid; c < w.K; c += 256) {
    const int node = w.nodes[c];
    const V3 n = ld3(w.normal, c);
    const bool active = !diverged && w.mult[c] > 0.0 && w.omega[c] >= 0.5;
    w.row_node[c] = node;
    st3(w.row_vec, c, w.omega[c] * n);
    w.row_valid[c] = active ? 1 : 0;
    if (w.tangents) {
      V3 t1, t2;
      backward_tangents(n, t1, t2);
      const int valid = (active && w.mult[c] > 1.0e-8) ? 1 : 0;
      for (int r = 0; r < 2; ++r) {
        const int row = w.K + 2 * c + r;
        w.row_node[row] = node;
        st3(w.row_vec, row, r == 0 ? t1 : t2);
        w.row_valid[row] = valid;
      }
    }
  }
  __syncthreads();
  __shared__ double s_max[4];
  double largest = 0.0;
  for (int r = tid; r < w.R; r += 256) {
    const double v = w.row_valid[r] ? dot(ld3(w.row_vec, r), ld3(mu, w.row_node[r])) : 0.0;
    w.value[r] = v;
    if (w.row_valid[r]) largest = fmax(largest, fabs(v));
  }
  for (int off = 32; off > 0; off >>= 1) largest = fmax(largest, __shfl_xor(largest, off, 64));
  if (lane == 0) s_max[wave] = largest;
  __syncthreads();
  const double gate = fmax(1.0e-12, 1.0e-3 * fmax(fmax(s_max[0], s_max[1]), fmax(s_max[2], s_max[3])));
  __shared__ int s_wave[4];
  __shared__ int s_base;
  if (tid == 0) s_base = 0;
  __syncthreads();
  for (int start = 0; start < w.R; start += 256) {
    const int r = start + tid;
    const bool flag = r < w.R && w.row_valid[r] && fabs(w.value[r]) >= gate;
    const uint64_t mask = __ballot(flag);
    if (lane == 0) s_wave[wave] = __popcll(mask);
    __syncthreads();
    int offset = s_base;
    for (int q = 0; q < wave; ++q) offset += s_wave[q];
    if (flag) w.order[offset + __popcll(mask & ((1ull << lane) - 1ull))] = r;
    __syncthreads();
    if (tid == 0) s_base += s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    __syncthreads();
  }
  if (tid == 0) *w.count = s_base;
}

__global__ __launch_bounds__(256) void k_wb_batch_seed(WoodburyDev w, int first, int bodies, int n, double* rhs,
                                                       double* sol, double* xr, double* fscale, int* aiter,
                                                       int* aactive, int* aa) {
  const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<long long>(bodies) * n) return;
  const int k = static_cast<int>(t / n), i = static_cast<int>(t % n);
  const bool on = first + k < *w.count;
  V3 v{0, 0, 0};
  if (on) {
    const int r = w.order[first + k];
    if (i == w.row_node[r]) v = ld3(w.row_vec, r);
  }
  const int j = static_cast<int>(t);
  st3(rhs, j, v);
  st3(sol, j, V3{0, 0, 0});
  st3(xr, j, V3{0, 0, 0});
  if (i == 0) {
    aiter[k] = 0;
    aactive[k] = on ? 1 : 0;
    aa[2 * k] = aa[2 * k + 1] = 0;
    fscale[k] = 1.0;
  }
}

__global__ void k_status_merge(int* from, int* to) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const int bits = *from;
    if (bits) atomicOr(to, bits);
    *from = 0;
  }
}

// _gram (adjoint.py:404-425) and the Eigen LDLT of G over the selected rows.
__global__ __launch_bounds__(256) void k_wb_solve(WoodburyDev w) {
#pragma clang fp contract(off)
  const int count = *w.count;
  if (count == 0) return;
  const size_t stride = 3 * static_cast<size_t>(w.n);
  for (long long t = threadIdx.x; t < static_cast<long long>(count) * count; t += 256) {
    const int i = static_cast<int>(t / count), j = static_cast<int>(t % count);
    const int hi = i > j ? i : j, lo = i > j ? j : i;
    const int rh = w.order[hi];
    w.gram[t] = dot(ld3(w.row_vec, rh), ld3(w.columns + lo * stride, w.row_node[rh]));
    if (j == 0) w.coef[i] = w.value[w.order[i]];
  }
  __syncthreads();
  ldlt_solve_block(w.gram, count, w.coef, w.perm, w.work);
}

// _project_adjoint (adjoint.py:457-471): z = mu - sum_k alpha_k column_k in row order.
__global__ __launch_bounds__(256) void k_wb_project(WoodburyDev w, const double* mu, double* z) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.n) return;
  const int count = *w.count;
  const size_t stride = 3 * static_cast<size_t>(w.n);
  V3 v = ld3(mu, i);
  for (int k = 0; k < count; ++k) v = v - w.coef[k] * ld3(w.columns + k * stride, i);
  st3(z, i, v);
}

// _safe_clip of one vector (adjoint.py:179-221): non-finite entries of free
// rows scrubbed to zero, Dirichlet rows zeroed, the norm capped at 1e6;
// either safeguard sets status 16.
__global__ __launch_bounds__(256) void k_wb_clip(int n, const uint8_t* free, double* z, int* status) {
#pragma clang fp contract(off)
  __shared__ double s_sum[4];
  double total = 0.0;
  bool scrubbed = false;
  for (int i = threadIdx.x; i < n; i += 256) {
    V3 v = free[i] ? ld3(z, i) : V3{0, 0, 0};
    for (int d = 0; d < 3; ++d)
      if (!isfinite(vget(v, d))) {
        vset(v, d, 0.0);
        scrubbed = true;
      }
    st3(z, i, v);
    total += dot(v, v);
  }
  if (scrubbed) atomicOr(status, 16);
  for (int off = 32; off > 0; off >>= 1) total += __shfl_xor(total, off, 64);
  if ((threadIdx.x & 63) == 0) s_sum[threadIdx.x >> 6] = total;
  __syncthreads();
  const double length = sqrt(((s_sum[0] + s_sum[1]) + s_sum[2]) + s_sum[3]);
  const double limit = 1.0e6;
  if (isfinite(length) && length > limit) {
    const double scale = limit / length;
    for (int t = threadIdx.x; t < 3 * n; t += 256) z[t] = scale * z[t];
    if (threadIdx.x == 0) atomicOr(status, 16);
  }
}

}  // namespace

void launch_mesh_contact(const MeshContactDev& mc, const float* pos32, const float* vel32, const double* pos64,
                         const double* vel64, double* target, const int* vbody, const double* mass, double dt,
                         hipStream_t stream) {
  const Vec3Src pos{pos32, pos64}, vel{vel32, vel64};
  // Widened by 1 + 1e-6: the margin covers the rounding of the closest-point
  // distance, so no pair the exact test hits lies outside a face's box.
  const double reach = mc.radius * (1.0 + 1.0e-6);
  hipLaunchKernelGGL(k_mc_zero, dim3(1), dim3(1), 0, stream, mc.pairs);
  for (int d = 0; d < 2; ++d) {
    const MeshContactSide& side = d == 0 ? mc.ab : mc.ba;
    if (side.nv == 0 || side.nf == 0) continue;
    hipLaunchKernelGGL(k_mc_grid, dim3(1), dim3(256), 0, stream, side, pos, reach);
    hipLaunchKernelGGL(k_mc_clear, dim3(blocks_for(kGridCells + 1)), dim3(256), 0, stream, side);
    const long long items = static_cast<long long>(side.nf) + side.nv;
    hipLaunchKernelGGL(k_mc_bin, dim3(blocks_for(items)), dim3(256), 0, stream, side, pos, reach, 0);
    hipLaunchKernelGGL(k_mc_scan, dim3(1), dim3(256), 0, stream, side);
    hipLaunchKernelGGL(k_mc_bin, dim3(blocks_for(items)), dim3(256), 0, stream, side, pos, reach, 1);
    hipLaunchKernelGGL(k_mc_rank, dim3(blocks_for(8LL * side.nf + side.nv)), dim3(256), 0, stream, side);
  }
  hipLaunchKernelGGL(k_mc_gather, dim3(blocks_for(mc.n)), dim3(256), 0, stream, mc, pos, vel, target, vbody, mass,
                     dt, reach);
}

}  // namespace dphd

namespace dphd {

void launch_contact_units(const FrictionDev& fc, int j0, double* rhs, hipStream_t stream) {
  hipLaunchKernelGGL(k_fc_units, dim3(blocks_for(fc.n)), dim3(256), 0, stream, fc, j0, rhs);
}

void launch_contact_store(const FrictionDev& fc, int j0, const double* out, hipStream_t stream) {
  hipLaunchKernelGGL(k_fc_store, dim3(blocks_for(fc.n)), dim3(256), 0, stream, fc, j0, out);
}

void launch_contact_block(const FrictionDev& fc, hipStream_t stream) {
  hipLaunchKernelGGL(k_fc_block, dim3(blocks_for(static_cast<long long>(fc.K) * fc.K)), dim3(256), 0, stream, fc);
}

void launch_contact_begin(const FrictionDev& fc, const double* position, double dt, hipStream_t stream) {
  if (fc.K == 0) return;
  if (fc.kind == 2 && fc.words > 0)
    hipLaunchKernelGGL(k_fc_face_boxes, dim3(fc.words), dim3(64), 0, stream, fc, position, nullptr);
  hipLaunchKernelGGL(k_fc_begin, dim3(blocks_for(fc.K)), dim3(256), 0, stream, fc, position);
  hipLaunchKernelGGL(k_fc_delassus, dim3(blocks_for(static_cast<long long>(fc.K) * fc.K)), dim3(256), 0, stream, fc,
                     dt);
}

void launch_contact_iteration(const FrictionDev& fc, const double* x, double* xr, const double* position,
                              const int* active, hipStream_t stream) {
  if (fc.K == 0) return;
  if (fc.kind == 2 && fc.restitution > 0.0)
    hipLaunchKernelGGL(k_fc_surface, dim3(blocks_for(3LL * fc.n)), dim3(256), 0, stream, fc, x, active);
  if (fc.kind == 2 && fc.words > 0)
    hipLaunchKernelGGL(k_fc_face_boxes, dim3(fc.words), dim3(64), 0, stream, fc, x, active);
  hipLaunchKernelGGL(k_fc_scalars, dim3(fc.K), dim3(64), 0, stream, fc, x, active);
  hipLaunchKernelGGL(k_fc_free, dim3(blocks_for(fc.K)), dim3(256), 0, stream, fc, xr, active);
  hipLaunchKernelGGL(k_fc_newton, dim3(1), dim3(256), 0, stream, fc, x, position, active);
  hipLaunchKernelGGL(k_fc_solution, dim3(blocks_for(fc.n)), dim3(256), 0, stream, fc, xr, active);
  if (fc.defect) hipLaunchKernelGGL(k_fc_keep, dim3(blocks_for(fc.K)), dim3(256), 0, stream, fc, active);
}

void launch_contact_finish(const FrictionDev& fc, const double* position, const float* velocity, double dt, double* x,
                           hipStream_t stream) {
  if (fc.K == 0 || fc.restitution <= 0.0) return;
  if (fc.kind == 2 && fc.words > 0)
    hipLaunchKernelGGL(k_fc_face_boxes, dim3(fc.words), dim3(64), 0, stream, fc, fc.surface, nullptr);
  hipLaunchKernelGGL(k_fc_restitution, dim3(blocks_for(fc.K)), dim3(256), 0, stream, fc, position, velocity, dt, x);
}

}  // namespace dphd

namespace dphd {

void launch_woodbury_rows(const WoodburyDev& w, const double* mu, const double* route, hipStream_t stream) {
  hipLaunchKernelGGL(k_wb_rows, dim3(1), dim3(256), 0, stream, w, mu, route);
}

void launch_woodbury_batch_seed(const WoodburyDev& w, int first, int bodies, int n, double* rhs, double* sol,
                                double* xr, double* fscale, int* aiter, int* aactive, int* aa, hipStream_t stream) {
  const long long total = static_cast<long long>(bodies) * n;
  hipLaunchKernelGGL(k_wb_batch_seed, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, stream, w,
                     first, bodies, n, rhs, sol, xr, fscale, aiter, aactive, aa);
}

void launch_status_merge(int* from, int* to, hipStream_t stream) {
  hipLaunchKernelGGL(k_status_merge, dim3(1), dim3(64), 0, stream, from, to);
}

void launch_woodbury_project(const WoodburyDev& w, const double* mu, const uint8_t* free, double* z, int* status,
                             hipStream_t stream) {
  hipLaunchKernelGGL(k_wb_solve, dim3(1), dim3(256), 0, stream, w);
  hipLaunchKernelGGL(k_wb_project, dim3(blocks_for(w.n)), dim3(256), 0, stream, w, mu, z);
  hipLaunchKernelGGL(k_wb_clip, dim3(1), dim3(256), 0, stream, w.n, free, z, status);
}

}  // namespace dphd
