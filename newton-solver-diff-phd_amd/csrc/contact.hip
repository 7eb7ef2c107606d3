// gfx950 kernels of the mesh contact state force (reference:
// src/newton_solver_diff_phd/_src/solvers/diff_phd/mesh_contact.py:34-179).
//
//   k_mc_rows       one wavefront per (vertex, 64 faces): the pair test of
//                   _pair_forces (:90-137) per lane, hits as a ballot word
//   k_mc_transpose  the hit words by face
//   k_mc_gather     _gather_forces (:140-170) + add_state_force (:173-177):
//                   a particle's own hit pairs in face order, then the
//                   reactions on the faces it is a corner of in vertex order
//
// Pair forces are recomputed where they are summed from the same inputs with
// floating-point contraction off, so the hit test and the force agree
// bitwise between the kernels and every sum has the reference's order
// (pairs outside the radius contribute exact zeros there and are skipped).
#include "contact.h"
#include "pd_math.h"

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

__global__ void k_mc_zero(int* pairs) { *pairs = 0; }

__global__ __launch_bounds__(256) void k_mc_rows(MeshContactSide side, Vec3Src pos, Vec3Src vel, double radius,
                                                 double kf, double mu, int* pairs) {
  const long long wave = (static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  if (wave >= static_cast<long long>(side.nv) * side.row_words) return;  // uniform per wavefront
  const int s = static_cast<int>(wave / side.row_words), word = static_cast<int>(wave % side.row_words);
  const int f = word * 64 + lane;
  bool hit = false;
  if (f < side.nf) hit = pair_force(pos, vel, side.verts[s], side.faces + 3 * f, radius, side.kn, kf, mu).hit;
  const uint64_t mask = __ballot(hit);
  if (lane == 0) {
    side.rows[static_cast<size_t>(s) * side.row_words + word] = mask;
    if (mask) atomicAdd(pairs, __popcll(mask));
  }
}

__global__ __launch_bounds__(256) void k_mc_transpose(MeshContactSide side) {
  const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<long long>(side.nf) * side.col_words) return;
  const int f = static_cast<int>(t / side.col_words), cw = static_cast<int>(t % side.col_words);
  const int fw = f >> 6, fb = f & 63;
  uint64_t word = 0;
  for (int b = 0; b < 64; ++b) {
    const int s = cw * 64 + b;
    if (s >= side.nv) break;
    word |= ((side.rows[static_cast<size_t>(s) * side.row_words + fw] >> fb) & 1ull) << b;
  }
  side.cols[static_cast<size_t>(f) * side.col_words + cw] = word;
}

// The particle's own pairs (as a vertex of `side`) in face order.
__device__ __forceinline__ V3 own_pairs(const MeshContactSide& side, int s, const Vec3Src& pos, const Vec3Src& vel,
                                        double radius, double kf, double mu, V3 total) {
  for (int word = 0; word < side.row_words; ++word) {
    uint64_t m = side.rows[static_cast<size_t>(s) * side.row_words + word];
    while (m) {
      const int f = word * 64 + __builtin_ctzll(m);
      m &= m - 1;
      total = total + pair_force(pos, vel, side.verts[s], side.faces + 3 * f, radius, side.kn, kf, mu).force;
    }
  }
  return total;
}

// Reactions on the particle as corner `corner` of face f of `side`'s faces,
// over the side's vertices in order.
__device__ __forceinline__ V3 face_reactions(const MeshContactSide& side, int f, int corner, const Vec3Src& pos,
                                             const Vec3Src& vel, double radius, double kf, double mu, V3 total) {
  for (int cw = 0; cw < side.col_words; ++cw) {
    uint64_t m = side.cols[static_cast<size_t>(f) * side.col_words + cw];
    while (m) {
      const int s = cw * 64 + __builtin_ctzll(m);
      m &= m - 1;
      const PairForce h = pair_force(pos, vel, side.verts[s], side.faces + 3 * f, radius, side.kn, kf, mu);
      total = total - vget(h.bary, corner) * h.force;
    }
  }
  return total;
}

__global__ __launch_bounds__(256) void k_mc_gather(MeshContactDev mc, Vec3Src pos, Vec3Src vel, double* target,
                                                   const int* vbody, const double* mass, double dt) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= mc.n) return;
  V3 total{0, 0, 0};
  if (mc.slot_a[i] >= 0) total = own_pairs(mc.ab, mc.slot_a[i], pos, vel, mc.radius, mc.kf, mc.mu, total);
  if (mc.slot_b[i] >= 0) total = own_pairs(mc.ba, mc.slot_b[i], pos, vel, mc.radius, mc.kf, mc.mu, total);
  // As a corner of a face of B pushed by A's vertices, then of A by B's.
  for (int k = mc.inc_b_off[i]; k < mc.inc_b_off[i + 1]; ++k)
    total = face_reactions(mc.ab, mc.inc_b[2 * k], mc.inc_b[2 * k + 1], pos, vel, mc.radius, mc.kf, mc.mu, total);
  for (int k = mc.inc_a_off[i]; k < mc.inc_a_off[i + 1]; ++k)
    total = face_reactions(mc.ba, mc.inc_a[2 * k], mc.inc_a[2 * k + 1], pos, vel, mc.radius, mc.kf, mc.mu, total);
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

}  // namespace

void launch_mesh_contact(const MeshContactDev& mc, const float* pos32, const float* vel32, const double* pos64,
                         const double* vel64, double* target, const int* vbody, const double* mass, double dt,
                         hipStream_t stream) {
  const Vec3Src pos{pos32, pos64}, vel{vel32, vel64};
  hipLaunchKernelGGL(k_mc_zero, dim3(1), dim3(1), 0, stream, mc.pairs);
  for (const MeshContactSide* side : {&mc.ab, &mc.ba}) {
    const long long waves = static_cast<long long>(side->nv) * side->row_words;
    if (waves == 0) continue;
    hipLaunchKernelGGL(k_mc_rows, dim3(blocks_for(waves * 64)), dim3(256), 0, stream, *side, pos, vel, mc.radius,
                       mc.kf, mc.mu, mc.pairs);
    hipLaunchKernelGGL(k_mc_transpose, dim3(blocks_for(static_cast<long long>(side->nf) * side->col_words)),
                       dim3(256), 0, stream, *side);
  }
  hipLaunchKernelGGL(k_mc_gather, dim3(blocks_for(mc.n)), dim3(256), 0, stream, mc, pos, vel, target, vbody, mass,
                     dt);
}

}  // namespace dphd
