// gfx950 kernels of the DiffPhD projective-dynamics step and its adjoint.
//
// Reference kernels (paths under src/newton_solver_diff_phd/_src/solvers/diff_phd/):
//   begin_step_chunk    kernels.py:155-207 (begin_step, explicit_boundary)
//   local_tet4_chunk    kernels.py:222-272 (local_step)
//   local_hex8_chunk    hex.py:122-169 (one lane per Gauss sample, 8 lanes
//                       per cell summed by a lane butterfly)
//   gather_chunk        kernels.py:75-91 + global_rhs :275-284 + the
//                       projection half of the convergence gate :354-401
//   spmv_task           sparse.py:209-227 (CSR-vector, rows binned by length)
//   push_gram_chunk     anderson.py:46-94 (push + Gram partials) with
//                       impose_boundary kernels.py:287-291
//   gram_solve          anderson.py:150-215 (reduce + pivoted LDL^T solve)
//   apply_conv_chunk    anderson.py:218-243 + convergence_tiled :354-401
//   k_conv_check        kernels.py:322-347
//   k_write_state       kernels.py:404-414
//   adjoint kernels     adjoint.py:70-321, 474-659
//
// Iteration kernels take a work list: a fixed grid whose workgroups loop over
// the items (vertex chunks, element chunks or SpMV tasks) of the bodies that
// are still iterating, found through a prefix sum of the per-body active
// flags built in LDS at kernel start. Every reduction is a fixed-shape
// butterfly or a chunk-ordered sum: results are bitwise repeatable run to
// run, and no float atomics are used.
#include <stdexcept>

#include "contact.h"
#include "kernels.h"
#include "pd_math.h"

namespace dphd {
namespace {

constexpr int kRhsIndex = kMaxWindow * (kMaxWindow + 1) / 2;  // first right-hand-side entry of the Gram partials

__device__ __forceinline__ V3 ld3(const double* p, int i) { return V3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ void st3(double* p, int i, V3 v) {
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}

// parts[stride*c + off] summed over the chunks c0 <= c < c1 in chunk order,
// with the loads issued eight at a time (one round trip per eight chunks).
__device__ __forceinline__ double chunk_sum(const double* parts, int stride, int off, int c0, int c1) {
  double t = 0.0;
  for (int c = c0; c < c1; c += 8) {
    double p[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) p[u] = parts[static_cast<size_t>(stride) * min(c + u, c1 - 1) + off];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (c + u < c1) t += p[u];
  }
  return t;
}

// chunk_sum of a[stride_a * c] and of b[stride_b * c] at once, each in chunk
// order, the 16 loads of a round in flight together.
__device__ __forceinline__ void chunk_sum2(const double* a, int stride_a, const double* b, int stride_b, int c0, int c1,
                                           double& ta, double& tb) {
  ta = 0.0;
  tb = 0.0;
  for (int c = c0; c < c1; c += 8) {
    double pa[8], pb[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      pa[u] = a[static_cast<size_t>(stride_a) * min(c + u, c1 - 1)];
      pb[u] = b[static_cast<size_t>(stride_b) * min(c + u, c1 - 1)];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (c + u < c1) {
        ta += pa[u];
        tb += pb[u];
      }
  }
}

// chunk_sum of components 0 .. N-1 at once (each in chunk order, as
// chunk_sum adds it), all 8 x N loads of a round in flight together.
template <int N>
__device__ __forceinline__ void chunk_sums(const double* parts, int stride, int c0, int c1, double (&t)[N]) {
#pragma unroll
  for (int q = 0; q < N; ++q) t[q] = 0.0;
  for (int c = c0; c < c1; c += 8) {
    double p[8][N];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int q = 0; q < N; ++q) p[u][q] = parts[static_cast<size_t>(stride) * min(c + u, c1 - 1) + q];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (c + u < c1) {
#pragma unroll
        for (int q = 0; q < N; ++q) t[q] += p[u][q];
      }
  }
}

// Lane exchanges without LDS: the gfx950 row swaps (v_permlane32_swap
// exchanges the upper half of its first operand with the lower half of its
// second, v_permlane16_swap the odd 16-lane rows of the first with the even
// rows of the second) and DPP moves within 16-lane rows (row_mirror: lane
// l <- l ^ 15, row_half_mirror: l <- l ^ 7, quad_perm).
constexpr int kDppRowMirror = 0x140, kDppHalfMirror = 0x141;
constexpr int kDppXor2 = 0x4e;  // quad_perm [2, 3, 0, 1]
constexpr int kDppXor1 = 0xb1;  // quad_perm [1, 0, 3, 2]

template <int Ctrl>
__device__ __forceinline__ float dpp_move(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), Ctrl, 0xf, 0xf, false));
}
template <int Ctrl>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(b & 0xffffffffll), Ctrl, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(b >> 32), Ctrl, 0xf, 0xf, false);
  return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// a + b after swapping a's upper half-wavefront (Half = 32) or its odd
// 16-lane rows (Half = 16) with the matching lanes of b: lanes whose Half bit
// is clear get a(l) + a(l ^ Half), the others b(l ^ Half) + b(l).
template <int Half>
__device__ __forceinline__ float swap_sum(float a, float b) {
  const unsigned ua = __builtin_bit_cast(unsigned, a), ub = __builtin_bit_cast(unsigned, b);
  const auto r = Half == 32 ? __builtin_amdgcn_permlane32_swap(ua, ub, false, false)
                            : __builtin_amdgcn_permlane16_swap(ua, ub, false, false);
  return __builtin_bit_cast(float, static_cast<unsigned>(r[0])) + __builtin_bit_cast(float, static_cast<unsigned>(r[1]));
}
template <int Half>
__device__ __forceinline__ double swap_sum(double a, double b) {
  const unsigned long long ua = __builtin_bit_cast(unsigned long long, a), ub = __builtin_bit_cast(unsigned long long, b);
  const unsigned alo = static_cast<unsigned>(ua), ahi = static_cast<unsigned>(ua >> 32);
  const unsigned blo = static_cast<unsigned>(ub), bhi = static_cast<unsigned>(ub >> 32);
  const auto lo = Half == 32 ? __builtin_amdgcn_permlane32_swap(alo, blo, false, false)
                             : __builtin_amdgcn_permlane16_swap(alo, blo, false, false);
  const auto hi = Half == 32 ? __builtin_amdgcn_permlane32_swap(ahi, bhi, false, false)
                             : __builtin_amdgcn_permlane16_swap(ahi, bhi, false, false);
  const double na = __builtin_bit_cast(
      double, (static_cast<unsigned long long>(static_cast<unsigned>(hi[0])) << 32) | static_cast<unsigned>(lo[0]));
  const double nb = __builtin_bit_cast(
      double, (static_cast<unsigned long long>(static_cast<unsigned>(hi[1])) << 32) | static_cast<unsigned>(lo[1]));
  return na + nb;
}

// One halving step over a lane pairing l <-> l ^ mask (a DPP move): lanes
// with `upper` set keep b, the others a, each plus the partner's copy.
template <int Ctrl, typename T>
__device__ __forceinline__ T mirror_sum(T a, T b, bool upper) {
  const T keep = upper ? b : a, send = upper ? a : b;
  return keep + dpp_move<Ctrl>(send);
}

// Sum over the 64 lanes of each of 64 values: lane l returns the total of
// v[l]; the same halving with one more step (l ^ 1) at the end.
template <typename T>
__device__ __forceinline__ T wave_sum64(T (&v)[64], int lane) {
  T a[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) a[i] = swap_sum<32>(v[i], v[32 + i]);  // value 32*b5 + i
  T b[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) b[i] = swap_sum<16>(a[i], a[16 + i]);  // + 16*b4
  const bool b3 = (lane >> 3) & 1, b2 = (lane >> 2) & 1, b1 = (lane >> 1) & 1, b0 = lane & 1;
  T c[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) c[i] = mirror_sum<kDppRowMirror>(b[i], b[8 + i], b3);  // + 8*b3
  T d[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) d[i] = mirror_sum<kDppHalfMirror>(c[i], c[4 + i], b2);  // + 4*b2
  T e[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) e[i] = mirror_sum<kDppXor2>(d[i], d[2 + i], b1);  // + 2*b1
  return mirror_sum<kDppXor1>(e[0], e[1], b0);  // + b0
}

// Sum of one value over the 64 lanes, in every lane.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_move<kDppXor1>(v);
  v += dpp_move<kDppXor2>(v);
  v += dpp_move<kDppHalfMirror>(v);
  v += dpp_move<kDppRowMirror>(v);
  v = swap_sum<16>(v, v);
  return swap_sum<32>(v, v);
}

// Sum over the 64 lanes of each of 32 values by recursive halving (every
// step pairs lanes across one more bit: 32, 16, then l ^ 15, l ^ 7, l ^ 2
// within rows, and l ^ 1): lanes 2k and 2k + 1 return the total of v[k].
// 32 exchanges instead of 32 x 6 shuffles, none through LDS.
template <typename T>
__device__ __forceinline__ T wave_sum32(T (&v)[32], int lane) {
  T a[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) a[i] = swap_sum<32>(v[i], v[16 + i]);  // value 16*b5 + i
  T b[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) b[i] = swap_sum<16>(a[i], a[8 + i]);  // + 8*b4
  const bool b3 = (lane >> 3) & 1, b2 = (lane >> 2) & 1, b1 = (lane >> 1) & 1;
  T c[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) c[i] = mirror_sum<kDppRowMirror>(b[i], b[4 + i], b3);  // + 4*b3
  T d[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) d[i] = mirror_sum<kDppHalfMirror>(c[i], c[2 + i], b2);  // + 2*b2
  const T e = mirror_sum<kDppXor2>(d[0], d[1], b1);  // + b1
  return e + dpp_move<kDppXor1>(e);
}

// Sum NV values over the 256-thread block into result[0..NV) (LDS), visible
// to every thread after the final barrier. The wavefront sums run without
// LDS; a wavefront's totals then go to smem[wave * NV + k].
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* smem, double* result) {
  static_assert(NV <= 64, "at most one value per lane");
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  if constexpr (NV == 1) {
    const double t = wave_sum(v[0]);
    if (lane == 0) smem[wave] = t;
  } else if constexpr (NV <= 32) {
    double all[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) all[k] = k < NV ? v[k] : 0.0;
    const double t = wave_sum32(all, lane);
    if ((lane & 1) == 0 && (lane >> 1) < NV) smem[wave * NV + (lane >> 1)] = t;
  } else {
    double all[64];
#pragma unroll
    for (int k = 0; k < 64; ++k) all[k] = k < NV ? v[k] : 0.0;
    const double t = wave_sum64(all, lane);
    if (lane < NV) smem[wave * NV + lane] = t;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < NV; k += blockDim.x) {
    double t = smem[k];
    for (int w = 1; w < kBlock / kWave; ++w) t += smem[w * NV + k];
    result[k] = t;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Work lists.

// pref[b] = number of items of active bodies before body b; returns the
// total. `base` holds each body's item range. Whole block participates.
__device__ int active_prefix(const int* active, const int* base, int B, int* pref) {
  __shared__ int scan[kBlock];
  const int per = (B + kBlock - 1) / kBlock;
  const int b0 = min(B, static_cast<int>(threadIdx.x) * per), b1 = min(B, b0 + per);
  int local = 0;
  for (int b = b0; b < b1; ++b) local += active[b] ? base[b + 1] - base[b] : 0;
  scan[threadIdx.x] = local;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    const int t = threadIdx.x >= off ? scan[threadIdx.x - off] : 0;
    __syncthreads();
    scan[threadIdx.x] += t;
    __syncthreads();
  }
  int run = scan[threadIdx.x] - local;
  for (int b = b0; b < b1; ++b) {
    pref[b] = run;
    run += active[b] ? base[b + 1] - base[b] : 0;
  }
  if (threadIdx.x == kBlock - 1) pref[B] = scan[kBlock - 1];
  __syncthreads();
  return pref[B];
}

// Body whose active item range holds v: pref[b] <= v < pref[b + 1].
__device__ __forceinline__ int find_body(const int* pref, int B, int v) {
  int lo = 0, hi = B;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pref[mid] <= v)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Static split of `total` work-list items over the grid's workers
// (`workers` per workgroup). Workgroups are dispatched round-robin over the
// 8 XCDs (blockIdx % 8), so when the grid is a multiple of 8 each XCD takes
// one contiguous eighth of the items: the neighbouring items of a body, which
// read the same vertex data, then share one L2.
struct ItemRange {
  int begin, end, first, stride;
};

__device__ __forceinline__ ItemRange xcd_items(int total, int workers, int worker) {
  const int g = static_cast<int>(gridDim.x);
  if (g % 8 != 0) return ItemRange{0, total, static_cast<int>(blockIdx.x) * workers + worker, g * workers};
  const int x = static_cast<int>(blockIdx.x) % 8, r = static_cast<int>(blockIdx.x) / 8;
  const long long t = total;
  return ItemRange{static_cast<int>(t * x / 8), static_cast<int>(t * (x + 1) / 8), r * workers + worker,
                   (g / 8) * workers};
}

// Writes the prefix of `active` over chunks `base` to out [B + 1]; every
// thread of the workgroup calls it (dynamic LDS of B + 1 ints).
__device__ void write_prefix(const int* active, const int* base, int B, int* out) {
  extern __shared__ int pref[];
  __syncthreads();
  active_prefix(active, base, B, pref);
  for (int b = threadIdx.x; b <= B; b += blockDim.x) out[b] = pref[b];
}

// The S^T list for the next products (t_mode[0]): sliced while the active
// bodies hold fewer than `threshold` column groups (tcb: each body's groups).
// Every thread of the workgroup calls it.
__device__ void write_mode(const int* active, const int* tcb, int B, int t_switch, int s_switch, int* t_mode) {
  extern __shared__ int pref[];
  __syncthreads();
  const int groups = active_prefix(active, tcb, B, pref);
  if (threadIdx.x == 0) {
    t_mode[0] = groups < t_switch ? 1 : 0;
    t_mode[2] = groups < s_switch ? 1 : 0;
  }
}

__global__ __launch_bounds__(kBlock) void k_active_prefix(const int* active, const int* base0, int* out0,
                                                          const int* base1, int* out1, int B, const int* tcb,
                                                          int t_switch, int s_switch, int* t_mode) {
  write_prefix(active, base0, B, out0);
  if (base1) write_prefix(active, base1, B, out1);
  write_mode(active, tcb, B, t_switch, s_switch, t_mode);
}

// Loop `fn(item, body, part)` over P parts of each active item listed by the
// prefix `pref_g` [B + 1] (chunks `base`); the whole block iterates together
// (block-uniform bounds), so fn may use block barriers. A workgroup without
// items reads one value; the others copy the prefix to LDS for the body
// search.
template <int P, typename Fn>
__device__ __forceinline__ void for_active_parts(const int* pref_g, const int* base, int B, Fn fn) {
  extern __shared__ int pref[];
  const int total = P * pref_g[B];
  const ItemRange r = xcd_items(total, 1, 0);
  if (r.begin + r.first >= r.end) return;
  for (int b = threadIdx.x; b <= B; b += blockDim.x) pref[b] = pref_g[b];
  __syncthreads();
  for (int v = r.begin + r.first; v < r.end; v += r.stride) {
    const int cv = v / P;
    const int b = find_body(pref, B, cv);
    fn(base[b] + (cv - pref[b]), b, v % P);
  }
}

// `fn(item, body)` over the active items, one part each.
template <typename Fn>
__device__ __forceinline__ void for_active(const int* pref_g, const int* base, int B, Fn fn) {
  for_active_parts<1>(pref_g, base, B, [&](int item, int b, int) { fn(item, b); });
}

// Right-hand side of the S product, 4-wide at the vertex's postorder index:
// float in the FP32-factor mode, double otherwise.
__device__ __forceinline__ void store_rhs(const DeviceState& d, int i, V3 v) {
  const size_t r = static_cast<size_t>(d.gpost[i]);
  if (d.f32) {
    static_cast<float4*>(d.rhs4)[r] =
        make_float4(static_cast<float>(v.x), static_cast<float>(v.y), static_cast<float>(v.z), 0.0f);
  } else {
    double2* p = static_cast<double2*>(d.rhs4) + 2 * r;
    p[0] = make_double2(v.x, v.y);
    p[1] = make_double2(v.z, 0.0);
  }
}

__device__ __forceinline__ void tet_grads(const float* poses, int e, V3 (&g)[4]) {
  const float* p = poses + 9 * e;
  double r[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) r[t] = static_cast<double>(p[t]);
  g[0] = V3{-((r[0] + r[3]) + r[6]), -((r[1] + r[4]) + r[7]), -((r[2] + r[5]) + r[8])};
  g[1] = V3{r[0], r[1], r[2]};
  g[2] = V3{r[3], r[4], r[5]};
  g[3] = V3{r[6], r[7], r[8]};
}

__device__ __forceinline__ V3 hex_grad(const DeviceState& d, int b, int s, int j) {
  const double* G = d.hexg + 192 * b + 24 * s + 3 * j;
  return V3{G[0], G[1], G[2]};
}

__device__ __forceinline__ void add_outer(M3& f, V3 x, V3 g) {
  const double xx[3] = {x.x, x.y, x.z}, gg[3] = {g.x, g.y, g.z};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) f.a[i][j] = f.a[i][j] + xx[i] * gg[j];
}

// Sum the per-corner vectors of the 8 Gauss-sample lanes of a Hex8 cell; lane
// s then holds corner s's total.
__device__ __forceinline__ V3 hex_lane_sum(double (&out)[8][3], int s) {
#pragma unroll
  for (int off = 1; off < 8; off <<= 1)
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) out[j][k] += __shfl_xor(out[j][k], off, kWave);
  V3 r{0, 0, 0};
#pragma unroll
  for (int j = 0; j < 8; ++j)
    if (j == s) r = V3{out[j][0], out[j][1], out[j][2]};
  return r;
}

// ---------------------------------------------------------------------------
// Step begin / end (one workgroup per vertex chunk, every body).

__global__ __launch_bounds__(kBlock) void k_begin_step(DeviceState d, StepScalars s, const float* q,
                                                        const float* qd, const float* f, const float* dir) {
  const Chunk c = d.vchunks[blockIdx.x];
  const int b = c.body;
  if (threadIdx.x == 0 && static_cast<int>(blockIdx.x) == d.vcb[b]) {
    d.iter[b] = 0;
    d.active[b] = 1;
    d.aa[2 * b] = d.aa[2 * b + 1] = 0;
  }
  const int i = c.begin + threadIdx.x;
  if (i >= c.end) return;
  const double mass = d.mass[b];
  const double dt = s.dt;
  const bool pinned = d.free[i] == 0;
  for (int k = 0; k < 3; ++k) {
    const double start = static_cast<double>(q[3 * i + k]);
    const double vel = static_cast<double>(qd[3 * i + k]);
    const double acc = static_cast<double>(f[3 * i + k]) / mass + d.gravity[3 * b + k];
    d.q0[3 * i + k] = start;
    d.qt[3 * i + k] = start + dt * vel + dt * dt * acc;
    double bnd = 0.0;
    if (pinned) {
      bnd = d.integrate ? start + dt * vel : start;
      if (dir) bnd = static_cast<double>(dir[3 * i + k]);
    }
    d.xd[3 * i + k] = bnd;
    d.x[3 * i + k] = start;
    d.p[3 * i + k] = 0.0;
  }
}

__global__ __launch_bounds__(kBlock) void k_write_state(DeviceState d, double dt, float* q, float* qd,
                                                         double* checkpoint) {
  const Chunk c = d.vchunks[blockIdx.x];
  const int i = c.begin + threadIdx.x;
  if (i >= c.end) return;
  for (int k = 0; k < 3; ++k) {
    const double x = d.x[3 * i + k];
    q[3 * i + k] = static_cast<float>(x);
    qd[3 * i + k] = static_cast<float>((x - d.q0[3 * i + k]) / dt);
    if (checkpoint) checkpoint[3 * i + k] = x;
  }
}

// out[0]: whether any body is still active; out[1]: the most iterations any
// body has run (the host sizes the next solve's first burst of chunks by it).
__global__ __launch_bounds__(kBlock) void k_any_active(const int* active, const int* iters, int B, int* out) {
  __shared__ int flag[kBlock], most[kBlock];
  int mine = 0, top = 0;
  for (int b = threadIdx.x; b < B; b += blockDim.x) {
    mine |= active[b];
    top = max(top, iters[b]);
  }
  flag[threadIdx.x] = mine;
  most[threadIdx.x] = top;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (static_cast<int>(threadIdx.x) < off) {
      flag[threadIdx.x] |= flag[threadIdx.x + off];
      most[threadIdx.x] = max(most[threadIdx.x], most[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = flag[0];
    out[1] = most[0];
  }
}

// ---------------------------------------------------------------------------
// Local step.

// Copies the positions of element chunk ci's distinct vertices into the
// stage xs (3 doubles each, in the chunk's list order); every thread of the
// workgroup calls it. The first barrier keeps the previous chunk's readers
// ahead of the overwrite. Consecutive threads load consecutive components, so
// runs of consecutive vertex indices are read coalesced.
__device__ __forceinline__ void stage_vertices(const DeviceState& d, int ci, double* xs) {
  const int v0 = d.lv_off[ci], nv = d.lv_off[ci + 1] - v0;
  __syncthreads();
  for (int t = threadIdx.x; t < 3 * nv; t += blockDim.x) {
    const int v = t / 3;
    xs[t] = d.x[3 * d.lv_list[v0 + v] + (t - 3 * v)];
  }
  __syncthreads();
}

__device__ __forceinline__ V3 staged(const double* xs, unsigned j) { return V3{xs[3 * j], xs[3 * j + 1], xs[3 * j + 2]}; }

__device__ __forceinline__ void local_tet4_chunk(const DeviceState& d, int ci, int b, const double* targets,
                                                 const double* xs) {
  const Chunk c = d.echunks[ci];
  const int e = c.begin + threadIdx.x;
  if (e >= c.end) return;
  V3 g[4];
  tet_grads(d.poses, e, g);
  // The corners' positions in the stage: four 16-bit entries in one load.
  const uint2 li = reinterpret_cast<const uint2*>(d.lidx)[e];
  const unsigned l4[4] = {li.x & 0xffffu, li.x >> 16, li.y & 0xffffu, li.y >> 16};
  M3 f = zero3();
#pragma unroll
  for (int k = 0; k < 4; ++k) add_outer(f, staged(xs, l4[k]), g[k]);
  M3 fd = zero3();
  const unsigned mask = d.pinmask[e];
  if (mask) {
    const int* nd = d.idx + 4 * e;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (mask & (1u << k)) add_outer(fd, ld3(d.xd, nd[k]), g[k]);
  }
  SampleWeights w{d.wc[e], d.wv[e], d.wn[e], 0.0};
  V3 m{0, 0, 0};
  double target = 1.0;
  if (d.has_muscles) {
    w.wm = d.mw[e];
    m = ld3(d.mdir, e);
    target = targets ? targets[e] : 1.0;
  }
  const NhParams nh{d.nh[3 * b], d.nh[3 * b + 1], d.nh[3 * b + 2]};
  M3 stress;
  const int st = sample_stress(f, fd, w, nh, m, target, false, stress);
  if (st) atomicOr(d.status, st);
#pragma unroll
  for (int k = 0; k < 4; ++k) st3(d.contrib, 4 * e + k, mul(stress, g[k]));
}

// 32 cells per chunk, 8 lanes per cell (one per Gauss sample).
__device__ __forceinline__ void local_hex8_chunk(const DeviceState& d, int ci, int b, const double* targets,
                                                 const double* xs) {
  const Chunk c = d.echunks[ci];
  const int s = threadIdx.x & 7;
  const int e = c.begin + (threadIdx.x >> 3);
  const bool valid = e < c.end;
  double out[8][3];
#pragma unroll
  for (int j = 0; j < 8; ++j) out[j][0] = out[j][1] = out[j][2] = 0.0;
  if (valid) {
    // The eight corners' positions in the stage (read alike by the cell's 8 lanes).
    const uint4 li = reinterpret_cast<const uint4*>(d.lidx)[e];
    const unsigned l8[8] = {li.x & 0xffffu, li.x >> 16, li.y & 0xffffu, li.y >> 16,
                            li.z & 0xffffu, li.z >> 16, li.w & 0xffffu, li.w >> 16};
    M3 f = zero3();
#pragma unroll
    for (int j = 0; j < 8; ++j) add_outer(f, staged(xs, l8[j]), hex_grad(d, b, s, j));
    M3 fd = zero3();
    const unsigned mask = d.pinmask[e];
    const int* nd = d.idx + 8 * e;
    if (mask) {
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (mask & (1u << j)) add_outer(fd, ld3(d.xd, nd[j]), hex_grad(d, b, s, j));
    }
    SampleWeights w{d.wc[e], d.wv[e], d.wn[e], 0.0};
    V3 m{0, 0, 0};
    double target = 1.0;
    if (d.has_muscles) {
      w.wm = d.mw[e];
      m = ld3(d.mdir, e);
      target = targets ? targets[e] : 1.0;
    }
    const NhParams nh{d.nh[3 * b], d.nh[3 * b + 1], d.nh[3 * b + 2]};
    M3 stress;
    const int st = sample_stress(f, fd, w, nh, m, target, true, stress);
    if (st) atomicOr(d.status, st);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const V3 r = mul(stress, hex_grad(d, b, s, j));
      out[j][0] = r.x;
      out[j][1] = r.y;
      out[j][2] = r.z;
    }
  }
  const V3 sum = hex_lane_sum(out, s);
  if (valid) st3(d.contrib, 8 * e + s, sum);
}

// Dynamic LDS: the work-list prefix (B + 1 ints), then the vertex stage at
// the next 8-byte boundary.
__host__ __device__ __forceinline__ size_t stage_offset(int B) {
  return (sizeof(int) * static_cast<size_t>(B + 1) + 7) / 8 * 8;
}

// One element chunk per workgroup at a time: its distinct vertices are loaded
// once into LDS, then every element (Tet4) or Gauss sample (Hex8) reads its
// corners there.
__global__ __launch_bounds__(kBlock) void k_local(DeviceState d, const double* targets) {
  extern __shared__ int lds_base[];
  double* xs = reinterpret_cast<double*>(reinterpret_cast<char*>(lds_base) + stage_offset(d.B));
  for_active(d.pref_fe, d.ecb, d.B, [&](int ci, int b) {
    stage_vertices(d, ci, xs);
    if (d.Q == 1)
      local_tet4_chunk(d, ci, b, targets, xs);
    else
      local_hex8_chunk(d, ci, b, targets, xs);
  });
}

// ---------------------------------------------------------------------------
// Gather + right-hand side (+ projection-change partials).

// Row i of the FP64 constant matrix times x (FP32-factor mode): the row's
// entries eight at a time (columns and values, then the x rows, all in
// flight), added in column order.
__device__ __forceinline__ V3 a_row_times(const DeviceState& d, int i, const double* x) {
  V3 ax{0, 0, 0};
  const int q0 = d.a_off[i], q1 = d.a_off[i + 1];
  for (int q = q0; q < q1; q += 8) {
    int col[8];
    double w[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int qq = min(q + u, q1 - 1);
      col[u] = d.a_col[qq];
      w[u] = d.a_val[qq];
    }
    V3 xv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) xv[u] = ld3(x, col[u]);
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (q + u < q1) ax = ax + w[u] * xv[u];
  }
  return ax;
}

__global__ __launch_bounds__(kBlock) void k_gather(DeviceState d) {
  __shared__ double smem[2 * kBlock / kWave];
  __shared__ double res[2];
  for_active(d.pref_fv, d.vcb, d.B, [&](int ci, int b) {
    const Chunk c = d.vchunks[ci];
    const int i = c.begin + threadIdx.x;
    double part[2] = {0.0, 0.0};
    if (i < c.end) {
      // The incident entries eight at a time (entry loads, then their
      // contributions, all in flight), added in incidence order.
      V3 t{0, 0, 0};
      const int a0 = d.inc_off[i], a1 = d.inc_off[i + 1];
      for (int a = a0; a < a1; a += 8) {
        int ent[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) ent[u] = d.inc_ent[min(a + u, a1 - 1)];
        V3 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = ld3(d.contrib, ent[u]);
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (a + u < a1) t = t + v[u];
      }
      if (d.planar_w) {
        // planar_projection (hex.py:246-260): k (x - min(0, n.x + d) n - x_D).
        const double kw = d.planar_w[i];
        if (kw != 0.0) {
          const V3 xi = ld3(d.x, i), nrm{d.plane[0], d.plane[1], d.plane[2]};
          const double dist = dot(nrm, xi) + d.plane[3];
          const V3 projected = xi - (dist < 0.0 ? dist : 0.0) * nrm;
          t = t + kw * (projected - ld3(d.xd, i));
        }
      }
      const V3 old = ld3(d.p, i);
      V3 r = d.inertia[b] * ld3(d.qt, i) + t;
      if (d.defect) {
        const V3 ax = a_row_times(d, i, d.x);
        r = r - ax;
        if (d.fc_vslot) {
          const int c = d.fc_vslot[i];
          if (c >= 0) r = r + ld3(d.fc_fprev, c);
        }
      }
      st3(d.p, i, t);
      store_rhs(d, i, r);
      const V3 dp = t - old;
      part[0] = dot(dp, dp);
      part[1] = dot(old, old);
    }
    block_sum<2>(part, smem, res);
    if (threadIdx.x < 2) d.conv_part[4 * ci + 2 + threadIdx.x] = res[threadIdx.x];
  });
}

// ---------------------------------------------------------------------------
// Factor products y = S x and z = S^T y (+ base), sparse.py:209-227, over
// the dense supernode blocks. In both products a lane owns four consecutive
// columns of a block (one quad: a 16-byte float or 32-byte double value
// load per row, read nontemporally so that the vectors stay in L2):
//  - S: a wavefront takes up to 8 rows of one block, keeps the inputs of its
//    columns in registers across the rows, and reduces each row over the
//    lanes at the end (a butterfly over the lanes the block spans);
//  - S^T: a workgroup takes 256 columns of one body and walks every block
//    that covers them; the row value y_r is uniform, each lane accumulates
//    its own columns, and the four wavefronts (rows r = w mod 4) are summed
//    in LDS in wavefront order.
// Accumulation is in Acc (FP64, or FP32 in the FP32-factor mode, whose
// products only feed the defect correction); every sum has a fixed order,
// so results are bitwise repeatable. Values the factor does not hold are
// stored zeros and multiply vertices of the same elimination tree or zero
// gap slots, never another body's values.

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

template <bool NT>
__device__ __forceinline__ void load_quad(const double* p, double (&v)[4]) {
  const f64x2 a = NT ? __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(p)) : reinterpret_cast<const f64x2*>(p)[0];
  const f64x2 b = NT ? __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(p) + 1) : reinterpret_cast<const f64x2*>(p)[1];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
}

template <bool NT>
__device__ __forceinline__ void load_quad(const float* p, float (&v)[4]) {
  const f32x4 a = NT ? __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(p)) : reinterpret_cast<const f32x4*>(p)[0];
  v[0] = a.x;
  v[1] = a.y;
  v[2] = a.z;
  v[3] = a.w;
}

template <typename Acc>
__device__ __forceinline__ void load_vec4(const double* in, size_t j, Acc (&x)[3]) {
  const f64x2 a = *reinterpret_cast<const f64x2*>(in + 4 * j);
  x[0] = static_cast<Acc>(a.x);
  x[1] = static_cast<Acc>(a.y);
  x[2] = static_cast<Acc>(in[4 * j + 2]);
}

template <typename Acc>
__device__ __forceinline__ void load_vec4(const float* in, size_t j, Acc (&x)[3]) {
  const f32x4 a = *reinterpret_cast<const f32x4*>(in + 4 * j);
  x[0] = static_cast<Acc>(a.x);
  x[1] = static_cast<Acc>(a.y);
  x[2] = static_cast<Acc>(a.z);
}

template <typename T>
__device__ __forceinline__ void store_vec4(T* out, size_t row, V3 r) {
  T* p = out + 4 * row;
  p[0] = static_cast<T>(r.x);
  p[1] = static_cast<T>(r.y);
  p[2] = static_cast<T>(r.z);
  p[3] = static_cast<T>(0);
}

static_assert(kColsPerGroup == 4 * kWave, "a lane owns four columns of a column group");

// Lane `lane` of the calling wavefront's value v (a wavefront-uniform lane).
__device__ __forceinline__ int lane_int(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ long long lane_i64(long long v, int lane) {
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v & 0xffffffffll), lane));
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(v >> 32), lane);
  return (static_cast<long long>(hi) << 32) | lo;
}
__device__ __forceinline__ float lane_real(float v, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}
__device__ __forceinline__ double lane_real(double v, int lane) {
  return __builtin_bit_cast(double, lane_i64(__builtin_bit_cast(long long, v), lane));
}

// One S item on one wavefront: ROWS rows; a lane takes one quad of every
// row per pass, and all loads of a pass are issued before its products
// (rows past the item's count re-read its last row and are dropped). A
// chunk of a split item (part >= 0) stores its 32 sums to its partial slot
// instead of the rows.
template <int ROWS, typename Real, typename Acc, typename In>
__device__ __forceinline__ void row_item(const Real* __restrict__ val, const In* __restrict__ in, In* __restrict__ out,
                                         long long offset, int col0, int quads, int stride, int out_row, int count,
                                         int part, Acc* __restrict__ part_buf, int lane) {
  Acc acc[ROWS][3];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) acc[r][0] = acc[r][1] = acc[r][2] = Acc(0);
  const size_t width = 4 * static_cast<size_t>(stride);
  const Real* rows = val + offset;
  for (int qi = lane; qi < quads; qi += kWave) {
    Acc x[4][3];
#pragma unroll
    for (int u = 0; u < 4; ++u) load_vec4(in, static_cast<size_t>(col0 + 4 * qi + u), x[u]);
    Real v[ROWS][4];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) load_quad<false>(rows + width * min(r, count - 1) + 4 * qi, v[r]);
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      if (r < count) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          acc[r][0] += static_cast<Acc>(v[r][u]) * x[u][0];
          acc[r][1] += static_cast<Acc>(v[r][u]) * x[u][1];
          acc[r][2] += static_cast<Acc>(v[r][u]) * x[u][2];
        }
      }
    }
  }
  // Row r's component c is value 4r + c (c = 3 is the zero fourth lane of
  // the 4-wide output); after the wavefront sum, lanes 2k and 2k + 1 hold
  // value k.
  static_assert(ROWS * 4 == 32, "the wavefront sum reduces 32 values");
  Acc vals[32];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    vals[4 * r + 0] = acc[r][0];
    vals[4 * r + 1] = acc[r][1];
    vals[4 * r + 2] = acc[r][2];
    vals[4 * r + 3] = Acc(0);
  }
  const Acc total = wave_sum32(vals, lane);
  const int k = lane >> 1;
  if ((lane & 1) == 0) {
    if (part >= 0)
      part_buf[static_cast<size_t>(part) * 32 + k] = total;
    else if ((k >> 2) < count)
      out[4 * static_cast<size_t>(out_row) + k] = static_cast<In>(total);
  }
}

// Sum over aligned groups of 2^s lanes (s <= 4), in every lane of the
// group: the 16- and 8-lane mirrors (l ^ 15, l ^ 7), then quad_perm xor 2
// and xor 1, the steps of wave_sum32 below its row swaps. A row whose quads
// sit one per lane from its group's first lane is thereby summed exactly as
// wave_sum32 sums a row with one quad per lane (the other lanes add zeros),
// so packing leaves every S-product value bitwise unchanged.
template <typename T>
__device__ __forceinline__ T segment_sum(T v, int s) {
  if (s > 3) v += dpp_move<kDppRowMirror>(v);
  if (s > 2) v += dpp_move<kDppHalfMirror>(v);
  if (s > 1) v += dpp_move<kDppXor2>(v);
  if (s > 0) v += dpp_move<kDppXor1>(v);
  return v;
}

__device__ __forceinline__ void store_sums(float* out, size_t row, const float (&t)[3]) {
  *reinterpret_cast<f32x4*>(out + 4 * row) = f32x4{t[0], t[1], t[2], 0.0f};
}
__device__ __forceinline__ void store_sums(double* out, size_t row, const double (&t)[3]) {
  f64x2* p = reinterpret_cast<f64x2*>(out + 4 * row);
  p[0] = f64x2{t[0], t[1]};
  p[1] = f64x2{t[2], 0.0};
}

// One S item of a narrow block (at most kPackedQuads quads, padded to 2^s
// lanes per row) on one wavefront: lane l takes quad l & (2^s - 1) of row
// l >> s in each of kPackedPasses passes (64 >> s rows a pass), so a lane's
// four input columns are loaded once for the item and every pass's value
// load is in flight together; a row's quads are then summed over its 2^s
// lanes and the row's first lane stores it.
template <typename Real, typename Acc, typename In>
__device__ __forceinline__ void row_item_packed(const Real* __restrict__ val, const In* __restrict__ in,
                                                In* __restrict__ out, long long offset, int col0, int quads,
                                                int stride, int out_row, int count, int s, int lane) {
  const int q = lane & ((1 << s) - 1), rl = lane >> s, rp = kWave >> s;
  const int qc = min(q, quads - 1);
  Acc x[4][3];
#pragma unroll
  for (int u = 0; u < 4; ++u) load_vec4(in, static_cast<size_t>(col0 + 4 * qc + u), x[u]);
  const Real* rows = val + offset + 4 * qc;
  const size_t width = 4 * static_cast<size_t>(stride);
  Real v[kPackedPasses][4];
#pragma unroll
  for (int p = 0; p < kPackedPasses; ++p) load_quad<false>(rows + width * min(p * rp + rl, count - 1), v[p]);
  Acc acc[kPackedPasses][3];
#pragma unroll
  for (int p = 0; p < kPackedPasses; ++p) {
    const bool use = q < quads && p * rp + rl < count;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      Acc t = Acc(0);
#pragma unroll
      for (int u = 0; u < 4; ++u) t += static_cast<Acc>(v[p][u]) * x[u][c];
      acc[p][c] = segment_sum(use ? t : Acc(0), s);
    }
  }
#pragma unroll
  for (int p = 0; p < kPackedPasses; ++p) {
    const int r = p * rp + rl;
    if (q == 0 && r < count) {
      In t[3] = {static_cast<In>(acc[p][0]), static_cast<In>(acc[p][1]), static_cast<In>(acc[p][2])};
      store_sums(out, static_cast<size_t>(out_row + r), t);
    }
  }
}

// An S or S^T item list: items, per-worker ranges woff into list, and the
// workgroups it was balanced over.
struct RowLists {
  const RowItem* items;
  const int* woff;
  const int* list;
  int grid;
};

// S product over the list *mode selects (0 whole items, 1 the split list):
// wavefront w streams the items list[woff[w] .. woff[w + 1]) whose body is
// active (workgroups beyond the list's grid return). Lane l holds the
// descriptor of the l-th item of a window of 64, so the descriptors of a
// window cost one load round trip.
template <typename Real, typename Acc, typename In>
__global__ __launch_bounds__(kBlock) void k_factor_rows(
    RowLists whole, RowLists split, const Real* __restrict__ val, const In* __restrict__ in, In* __restrict__ out,
    const int* __restrict__ active, const int* __restrict__ mode, Acc* __restrict__ part_buf) {
  const RowLists L = *mode != 0 ? split : whole;
  if (static_cast<int>(blockIdx.x) >= L.grid) return;
  const RowItem* __restrict__ items = L.items;
  const int* __restrict__ list = L.list;
  const int lane = static_cast<int>(threadIdx.x) % kWave;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(blockIdx.x) * (kBlock / kWave) +
                                                  static_cast<int>(threadIdx.x) / kWave);
  const int l0 = L.woff[wave], l1 = L.woff[wave + 1];
  for (int w0 = l0; w0 < l1; w0 += kWave) {
    RowItem mine{};
    bool on = false;
    if (w0 + lane < l1) {
      mine = items[list[w0 + lane]];
      on = active[mine.body] != 0;
    }
    unsigned long long pending = __ballot(on);
    while (pending) {
      const int j = __builtin_amdgcn_readfirstlane(static_cast<int>(__builtin_ctzll(pending)));
      pending &= pending - 1;
      const long long offset = lane_i64(mine.offset, j);
      const int col0 = lane_int(mine.col0, j), quads = lane_int(mine.quads, j);
      const int out_row = lane_int(mine.out_row, j), count = lane_int(mine.count, j);
      const int part = lane_int(mine.part, j), stride = lane_int(mine.stride, j);
      const int packed = lane_int(mine.packed, j);
      if (packed > 0)
        row_item_packed<Real, Acc>(val, in, out, offset, col0, quads, stride, out_row, count, packed - 1, lane);
      else
        row_item<kRowsPerGroup, Real, Acc>(val, in, out, offset, col0, quads, stride, out_row, count, part, part_buf,
                                           lane);
    }
  }
}

// The split S items: a thread per (item, value), striding over the grid,
// adds the chunks' partial sums in column order and stores the value (rows
// within the item's count). The grid is small, so that while the whole list
// runs this launch costs little more than its dispatch.
template <typename Acc, typename In>
__global__ __launch_bounds__(kBlock) void k_rows_combine(const RowCombine* __restrict__ items, int count,
                                                         const Acc* __restrict__ part_buf, In* __restrict__ out,
                                                         const int* __restrict__ active, const int* __restrict__ mode) {
  if (*mode == 0) return;
  const int stride = static_cast<int>(gridDim.x * blockDim.x);
  for (int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); t < 32 * count; t += stride) {
    const int w = t >> 5, k = t & 31;
    const RowCombine it = items[w];
    if (!active[it.body] || (k >> 2) >= it.count) continue;
    // Chunks eight at a time, all their loads in flight, added in column order.
    Acc acc = Acc(0);
    const int p1 = it.part0 + it.parts;
    for (int p = it.part0; p < p1; p += 8) {
      Acc v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = part_buf[static_cast<size_t>(min(p + u, p1 - 1)) * 32 + k];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (p + u < p1) acc += v[u];
    }
    out[4 * static_cast<size_t>(it.out_row) + k] = static_cast<In>(acc);
  }
}

// The adjoint's safe-clip pass over the S^T output (adjoint.py:184-204,
// _clip_partials): with `free` set, Dirichlet rows are stored as zero,
// non-finite components as zero (status 16), and each column group's
// squared norm goes to norm[item] (its lanes' columns in order, then the
// wavefront sum), which the next kernel sums over the body's groups.
struct AdjointEpilogue {
  const uint8_t* free = nullptr;
  double* norm = nullptr;
  int* status = nullptr;
};

// S^T product: workgroup g takes the column groups t_list[woff[g] ..
// woff[g + 1]) whose body is active (lane l of every wavefront holds the
// l-th group of a window of 64); its kColWaves wavefronts split a group's
// rows.
// The lane's four columns c .. c + 3 of a column group (body columns end
// at gend) from their sums acc: plus the base rows, the adjoint's clip pass,
// stored at their vertices; with the clip pass, the group's squared norm to
// adj.norm[gid]. The calling wavefront's 64 lanes all take part.
template <typename Acc>
__device__ __forceinline__ void cols_store(const Acc (&acc)[4][3], int c, int gend, const int* __restrict__ gorig,
                                           const double* base, double* __restrict__ out, const AdjointEpilogue& adj,
                                           int gid, int lane) {
  // A body's slots are a multiple of 4, so the lane's four columns are all
  // the body's or none; their vertices in one load, then the base rows of
  // all four together.
  double sq = 0.0;
  bool scrubbed = false;
  if (c < gend) {
    const int4 o4 = *reinterpret_cast<const int4*>(gorig + c);
    const int o[4] = {o4.x, o4.y, o4.z, o4.w};
    V3 add[4] = {};
    if (base) {
#pragma unroll
      for (int u = 0; u < 4; ++u) add[u] = ld3(base, o[u] >= 0 ? o[u] : 0);
    }
    uint8_t keep[4] = {1, 1, 1, 1};
    if (adj.free) {
#pragma unroll
      for (int u = 0; u < 4; ++u) keep[u] = o[u] >= 0 ? adj.free[o[u]] : 0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (o[u] >= 0) {
        V3 v = V3{static_cast<double>(acc[u][0]), static_cast<double>(acc[u][1]), static_cast<double>(acc[u][2])} +
               add[u];
        if (adj.free) {
          if (!keep[u]) v = V3{0.0, 0.0, 0.0};
          if (!isfinite(v.x)) { v.x = 0.0; scrubbed = true; }
          if (!isfinite(v.y)) { v.y = 0.0; scrubbed = true; }
          if (!isfinite(v.z)) { v.z = 0.0; scrubbed = true; }
          sq += dot(v, v);
        }
        st3(out, o[u], v);
      }
  }
  if (adj.free) {
    const double total = wave_sum(sq);
    const bool any_scrubbed = __ballot(scrubbed) != 0;
    if (lane == 0) {
      adj.norm[gid] = total;
      if (any_scrubbed) atomicOr(adj.status, 16);
    }
  }
}

struct ColLists {
  const ColGroup* items;
  const int* woff;
  const int* list;
  int grid;
};

// S^T over the column-group list *mode selects (0 whole groups, 1 the sliced
// list; workgroups beyond the list's grid return). Sliced items with a
// partial slot store their sums to part_buf[slot][12][64] (coalesced) for the
// combine kernel.
template <typename Real, typename Acc, typename In>
__global__ __launch_bounds__(kColWaves * kWave) void k_factor_cols(
    ColLists whole, ColLists sliced, const ColEntry* __restrict__ entries, const Real* __restrict__ val,
    const In* __restrict__ y, double* __restrict__ out, const int* __restrict__ gorig, const double* base,
    const int* __restrict__ active, AdjointEpilogue adj, const int* __restrict__ mode, Acc* __restrict__ part_buf) {
  const ColLists L = *mode != 0 ? sliced : whole;
  if (static_cast<int>(blockIdx.x) >= L.grid) return;
  const ColGroup* __restrict__ items = L.items;
  const int* __restrict__ woff = L.woff;
  const int* __restrict__ list = L.list;
  // Rows per batch: 16 with FP32 values, 8 with FP64 (the same bytes).
  constexpr int kWaves = kColWaves, kBatch = sizeof(Real) == 4 ? 16 : 8;
  static_assert(kBatch <= kWave, "a batch's entries are held one per lane");
  __shared__ Acc part[kWaves - 1][12][kWave];
  const int lane = static_cast<int>(threadIdx.x) % kWave;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
  const int l0 = woff[blockIdx.x], l1 = woff[blockIdx.x + 1];
  for (int g0 = l0; g0 < l1; g0 += kWave) {
    ColGroup mine{};
    int mine_id = 0;
    bool on = false;
    if (g0 + lane < l1) {
      mine_id = list[g0 + lane];
      mine = items[mine_id];
      on = active[mine.body] != 0;
    }
    // The same in every wavefront of the workgroup, so the barriers below
    // are reached by all of them.
    unsigned long long pending = __ballot(on);
    while (pending) {
      const int j = __builtin_amdgcn_readfirstlane(static_cast<int>(__builtin_ctzll(pending)));
      pending &= pending - 1;
      const int gc0 = lane_int(mine.c0, j), ge0 = lane_int(mine.e0, j), ge1 = lane_int(mine.e1, j);
      const int gend = lane_int(mine.end, j), gpart = lane_int(mine.part, j), gid = lane_int(mine.gid, j);
      const int c = gc0 + 4 * lane;
      Acc acc[4][3];
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u][0] = acc[u][1] = acc[u][2] = Acc(0);
      // Wavefront w takes the w-th share of the group's rows, in batches of
      // kBatch rows (from any number of blocks). Lane k < kBatch holds the
      // k-th entry of a batch, loaded while the batch before it streams, so
      // a batch costs one round trip; rows past the share re-read its last
      // row and are dropped.
      const int count = ge1 - ge0, share = (count + kWaves - 1) / kWaves;
      const int w0 = ge0 + min(wave * share, count), w1 = ge0 + min((wave + 1) * share, count);
      const int slot = lane % kBatch;
      ColEntry cur{};
      if (w0 < w1) cur = entries[min(w0 + slot, w1 - 1)];
      for (int e = w0; e < w1; e += kBatch) {
        // The batch's values first, then lane k's S-product row (a vector
        // load, no scalar-cache round trip of its own) and the next batch's
        // entries: all in flight together.
        Real v[kBatch][4];
        bool use[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
          const long long rowoff = lane_i64(cur.rowoff, k);
          const int col0 = lane_int(cur.col0, k), quads = lane_int(cur.quads, k);
          const int qi = (c - col0) >> 2;
          use[k] = e + k < w1 && c >= col0 && qi < quads;
          load_quad<false>(val + rowoff + 4 * static_cast<long long>(use[k] ? qi : 0), v[k]);
        }
        Acc ymine[3];
        load_vec4(y, static_cast<size_t>(cur.yrow), ymine);
        // (After the last batch this re-reads the share's last entry, unused.)
        const ColEntry nxt = entries[min(e + kBatch + slot, w1 - 1)];
        Acc yr[kBatch][3];
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
          yr[k][0] = lane_real(ymine[0], k);
          yr[k][1] = lane_real(ymine[1], k);
          yr[k][2] = lane_real(ymine[2], k);
        }
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const Acc w = use[k] ? static_cast<Acc>(v[k][u]) : Acc(0);
            acc[u][0] += w * yr[k][0];
            acc[u][1] += w * yr[k][1];
            acc[u][2] += w * yr[k][2];
          }
        }
        cur = nxt;
      }
      if (wave > 0) {
#pragma unroll
        for (int t = 0; t < 12; ++t) part[wave - 1][t][lane] = acc[t / 3][t % 3];
      }
      __syncthreads();
      if (wave == 0) {
#pragma unroll
        for (int w = 0; w < kWaves - 1; ++w)
#pragma unroll
          for (int t = 0; t < 12; ++t) acc[t / 3][t % 3] += part[w][t][lane];
        if (gpart >= 0) {
          Acc* dst = part_buf + static_cast<size_t>(gpart) * 12 * kWave + lane;
#pragma unroll
          for (int t = 0; t < 12; ++t) dst[t * kWave] = acc[t / 3][t % 3];
        } else {
          cols_store<Acc>(acc, c, gend, gorig, base, out, adj, gid, lane);
        }
      }
      __syncthreads();
    }
  }
}

// The groups cut into slices: a wavefront per group, striding over the
// grid's wavefronts, adds its slices' partial sums in row order, then stores
// its columns as k_factor_cols does. A small grid, as for k_rows_combine.
template <typename Acc>
__global__ __launch_bounds__(kBlock) void k_factor_combine(const ColCombine* __restrict__ items, int count,
                                                           const Acc* __restrict__ part_buf, double* __restrict__ out,
                                                           const int* __restrict__ gorig, const double* base,
                                                           const int* __restrict__ active, AdjointEpilogue adj,
                                                           const int* __restrict__ mode) {
  if (*mode != 1) return;
  const int lane = static_cast<int>(threadIdx.x) % kWave;
  const int waves = static_cast<int>(gridDim.x) * (kBlock / kWave);
  for (int w = __builtin_amdgcn_readfirstlane(static_cast<int>(blockIdx.x) * (kBlock / kWave) +
                                              static_cast<int>(threadIdx.x) / kWave);
       w < count; w += waves) {
    const ColCombine it = items[w];
    if (!active[it.body]) continue;
    Acc acc[4][3];
#pragma unroll
    for (int t = 0; t < 12; ++t) acc[t / 3][t % 3] = Acc(0);
    // Slices four at a time, all their loads in flight, added in slice order.
    const int p1 = it.part0 + it.parts;
    for (int p = it.part0; p < p1; p += 4) {
      Acc v[4][12];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const Acc* src = part_buf + static_cast<size_t>(min(p + u, p1 - 1)) * 12 * kWave + lane;
#pragma unroll
        for (int t = 0; t < 12; ++t) v[u][t] = src[t * kWave];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (p + u < p1) {
#pragma unroll
          for (int t = 0; t < 12; ++t) acc[t / 3][t % 3] += v[u][t];
        }
    }
    cols_store<Acc>(acc, it.c0 + 4 * lane, it.end, gorig, base, out, adj, it.gid, lane);
  }
}

// r = rhs - A x with the FP64 constant matrix (FP32-factor mode), one thread
// per vertex, all three coordinates.
__global__ __launch_bounds__(kBlock) void k_residual_a(DeviceState d, const double* rhs, const double* x, double* r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.N) return;
  st3(r, i, ld3(rhs, i) - a_row_times(d, i, x));
}

__global__ __launch_bounds__(kBlock) void k_add3(int n, const double* x, double* y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) st3(y, i, ld3(y, i) + ld3(x, i));
}

// rhs [N*3] FP64 (original order) -> the 4-wide postorder right-hand side.
template <typename In>
__global__ __launch_bounds__(kBlock) void k_pack4(int n, const int* gpost, const double* rhs, In* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  store_vec4(out, static_cast<size_t>(gpost[i]), ld3(rhs, i));
}

// ---------------------------------------------------------------------------
// Anderson mixing. The ring state (count, start) is double-buffered by the
// parity of the body's iteration count: the push and apply of iteration k
// read aa[k & 1], and the apply's writer stores the updated state in
// aa[(k + 1) & 1], so no workgroup can observe a half-updated state.

__device__ __forceinline__ int gram_index(int a, int b) {  // a <= b < 8
  return a * kMaxWindow - (a * (a - 1)) / 2 + (b - a);
}

__device__ __forceinline__ int parity_base(const DeviceState& d, const int* iters, int b) {
  return 2 * ((iters[b] & 1) * d.B + b);
}

// Push (x_prev, g = raw - x_prev) and accumulate this vertex's Gram terms.
__device__ __forceinline__ void push_and_gram(const DeviceState& d, int b, int i, V3 prev, V3 raw, const int* iters,
                                              double (&acc)[kGramEntries]) {
  const int pb = parity_base(d, iters, b);
  const int count = d.aa[pb], start = d.aa[pb + 1];
  const int cap = d.W + 1;
  const int slot = (start + count) % cap;
  const int N = d.N;
  const V3 g = raw - prev;
  st3(d.hg, slot * N + i, g);
  st3(d.hr, slot * N + i, prev + g);
  V3 dg[kMaxWindow];
  V3 lo = ld3(d.hg, (start % cap) * N + i);
#pragma unroll
  for (int a = 0; a < kMaxWindow; ++a) {
    if (a < count) {
      const int s1 = (start + a + 1) % cap;
      const V3 hi = (s1 == slot) ? g : ld3(d.hg, s1 * N + i);
      dg[a] = hi - lo;
      lo = hi;
    } else {
      dg[a] = V3{0, 0, 0};
    }
  }
#pragma unroll
  for (int a = 0; a < kMaxWindow; ++a) {
#pragma unroll
    for (int bb = a; bb < kMaxWindow; ++bb) acc[gram_index(a, bb)] += dot(dg[a], dg[bb]);
    acc[kRhsIndex + a] += dot(dg[a], g);
  }
}

// Mixing coefficients of a ring of N differences (anderson.py:150-215):
// the regularized Gram system solved in registers; returns |gamma|^2. The
// Gram entries are read from the reduced partials with compile-time indices.
template <int N>
__device__ __forceinline__ double anderson_coefficients(const double* red, double* gamma) {
  double a[N][N], b[N], x[N];
  double trace = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) trace += red[gram_index(i, i)];
  const double reg = fmax(1.0e-6 * trace / static_cast<double>(N), 1.0e-12);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    b[i] = red[kRhsIndex + i];
#pragma unroll
    for (int j = 0; j < N; ++j) a[i][j] = red[i <= j ? gram_index(i, j) : gram_index(j, i)];
    a[i][i] = a[i][i] + reg;
  }
  ldlt_solve_regs<N>(a, b, x);
  double norm = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    gamma[i] = x[i];
    norm += x[i] * x[i];
  }
  return norm;
}

// Reduce the body's Gram partials in chunk order and run the reference
// _solve: mixing coefficients and the ring update. Results in LDS:
// gamma[0..count), meta[0] = accepted, meta[1] = count used by the apply,
// meta[2] = its start. The writer stores the updated ring state.
__device__ void gram_solve(const DeviceState& d, int b, const int* iters, bool writer, double* red, double* gamma,
                           int* meta, int* status) {
  const int pb = parity_base(d, iters, b);
  const int c0 = d.vcb[b], c1 = d.vcb[b + 1];
  for (int k = threadIdx.x; k < kGramEntries; k += blockDim.x) red[k] = chunk_sum(d.gram_part, kGramEntries, k, c0, c1);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int W = d.W, cap = W + 1;
    int count = d.aa[pb], start = d.aa[pb + 1];
    meta[1] = count;
    meta[2] = start;
    int accept = 0;
    if (count >= 1) {
      double norm = 0.0;
      switch (count) {
        case 1: norm = anderson_coefficients<1>(red, gamma); break;
        case 2: norm = anderson_coefficients<2>(red, gamma); break;
        case 3: norm = anderson_coefficients<3>(red, gamma); break;
        case 4: norm = anderson_coefficients<4>(red, gamma); break;
        case 5: norm = anderson_coefficients<5>(red, gamma); break;
        case 6: norm = anderson_coefficients<6>(red, gamma); break;
        case 7: norm = anderson_coefficients<7>(red, gamma); break;
        default: norm = anderson_coefficients<8>(red, gamma); break;
      }
      if (sqrt(norm) >= 10.0) {
        count = 0;
        start = 0;
      } else {
        accept = 1;
        count += 1;
      }
    } else {
      count += 1;
    }
    while (count > W) {
      start = (start + 1) % cap;
      count -= 1;
    }
    meta[0] = accept;
    if (writer) {
      const int nb = 2 * ((((iters[b] + 1) & 1)) * d.B + b);
      d.aa[nb] = count;
      d.aa[nb + 1] = start;
    }
  }
  __syncthreads();
  (void)status;
}

__device__ __forceinline__ V3 anderson_value(const DeviceState& d, int i, V3 raw, const double* gamma,
                                             const int* meta) {
  if (d.W == 0 || !meta[0]) return raw;
  const int cap = d.W + 1, N = d.N;
  const int sc = meta[1], ss = meta[2];
  V3 v = ld3(d.hr, ((ss + sc) % cap) * N + i);
  V3 lo = ld3(d.hr, (ss % cap) * N + i);
  for (int a = 0; a < sc; ++a) {
    const V3 hi = ld3(d.hr, ((ss + a + 1) % cap) * N + i);
    v = v - gamma[a] * (hi - lo);
    lo = hi;
  }
  return v;
}

__global__ __launch_bounds__(kBlock) void k_push_gram(DeviceState d) {
  __shared__ double smem[kGramEntries * kBlock / kWave];
  __shared__ double res[kGramEntries];
  for_active(d.pref_fv, d.vcb, d.B, [&](int ci, int b) {
    const Chunk c = d.vchunks[ci];
    const int i = c.begin + threadIdx.x;
    double acc[kGramEntries];
#pragma unroll
    for (int k = 0; k < kGramEntries; ++k) acc[k] = 0.0;
    if (i < c.end) {
      V3 raw = ld3(d.xr, i);
      if (!d.free[i]) {
        raw = ld3(d.xd, i);
        st3(d.xr, i, raw);
      }
      push_and_gram(d, b, i, ld3(d.x, i), raw, d.iter, acc);
    }
    block_sum<kGramEntries>(acc, smem, res);
    for (int k = threadIdx.x; k < kGramEntries; k += blockDim.x) d.gram_part[kGramEntries * ci + k] = res[k];
  });
}

__global__ __launch_bounds__(kBlock) void k_apply_conv(DeviceState d) {
  __shared__ double smem[2 * kBlock / kWave];
  __shared__ double res[2];
  __shared__ double red[kGramEntries];
  __shared__ double gamma[kMaxWindow];
  __shared__ int meta[3];
  for_active(d.pref_fv, d.vcb, d.B, [&](int ci, int b) {
    if (d.W > 0) {
      gram_solve(d, b, d.iter, ci == d.vcb[b], red, gamma, meta, d.status);
    } else if (threadIdx.x == 0) {
      meta[0] = 0;
    }
    __syncthreads();
    const Chunk c = d.vchunks[ci];
    const int i = c.begin + threadIdx.x;
    double part[2] = {0.0, 0.0};
    if (i < c.end) {
      const V3 prev = ld3(d.x, i);
      V3 v = anderson_value(d, i, ld3(d.xr, i), gamma, meta);
      if (!d.free[i]) v = ld3(d.xd, i);
      const V3 step = v - prev;
      part[0] = dot(step, step);
      part[1] = dot(prev, prev);
      st3(d.x, i, v);
    }
    block_sum<2>(part, smem, res);
    if (threadIdx.x < 2) d.conv_part[4 * ci + threadIdx.x] = res[threadIdx.x];
  });
}

// Prefix sums over the bodies of NC per-body counts at once, one workgroup
// (every thread calls it): count c of body b is base[c][b + 1] - base[c][b]
// when flag[b] (LDS) is set, else 0; out[c][b] receives the counts of the
// bodies before b and out[c][B] their total (a null out[c] is not stored);
// tot[c] = the total, in every thread. Thread t takes a contiguous range of
// bodies; its range totals are scanned within the wavefront by shuffles and
// across the four wavefronts through LDS (integer sums, exact in any order).
template <int NC>
__device__ void body_prefixes(const int* flag, const int* const (&base)[NC], int* const (&out)[NC], int B,
                              int (&tot)[NC]) {
  __shared__ int wtot[NC][kBlock / kWave];
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const int per = (B + kBlock - 1) / kBlock;
  const int b0 = min(B, static_cast<int>(threadIdx.x) * per), b1 = min(B, b0 + per);
  int local[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) local[c] = 0;
  for (int b = b0; b < b1; ++b)
    if (flag[b]) {
#pragma unroll
      for (int c = 0; c < NC; ++c) local[c] += base[c][b + 1] - base[c][b];
    }
  int incl[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) incl[c] = local[c];
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int t = __shfl_up(incl[c], off, kWave);
      if (lane >= off) incl[c] += t;
    }
  if (lane == kWave - 1) {
#pragma unroll
    for (int c = 0; c < NC; ++c) wtot[c][wave] = incl[c];
  }
  __syncthreads();
  int run[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    run[c] = incl[c] - local[c];
    tot[c] = 0;
    for (int w = 0; w < kBlock / kWave; ++w) {
      if (w < wave) run[c] += wtot[c][w];
      tot[c] += wtot[c][w];
    }
  }
  for (int b = b0; b < b1; ++b) {
    const bool on = flag[b] != 0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (out[c]) out[c][b] = run[c];
      if (on) run[c] += base[c][b + 1] - base[c][b];
    }
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < NC; ++c)
      if (out[c]) out[c][B] = tot[c];
  }
}

// Every body's convergence test (one thread per body, one workgroup), then
// the work-list prefixes of the bodies still active.
__global__ __launch_bounds__(kBlock) void k_conv_check(DeviceState d, StepScalars s) {
  extern __shared__ int flag[];  // [B]: still active after this test
  for (int b = threadIdx.x; b < d.B; b += kBlock) {
    flag[b] = 0;
    if (!d.active[b]) continue;
    const int c0 = d.vcb[b], c1 = d.vcb[b + 1];
    double red[4];
    chunk_sums<4>(d.conv_part, 4, c0, c1, red);
    const int k = d.iter[b];
    const bool step_small = sqrt(red[0]) <= s.rel_tol * sqrt(red[1]) + s.abs_tol;
    const bool rhs_small = sqrt(red[2]) <= s.rel_tol * sqrt(red[3]) + s.abs_tol;
    d.iter[b] = k + 1;
    if (k > 0 && step_small && rhs_small) {
      d.active[b] = 0;
    } else if (k + 1 >= s.max_iterations) {
      d.active[b] = 0;
      atomicOr(d.status, 1);
    } else {
      flag[b] = 1;
    }
  }
  __syncthreads();
  int tot[3];
  body_prefixes<3>(flag, {d.vcb, d.ecb, d.tcb}, {d.pref_fv, d.pref_fe, nullptr}, d.B, tot);
  if (threadIdx.x == 0) {
    d.t_mode[0] = tot[2] < d.t_switch ? 1 : 0;
    d.t_mode[2] = tot[2] < d.s_switch ? 1 : 0;
  }
}

// ---------------------------------------------------------------------------
// Adjoint.

__global__ __launch_bounds__(kBlock) void k_adj_seed(DeviceState d, double dt, const float* gq, const float* gqd) {
  const Chunk c = d.vchunks[blockIdx.x];
  const int b = c.body;
  if (threadIdx.x == 0 && static_cast<int>(blockIdx.x) == d.vcb[b]) {
    d.aiter[b] = 0;
    d.aactive[b] = 1;
    d.aa[2 * b] = d.aa[2 * b + 1] = 0;
    d.fscale[b] = 1.0;
  }
  const int i = c.begin + threadIdx.x;
  if (i >= c.end) return;
  for (int k = 0; k < 3; ++k) {
    const double vel = (gqd ? static_cast<double>(gqd[3 * i + k]) : 0.0) / dt;
    d.arhs[3 * i + k] = (gq ? static_cast<double>(gq[3 * i + k]) : 0.0) + vel;
    d.sol[3 * i + k] = 0.0;
    d.xr[3 * i + k] = 0.0;
  }
}

// Jacobian columns of every sample, stored structure-of-arrays [81][samples]
// so that each of the 81 loads of a wavefront is one coalesced line set.
__global__ __launch_bounds__(kBlock) void k_jacobian(DeviceState d, const double* x, const double* targets) {
  const int total = d.E * d.Q;
  const int sample = blockIdx.x * blockDim.x + threadIdx.x;
  if (sample >= total) return;
  const int e = sample / d.Q, s = sample % d.Q;
  const int b = d.ebody[e];
  M3 f = zero3();
  if (d.Q == 1) {
    V3 g[4];
    tet_grads(d.poses, e, g);
    for (int k = 0; k < 4; ++k) add_outer(f, ld3(x, d.idx[4 * e + k]), g[k]);
  } else {
    for (int j = 0; j < 8; ++j) add_outer(f, ld3(x, d.idx[8 * e + j]), hex_grad(d, b, s, j));
  }
  const NhParams nh{d.nh[3 * b], d.nh[3 * b + 1], d.nh[3 * b + 2]};
  double wm = 0.0, target = 1.0;
  V3 m{0, 0, 0};
  if (d.has_muscles) {
    wm = d.mw[e];
    m = ld3(d.mdir, e);
    target = targets ? targets[e] : 1.0;
  }
  // Each column is stored as it is formed (81 values held at once spill).
  const size_t stride = static_cast<size_t>(total);
  double* out = d.jac + sample;
  sample_jacobian_columns(f, d.material, d.jwc[e], d.jwv[e], d.jwn[e], nh, wm, m, target, [&](int c, const M3& dp) {
#pragma unroll
    for (int t = 0; t < 9; ++t) out[(c * 9 + t) * stride] = dp.a[t / 3][t % 3];
  });
}

// ds = sum_c dF_c J_c for one sample.
__device__ __forceinline__ M3 jac_apply(const DeviceState& d, int sample, const M3& df) {
  const size_t stride = static_cast<size_t>(d.E) * d.Q;
  M3 ds = zero3();
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    const double w = df.a[c % 3][c / 3];
#pragma unroll
    for (int t = 0; t < 9; ++t) ds.a[t / 3][t % 3] = ds.a[t / 3][t % 3] + w * d.jac[(c * 9 + t) * stride + sample];
  }
  return ds;
}

// Element rows of the Dirichlet coupling (A - K) mu, for elements with a
// pinned corner (adjoint.py:140-155 with the boundary rows of :474-520).
__device__ __forceinline__ void coupling_tet4_chunk(const DeviceState& d, int ci, int b, const double* mu) {
  const Chunk c = d.echunks[ci];
  const int e = c.begin + threadIdx.x;
  if (e >= c.end) return;
  if (d.pinmask[e] == 0) return;
  if (d.route[2 * b + 1] != 0.0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) st3(d.contrib, 4 * e + k, V3{0, 0, 0});
    return;
  }
  V3 g[4];
  tet_grads(d.poses, e, g);
  const int* nd = d.idx + 4 * e;
  M3 df = zero3();
#pragma unroll
  for (int k = 0; k < 4; ++k) add_outer(df, ld3(mu, nd[k]), g[k]);
  M3 ds = scale(d.wc[e] + d.wv[e] + d.wn[e], df);
  if (d.has_muscles) {
    const V3 m = ld3(d.mdir, e);
    ds = add(ds, scale(d.mw[e], outer(mul(df, m), m)));
  }
  ds = sub(ds, jac_apply(d, e, df));
#pragma unroll
  for (int k = 0; k < 4; ++k) st3(d.contrib, 4 * e + k, mul(ds, g[k]));
}

__device__ __forceinline__ void coupling_hex8_chunk(const DeviceState& d, int ci, int b, const double* mu) {
  const Chunk c = d.echunks[ci];
  const int s = threadIdx.x & 7;
  const int e = c.begin + (threadIdx.x >> 3);
  const bool valid = e < c.end && d.pinmask[e] != 0;
  const bool zero = d.route[2 * b + 1] != 0.0;
  double out[8][3];
#pragma unroll
  for (int j = 0; j < 8; ++j) out[j][0] = out[j][1] = out[j][2] = 0.0;
  if (valid && !zero) {
    const int* nd = d.idx + 8 * e;
    M3 df = zero3();
#pragma unroll
    for (int j = 0; j < 8; ++j) add_outer(df, ld3(mu, nd[j]), hex_grad(d, b, s, j));
    M3 ds = scale(d.wc[e] + d.wv[e] + d.wn[e], df);
    if (d.has_muscles) {
      const V3 m = ld3(d.mdir, e);
      ds = add(ds, scale(d.mw[e], outer(mul(df, m), m)));
    }
    ds = sub(ds, jac_apply(d, 8 * e + s, df));
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const V3 r = mul(ds, hex_grad(d, b, s, j));
      out[j][0] = r.x;
      out[j][1] = r.y;
      out[j][2] = r.z;
    }
  }
  const V3 sum = hex_lane_sum(out, s);
  if (valid) st3(d.contrib, 8 * e + s, sum);
}

// Dirichlet coupling rows over every element chunk (once per backward).
// One workgroup per element chunk that holds a pinned corner (the other
// chunks have no coupling rows).
__global__ __launch_bounds__(kBlock) void k_adj_coupling(DeviceState d) {
  const int ci = d.pin_chunks[blockIdx.x];
  const int b = d.echunks[ci].body;
  if (d.Q == 1)
    coupling_tet4_chunk(d, ci, b, d.sol);
  else
    coupling_hex8_chunk(d, ci, b, d.sol);
}

// The adjoint's element term mu -> sum_e G_e^T J_e (G_e mu) (adjoint.py:140-155)
// is linear in mu and fixed for the whole backward, so it is assembled once
// into 3x3 vertex blocks and every iteration applies it as one sparse
// product.
//
// Row block k of one element: K_e[k][l] = sum_s sum_{j,n} g_sk[j] g_sl[n]
// J_s[(beta, n) -> (alpha, j)], with J stored as jac[(c * 9 + t)] for
// c = beta + 3 n (entry of dF) and t = 3 alpha + j (entry of the stress
// change), the convention of jac_apply. One thread per (element, corner).
__global__ __launch_bounds__(kBlock) void k_element_operator(DeviceState d) {
  const int K = d.K;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= d.E * K) return;
  const int e = tid / K, k = tid % K;
  const int b = d.ebody[e];
  const size_t stride = static_cast<size_t>(d.E) * d.Q;
  double blk[8][9];
#pragma unroll
  for (int l = 0; l < 8; ++l)
#pragma unroll
    for (int t = 0; t < 9; ++t) blk[l][t] = 0.0;
  V3 tg[4];
  if (d.Q == 1) tet_grads(d.poses, e, tg);
  for (int s = 0; s < d.Q; ++s) {
    const size_t sample = static_cast<size_t>(e) * d.Q + s;
    const V3 gk = d.Q == 1 ? tg[k] : hex_grad(d, b, s, k);
    const double gkv[3] = {gk.x, gk.y, gk.z};
    // p[c][alpha] = sum_j g_k[j] J[c][3 alpha + j]
    double p[9][3];
#pragma unroll
    for (int c = 0; c < 9; ++c)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) t += gkv[j] * d.jac[(c * 9 + 3 * a + j) * stride + sample];
        p[c][a] = t;
      }
#pragma unroll
    for (int l = 0; l < 8; ++l) {
      if (l >= K) break;
      const V3 gl = d.Q == 1 ? tg[l & 3] : hex_grad(d, b, s, l);
      const double glv[3] = {gl.x, gl.y, gl.z};
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int be = 0; be < 3; ++be) {
          double t = blk[l][3 * a + be];
#pragma unroll
          for (int n = 0; n < 3; ++n) t += glv[n] * p[be + 3 * n][a];
          blk[l][3 * a + be] = t;
        }
    }
  }
  double* out = d.ke + static_cast<size_t>(tid) * K * 9;
#pragma unroll
  for (int l = 0; l < 8; ++l) {
    if (l >= K) break;
#pragma unroll
    for (int t = 0; t < 9; ++t) out[9 * l + t] = blk[l][t];
  }
}

// The Tet4 case (one sample, four corners): each block K_e[k][l] goes to LDS
// as soon as it is formed (no block array in registers), and the workgroup's
// 256 rows of 36 values, one contiguous range of ke, are then written with
// consecutive lanes on consecutive addresses.
__global__ __launch_bounds__(kBlock) void k_element_operator_tet4(DeviceState d) {
  __shared__ double stage[kBlock * 36];
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int first = blockIdx.x * blockDim.x;
  const int rows = min(kBlock, d.E * 4 - first);
  if (tid < d.E * 4) {
    const int e = tid / 4, k = tid % 4;
    const size_t stride = static_cast<size_t>(d.E);
    V3 tg[4];
    tet_grads(d.poses, e, tg);
    const double gkv[3] = {tg[k].x, tg[k].y, tg[k].z};
    double p[9][3];
#pragma unroll
    for (int c = 0; c < 9; ++c)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) t += gkv[j] * d.jac[(c * 9 + 3 * a + j) * stride + e];
        p[c][a] = t;
      }
    double* out = stage + threadIdx.x * 36;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const double glv[3] = {tg[l].x, tg[l].y, tg[l].z};
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int be = 0; be < 3; ++be) {
          double t = 0.0;
#pragma unroll
          for (int n = 0; n < 3; ++n) t += glv[n] * p[be + 3 * n][a];
          out[9 * l + 3 * a + be] = t;
        }
    }
  }
  __syncthreads();
  double* dst = d.ke + static_cast<size_t>(first) * 36;
  for (int t = threadIdx.x; t < rows * 36; t += kBlock) dst[t] = stage[t];
}

// Rows of the assembled operator: workgroup (chunk, slot), one thread per
// row of the vertex chunk: block (i, j) = sum over the incident elements of
// i, in incidence order, of K_e[k][l] with idx[e][k] = i and idx[e][l] = j
// (the element blocks listed for it by the host, four loads in flight);
// minus A_ij I in the defect mode. Workgroup g runs on XCD g % 8; the slots
// of one chunk are consecutive workgroups of one XCD, so the element rows
// (four blocks, one per slot) a chunk reads are shared in that XCD's L2.
__global__ __launch_bounds__(kBlock) void k_operator_assemble(DeviceState d) {
  const int g = static_cast<int>(blockIdx.x), x = g % 8, r = g / 8;
  const int ci = (r / d.kmaxwidth) * 8 + x, s = r % d.kmaxwidth;
  if (ci >= d.nvc) return;
  const int w = d.kwidth[ci];
  if (s >= w) return;
  const Chunk c = d.vchunks[ci];
  const int i = c.begin + threadIdx.x;
  const size_t base = static_cast<size_t>(d.kbase[ci]);
  const size_t blk = base + static_cast<size_t>(s) * kBlock + threadIdx.x;
  double acc[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) acc[t] = 0.0;
  const int j = i < c.end ? d.kcol[blk] : -1;
  if (j >= 0) {
    const int p1 = d.kmap_off[blk + 1];
    for (int p = d.kmap_off[blk]; p < p1; p += 4) {
      double kb[4][9];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double* src = d.ke + static_cast<size_t>(d.kmap_src[min(p + u, p1 - 1)]) * 9;
#pragma unroll
        for (int t = 0; t < 9; ++t) kb[u][t] = src[t];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (p + u < p1) {
#pragma unroll
          for (int t = 0; t < 9; ++t) acc[t] += kb[u][t];
        }
    }
    if (d.defect) {
      const double a = d.kdiag[blk];
      acc[0] -= a;
      acc[4] -= a;
      acc[8] -= a;
    }
  }
  double* out = d.kval + 9 * base + static_cast<size_t>(s) * 9 * kBlock + threadIdx.x;
#pragma unroll
  for (int t = 0; t < 9; ++t) out[t * kBlock] = acc[t];
}

// work = K mu + rhs (gather with accumulate, adjoint.py:939), or
// rhs + (K - A) mu in the defect mode; one thread per row, a workgroup of
// one wavefront per kRowParts-th of a chunk (so a single body's rows spread
// over that many more CUs). mu is the last iteration's mixed iterate times
// its safe-clip scale (fscale * xr), which this kernel also stores as the
// row's mu (sol).
static_assert(kRowParts * kWave == kBlock, "the parts of a chunk cover its kBlock rows");
__global__ __launch_bounds__(kWave) void k_adj_rhs(DeviceState d) {
  for_active_parts<kRowParts>(d.pref_av, d.vcb, d.B, [&](int ci, int b, int part) {
    const Chunk c = d.vchunks[ci];
    const int row = part * kWave + static_cast<int>(threadIdx.x);
    const int i = c.begin + row;
    if (i >= c.end) return;
    const double sc = d.fscale[b];
    st3(d.sol, i, sc * ld3(d.xr, i));
    const int w = d.kwidth[ci];
    const size_t base = static_cast<size_t>(d.kbase[ci]);
    const int* col = d.kcol + base + row;
    const double* val = d.kval + 9 * base + row;
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    // Slots in batches of 8 column loads, then 4 slots at a time with all
    // their vertex and block loads issued before their products; padding
    // (column -1, zero block) ends the row.
    for (int s0 = 0; s0 < w; s0 += 8) {
      int j[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) j[u] = s0 + u < w ? col[static_cast<size_t>(s0 + u) * kBlock] : -1;
      if (j[0] < 0) break;
#pragma unroll
      for (int h = 0; h < 8; h += 4) {
        V3 m[4];
        double v[4][9];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int slot = min(s0 + h + u, w - 1);
          m[u] = sc * ld3(d.xr, j[h + u] >= 0 ? j[h + u] : i);
          const double* vb = val + static_cast<size_t>(slot) * 9 * kBlock;
#pragma unroll
          for (int t = 0; t < 9; ++t) v[u][t] = vb[t * kBlock];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (j[h + u] >= 0) {
            t0 += v[u][0] * m[u].x + v[u][1] * m[u].y + v[u][2] * m[u].z;
            t1 += v[u][3] * m[u].x + v[u][4] * m[u].y + v[u][5] * m[u].z;
            t2 += v[u][6] * m[u].x + v[u][7] * m[u].y + v[u][8] * m[u].z;
          }
        }
      }
    }
    store_rhs(d, i, V3{t0, t1, t2} + ld3(d.arhs, i));
  });
}

__device__ __forceinline__ V3 scrub(V3 v, bool& scrubbed) {
  if (!isfinite(v.x)) { v.x = 0.0; scrubbed = true; }
  if (!isfinite(v.y)) { v.y = 0.0; scrubbed = true; }
  if (!isfinite(v.z)) { v.z = 0.0; scrubbed = true; }
  return v;
}

// Clip scale (adjoint.py:207-223, _clip_apply) of a vector with squared norm
// `sq`: limit / ||v|| above the 1e6 limit, else 1.
__device__ __forceinline__ double clip_factor(double sq) {
  const double len = sqrt(sq);
  const double limit = 1.0e6;
  return isfinite(len) && len > limit ? limit / len : 1.0;
}

// The raw map output after the S^T product's clip pass (Dirichlet rows and
// non-finite components already zero): scaled to the clip limit (the body's
// column-group norms summed in group order, every workgroup the same; the
// body's first chunk reports the safeguard), then its step from mu and the
// Anderson push with this vertex's Gram terms.
__global__ __launch_bounds__(kBlock) void k_adj_push(DeviceState d) {
  __shared__ double smem[(kGramEntries + 1) * kBlock / kWave];
  __shared__ double res[kGramEntries + 1];
  for_active(d.pref_av, d.vcb, d.B, [&](int ci, int b) {
    const double sc = clip_factor(chunk_sum(d.t_norm, 1, 0, d.tcb[b], d.tcb[b + 1]));
    if (sc != 1.0 && ci == d.vcb[b] && threadIdx.x == 0) atomicOr(d.status, 16);
    const Chunk c = d.vchunks[ci];
    const int i = c.begin + threadIdx.x;
    double acc[kGramEntries + 1];
#pragma unroll
    for (int k = 0; k < kGramEntries + 1; ++k) acc[k] = 0.0;
    if (i < c.end) {
      V3 raw = ld3(d.xr, i);
      if (sc != 1.0) {
        raw = sc * raw;
        st3(d.xr, i, raw);
      }
      const V3 prev = ld3(d.sol, i);
      const V3 step = raw - prev;
      acc[kGramEntries] = dot(step, step);
      if (d.W > 0) {
        double g[kGramEntries];
#pragma unroll
        for (int k = 0; k < kGramEntries; ++k) g[k] = 0.0;
        push_and_gram(d, b, i, prev, raw, d.aiter, g);
#pragma unroll
        for (int k = 0; k < kGramEntries; ++k) acc[k] = g[k];
      }
    }
    block_sum<kGramEntries + 1>(acc, smem, res);
    for (int k = threadIdx.x; k < kGramEntries; k += blockDim.x) d.gram_part[kGramEntries * ci + k] = res[k];
    if (threadIdx.x == 0) d.conv_part[4 * ci] = res[kGramEntries];
  });
}

__global__ __launch_bounds__(kBlock) void k_adj_apply(DeviceState d) {
  __shared__ double smem[kBlock / kWave];
  __shared__ double res[1];
  __shared__ double red[kGramEntries];
  __shared__ double gamma[kMaxWindow];
  __shared__ int meta[3];
  for_active(d.pref_av, d.vcb, d.B, [&](int ci, int b) {
    if (d.W > 0) {
      gram_solve(d, b, d.aiter, ci == d.vcb[b], red, gamma, meta, d.status);
    } else if (threadIdx.x == 0) {
      meta[0] = 0;
    }
    __syncthreads();
    const Chunk c = d.vchunks[ci];
    const int i = c.begin + threadIdx.x;
    double part[1] = {0.0};
    if (i < c.end) {
      V3 v = anderson_value(d, i, ld3(d.xr, i), gamma, meta);
      if (!d.free[i]) v = V3{0, 0, 0};
      bool scrubbed = false;
      v = scrub(v, scrubbed);
      if (scrubbed) atomicOr(d.status, 16);
      st3(d.xr, i, v);
      part[0] = dot(v, v);
    }
    block_sum<1>(part, smem, res);
    if (threadIdx.x == 0) d.clip_part2[ci] = res[0];
  });
}

// Every body's adjoint convergence test (one thread per body, one
// workgroup), then the work-list prefix of the bodies still active. The
// mixed iterate's safe clip (adjoint.py:207-223) is decided here: mu =
// fscale * xr (stored as mu by the next iteration's first kernel or after
// the loop), and its squared norm for the gate is fscale^2 times the
// iterate's (exactly the iterate's when no clip applies).
__global__ __launch_bounds__(kBlock) void k_adj_check(DeviceState d, StepScalars s) {
  extern __shared__ int flag[];  // [B]: still active after this test
  for (int b = threadIdx.x; b < d.B; b += kBlock) {
    flag[b] = 0;
    if (!d.aactive[b]) continue;
    double t0, sq;
    chunk_sum2(d.conv_part, 4, d.clip_part2, 1, d.vcb[b], d.vcb[b + 1], t0, sq);
    const double sc = clip_factor(sq);
    if (sc != 1.0) atomicOr(d.status, 16);
    d.fscale[b] = sc;
    const double t1 = sc == 1.0 ? sq : (sc * sc) * sq;
    const int k = d.aiter[b];
    d.aiter[b] = k + 1;
    if (k > 0 && sqrt(t0) < s.rel_tol * sqrt(t1) + s.abs_tol) {
      d.aactive[b] = 0;
    } else if (k + 1 >= s.max_iterations) {
      d.aactive[b] = 0;
      // Woodbury columns are truncated by design; only the backbone reports.
      if (s.report) atomicOr(d.status, 8);
    } else {
      flag[b] = 1;
    }
  }
  __syncthreads();
  int tot[2];
  body_prefixes<2>(flag, {d.vcb, d.tcb}, {d.pref_av, nullptr}, d.B, tot);
  if (threadIdx.x == 0) {
    d.t_mode[0] = tot[1] < d.t_switch ? 1 : 0;
    d.t_mode[2] = tot[1] < d.s_switch ? 1 : 0;
  }
}

// mu = fscale * xr of every vertex, after an adjoint loop (the bodies that
// converged in its last iteration had theirs only in this form).
__global__ __launch_bounds__(kBlock) void k_adj_finalize(DeviceState d) {
  const Chunk c = d.vchunks[blockIdx.x];
  const int i = c.begin + threadIdx.x;
  if (i < c.end) st3(d.sol, i, d.fscale[c.body] * ld3(d.xr, i));
}

// Routing guard per body (adjoint.py:226-261): one block loops over the body.
__global__ __launch_bounds__(kBlock) void k_adj_route(DeviceState d) {
  const int b = blockIdx.x;
  __shared__ double smem[2 * kBlock / kWave];
  __shared__ double res[2];
  double part[2] = {0.0, 0.0};
  for (int c = d.vcb[b]; c < d.vcb[b + 1]; ++c) {
    const Chunk ch = d.vchunks[c];
    const int i = ch.begin + threadIdx.x;
    if (i < ch.end) {
      const V3 r = ld3(d.arhs, i), m = ld3(d.sol, i);
      part[0] += dot(r, r);
      part[1] += dot(m, m);
    }
  }
  block_sum<2>(part, smem, res);
  if (threadIdx.x != 0) return;
  const double mu_norm = sqrt(res[1]);
  const double limit = 1.0e6;
  const bool failed = mu_norm >= 0.999 * limit;
  double value = 1.0;
  if (failed) {
    value = 0.0;
    atomicOr(d.status, 16);
  } else {
    const double cap = sqrt(res[0]) + 1.0;
    const double injection = d.inertia[b] * mu_norm;
    if (injection > cap) {
      value = cap / injection;
      atomicOr(d.status, 16);
    }
  }
  d.route[2 * b] = value;
  d.route[2 * b + 1] = failed ? 1.0 : 0.0;
}

__global__ __launch_bounds__(kBlock) void k_adj_state_grads(DeviceState d, double dt, const float* gqd_out,
                                                             float* gq, float* gqd, float* gf) {
  const Chunk c = d.vchunks[blockIdx.x];
  const int b = c.body;
  const int i = c.begin + threadIdx.x;
  if (i >= c.end) return;
  const double sc = d.route[2 * b];
  const double inertia = d.inertia[b];
  for (int k = 0; k < 3; ++k) {
    const double mu = d.sol[3 * i + k];
    const double base = -((gqd_out ? static_cast<double>(gqd_out[3 * i + k]) : 0.0) / dt);
    const double inj = sc * inertia * mu;
    gq[3 * i + k] = gq[3 * i + k] + static_cast<float>(base + inj);
    gqd[3 * i + k] = gqd[3 * i + k] + static_cast<float>(dt * inj);
    gf[3 * i + k] = gf[3 * i + k] + static_cast<float>(sc * mu);
  }
}

__global__ __launch_bounds__(kBlock) void k_adj_dirichlet(DeviceState d, double dt, float* gq, float* gqd,
                                                           float* gdir) {
  const Chunk c = d.vchunks[blockIdx.x];
  const int i = c.begin + threadIdx.x;
  if (i >= c.end || d.free[i]) return;
  V3 t{0, 0, 0};
  for (int a = d.inc_off[i]; a < d.inc_off[i + 1]; ++a) t = t + ld3(d.contrib, d.inc_ent[a]);
  const V3 value = ld3(d.arhs, i) - t;
  const double v[3] = {value.x, value.y, value.z};
  for (int k = 0; k < 3; ++k) {
    if (gdir) {
      gdir[3 * i + k] = gdir[3 * i + k] + static_cast<float>(v[k]);
    } else {
      gq[3 * i + k] = gq[3 * i + k] + static_cast<float>(v[k]);
      if (d.integrate) gqd[3 * i + k] = gqd[3 * i + k] + static_cast<float>(dt * v[k]);
    }
  }
}

// PD material-gradient terms of one sample (adjoint.py:556-591): F - R and
// F - U diag(sigma + d) V^T contracted with z.
__device__ __forceinline__ void material_terms(const M3& f, double& first_out, double& second_out, const V3* g,
                                               const V3* z, int K) {
  M3 u, v;
  V3 sg;
  unsigned_svd(f, u, sg, v);
  int singular = 0;
  const V3 off = volume_offset(sg, singular);
  const M3 first = sub(f, mul(u, transpose(v)));
  const M3 second = sub(f, udvt(u, sg + off, v));
  double fv = 0.0, sv = 0.0;
  for (int k = 0; k < K; ++k) {
    fv -= dot(z[k], mul(first, g[k]));
    sv -= dot(z[k], mul(second, g[k]));
  }
  first_out = fv;
  second_out = sv;
}

__global__ __launch_bounds__(kBlock) void k_adj_material_tet4(DeviceState d, const double* x, const double* zs) {
  const Chunk c = d.echunks[blockIdx.x];
  const int e = c.begin + threadIdx.x;
  if (e >= c.end) return;
  const int b = c.body;
  V3 g[4], z[4];
  tet_grads(d.poses, e, g);
  M3 f = zero3();
  for (int k = 0; k < 4; ++k) {
    const int n = d.idx[4 * e + k];
    add_outer(f, ld3(x, n), g[k]);
    z[k] = ld3(zs, n);
  }
  double fv, sv;
  if (d.material == 0) {
    material_terms(f, fv, sv, g, z, 4);
  } else {
    const double mu = d.nh[3 * b], lame = d.nh[3 * b + 1];
    const M3 first = scale(1.0 / (2.0 * mu), neo_hookean_stress(f, mu, lame));
    fv = 0.0;
    for (int k = 0; k < 4; ++k) fv -= dot(z[k], mul(first, g[k]));
    sv = 0.0;
  }
  const double vol = d.svol[b];
  d.gmat_tmp[2 * e] = 2.0 * vol * fv;
  d.gmat_tmp[2 * e + 1] = d.material == 0 ? vol * sv : 0.0;
}

__global__ __launch_bounds__(kBlock) void k_adj_material_hex8(DeviceState d, const double* x, const double* zs) {
  const Chunk c = d.echunks[blockIdx.x];
  const int b = c.body;
  const int s = threadIdx.x & 7;
  const int e = c.begin + (threadIdx.x >> 3);
  const bool valid = e < c.end;
  double fv = 0.0, sv = 0.0;
  if (valid) {
    V3 g[8], z[8];
    M3 f = zero3();
    for (int j = 0; j < 8; ++j) {
      const int n = d.idx[8 * e + j];
      g[j] = hex_grad(d, b, s, j);
      add_outer(f, ld3(x, n), g[j]);
      z[j] = ld3(zs, n);
    }
    if (d.material == 0) {
      material_terms(f, fv, sv, g, z, 8);
    } else {
      const double mu = d.nh[3 * b], lame = d.nh[3 * b + 1];
      const M3 first = scale(1.0 / (2.0 * mu), neo_hookean_stress(f, mu, lame));
      for (int j = 0; j < 8; ++j) fv -= dot(z[j], mul(first, g[j]));
    }
  }
  for (int off = 1; off < 8; off <<= 1) {
    fv += __shfl_xor(fv, off, kWave);
    sv += __shfl_xor(sv, off, kWave);
  }
  if (valid && s == 0) {
    const double vol = d.svol[b];
    d.gmat_tmp[2 * e] = 2.0 * vol * fv;
    d.gmat_tmp[2 * e + 1] = d.material == 0 ? vol * sv : 0.0;
  }
}

// Activation gradients of Tet4 fibres (adjoint.py:595-620): d/du of the
// contraction u with target length a = 1 - u is
// -sum_k z_k . (w_m / |F m|) (m . g_k) F m, zero for a disabled or collapsed
// fibre.
__global__ __launch_bounds__(kBlock) void k_adj_activation_tet4(DeviceState d, const double* x, const double* zs) {
  const Chunk c = d.echunks[blockIdx.x];
  const int e = c.begin + threadIdx.x;
  if (e >= c.end) return;
  const double wm = d.mw[e];
  double value = 0.0;
  if (wm != 0.0) {
    V3 g[4];
    tet_grads(d.poses, e, g);
    M3 f = zero3();
    for (int k = 0; k < 4; ++k) add_outer(f, ld3(x, d.idx[4 * e + k]), g[k]);
    const V3 m = ld3(d.mdir, e);
    const V3 fibre = mul(f, m);
    const double len = sqrt(dot(fibre, fibre));
    if (len > 2.220446049250313e-16) {
      for (int k = 0; k < 4; ++k) {
        const V3 z = ld3(zs, d.idx[4 * e + k]);
        value += dot(z, (wm / len * dot(m, g[k])) * fibre);
      }
      value = -value;
    }
  }
  d.gact_tmp[e] = value;
}

// Aggregate guard per body (adjoint.py:622-644), parallel over the body's
// elements: valid[2b] gates the material gradients, valid[2b+1] the
// activation gradients; a non-finite aggregate of either sets status 16.
__global__ __launch_bounds__(kBlock) void k_adj_parameter_guard(DeviceState d, int materials, int activations) {
  const int b = blockIdx.x;
  __shared__ double smem[3 * kBlock / kWave];
  __shared__ double res[3];
  double part[3] = {0.0, 0.0, 0.0};
  for (int e = d.eoff[b] + threadIdx.x; e < d.eoff[b + 1]; e += kBlock) {
    if (materials) {
      part[0] += d.gmat_tmp[2 * e];
      part[1] += d.gmat_tmp[2 * e + 1];
    }
    // Any non-finite value makes the sum non-finite (a NaN stays NaN, and
    // infinities of both signs give NaN), which is the per-element test.
    if (activations) part[2] += isfinite(d.gact_tmp[e]) ? 0.0 : NAN;
  }
  block_sum<3>(part, smem, res);
  if (threadIdx.x != 0) return;
  const bool zero_energy = materials && d.material == 0 && d.nh[3 * b + 1] == 0.0;
  const bool material_ok = !zero_energy && isfinite(res[0]) && isfinite(res[1]);
  const bool activation_ok = isfinite(res[2]);
  if (!material_ok || !activation_ok) atomicOr(d.status, 16);
  const bool live = d.route[2 * b + 1] == 0.0;
  d.valid[2 * b] = (material_ok && live) ? 1 : 0;
  d.valid[2 * b + 1] = (activation_ok && live) ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void k_adj_parameter_add(DeviceState d, double* gmat, double* gact) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d.E) return;
  const int b = d.ebody[e];
  if (gmat && d.valid[2 * b]) {
    gmat[2 * e] += d.gmat_tmp[2 * e];
    gmat[2 * e + 1] += d.gmat_tmp[2 * e + 1];
  }
  if (gact && d.valid[2 * b + 1]) gact[e] += d.gact_tmp[e];
}

// det(F) at the eight Gauss samples of every Hex8 cell from float32
// positions (deformation_determinants, hex.py:190-201); one lane per sample.
__global__ __launch_bounds__(kBlock) void k_hex_determinants(DeviceState d, const float* q, double* out) {
  const int sample = blockIdx.x * blockDim.x + threadIdx.x;
  if (sample >= d.E * 8) return;
  const int e = sample >> 3, s = sample & 7;
  const int b = d.ebody[e];
  M3 f = zero3();
  for (int j = 0; j < 8; ++j) {
    const int n = d.idx[8 * e + j];
    const V3 x{static_cast<double>(q[3 * n]), static_cast<double>(q[3 * n + 1]), static_cast<double>(q[3 * n + 2])};
    add_outer(f, x, hex_grad(d, b, s, j));
  }
  out[sample] = det3(f);
}

// ---------------------------------------------------------------------------
// Pointwise evaluation kernels.

__global__ void k_eval_projections(int count, const double* fs, const double* dirs, double mu, double lame,
                                   double penalty, double target, double* sigma, double* recon, double* rotation,
                                   double* volume, double* nhp, double* fibre) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  M3 f;
  for (int t = 0; t < 9; ++t) f.a[t / 3][t % 3] = fs[9 * i + t];
  M3 u, v;
  V3 s;
  unsigned_svd(f, u, s, v);
  st3(sigma, i, s);
  const M3 rc = udvt(u, s, v), rt = mul(u, transpose(v));
  int singular = 0;
  const V3 off = volume_offset(s, singular);
  const M3 vol = udvt(u, s + off, v);
  const M3 nhm = udvt(u, nh_prox_stretches(s, mu, lame, penalty), v);
  for (int t = 0; t < 9; ++t) {
    recon[9 * i + t] = rc.a[t / 3][t % 3];
    rotation[9 * i + t] = rt.a[t / 3][t % 3];
    volume[9 * i + t] = vol.a[t / 3][t % 3];
    nhp[9 * i + t] = nhm.a[t / 3][t % 3];
  }
  st3(fibre, i, muscle_projection(f, ld3(dirs, i), target));
}

__global__ void k_eval_jvps(int count, const double* fs, const double* dfs, const double* dirs, double mu,
                            double lame, double target, double* rotation, double* volume, double* nhp,
                            double* fibre) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  M3 f, df;
  for (int t = 0; t < 9; ++t) {
    f.a[t / 3][t % 3] = fs[9 * i + t];
    df.a[t / 3][t % 3] = dfs[9 * i + t];
  }
  M3 u, v;
  V3 s;
  unsigned_svd(f, u, s, v);
  const M3 r = corotated_jvp(u, s, v, df);
  int singular = 0;
  const V3 off = volume_offset(s, singular);
  const M3 vo = volume_jvp(u, s, v, off, df);
  const double pen = 2.0 * mu;
  const V3 st = nh_prox_stretches(s, mu, lame, pen);
  const M3 nj = nh_prox_jvp(u, s, v, st, df, mu, lame, pen);
  for (int t = 0; t < 9; ++t) {
    rotation[9 * i + t] = r.a[t / 3][t % 3];
    volume[9 * i + t] = vo.a[t / 3][t % 3];
    nhp[9 * i + t] = nj.a[t / 3][t % 3];
  }
  st3(fibre, i, muscle_jvp(f, ld3(dirs, i), target, df));
}

size_t prefix_lds(const DeviceState& d) { return sizeof(int) * static_cast<size_t>(d.B + 1); }

// The pair x = S^T (S rhs4) (+ base) over the items of the active bodies:
// the S product writes the 4-wide postorder y4, the S^T product the FP64
// [N*3] out in original order.
// The S product, then S^T from whichever list *mode selects: the whole
// groups, or the sliced list and its combine (the other launches return at
// once).
// S^T from whichever list *mode selects: one launch over both lists' grids
// (the workgroups of the other list, and those beyond the chosen list's
// grid, return at once), then the sliced list's combine (which returns at
// once unless the sliced list ran).
template <typename Real, typename Acc, typename In>
void cols_launch(const DeviceState& d, double* out, const double* base, const int* active, AdjointEpilogue adj,
                 const int* mode, hipStream_t stream) {
  const Real* val = static_cast<const Real*>(d.fval);
  In* y4 = static_cast<In*>(d.y4);
  Acc* part_buf = static_cast<Acc*>(d.t_part);
  const ColLists whole{d.t_items, d.t_woff, d.t_list, d.t_nitems ? d.g_t : 0};
  const ColLists sliced{d.ts_items, d.ts_woff, d.ts_list, d.ts_nitems ? d.g_ts : 0};
  const int grid = std::max(whole.grid, sliced.grid);
  if (grid)
    hipLaunchKernelGGL((k_factor_cols<Real, Acc, In>), dim3(grid), dim3(kColWaves * kWave), 0, stream, whole, sliced,
                       d.t_entries, val, static_cast<const In*>(y4), out, d.gorig, base, active, adj, mode, part_buf);
  if (d.tc_nitems)
    hipLaunchKernelGGL(k_factor_combine<Acc>, dim3(d.g_tc), dim3(kBlock), 0, stream, d.tc_items, d.tc_nitems,
                       static_cast<const Acc*>(part_buf), out, d.gorig, base, active, adj, mode);
}

// The S product from whichever list *mode selects, the same way, then the
// split list's combine.
template <typename Real, typename Acc, typename In>
void rows_launch(const DeviceState& d, const int* active, const int* mode, hipStream_t stream) {
  const Real* val = static_cast<const Real*>(d.fval);
  In* y4 = static_cast<In*>(d.y4);
  Acc* part_buf = static_cast<Acc*>(d.s_part);
  const RowLists whole{d.s_items, d.s_woff, d.s_list, d.s_nitems ? d.g_s : 0};
  const RowLists split{d.ss_items, d.ss_woff, d.ss_list, d.ss_nitems ? d.g_ss : 0};
  const int grid = std::max(whole.grid, split.grid);
  if (grid)
    hipLaunchKernelGGL((k_factor_rows<Real, Acc, In>), dim3(grid), dim3(kBlock), 0, stream, whole, split, val,
                       static_cast<const In*>(d.rhs4), y4, active, mode, part_buf);
  if (d.sc_nitems)
    hipLaunchKernelGGL((k_rows_combine<Acc, In>), dim3(std::min(kCombineGrid, (d.sc_nitems * 32 + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, stream, d.sc_items, d.sc_nitems, static_cast<const Acc*>(part_buf), y4, active,
                       mode);
}

template <typename Real, typename Acc, typename In>
void spmv_pair(const DeviceState& d, double* out, const double* base, const int* active, AdjointEpilogue adj,
               const int* t_mode, const int* s_mode, hipStream_t stream) {
  rows_launch<Real, Acc, In>(d, active, s_mode, stream);
  cols_launch<Real, Acc, In>(d, out, base, active, adj, t_mode, stream);
}

// Dispatch on the stored precision; `all` selects the lists chosen for every
// body at once instead of the active bodies' (the loops' prefix kernels set
// those).
void spmv_dispatch(const DeviceState& d, double* out, const double* base, const int* active, hipStream_t stream,
                   AdjointEpilogue adj = AdjointEpilogue{}, bool all = false) {
  if (!d.t_mode) throw std::logic_error("factor products launched before their list choices were allocated");
  const int* t_mode = d.t_mode + (all ? 1 : 0);
  const int* s_mode = d.t_mode + (all ? 3 : 2);
  if (d.f32)
    spmv_pair<float, float, float>(d, out, base, active, adj, t_mode, s_mode, stream);
  else
    spmv_pair<double, double, double>(d, out, base, active, adj, t_mode, s_mode, stream);
}

// Iteration products: x_raw = S^T (S rhs) (+ base, the defect-correction form).
void spmv_iteration(const DeviceState& d, const double* base, const int* active, hipStream_t stream) {
  spmv_dispatch(d, d.xr, base, active, stream);
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers.

hipError_t product_capacity(bool f32, int* rows_blocks, int* cols_blocks) {
  int dev = 0, cus = 0, nr = 0, nc = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess)
    e = f32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nr, k_factor_rows<float, float, float>, kBlock, 0)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nr, k_factor_rows<double, double, double>, kBlock, 0);
  if (e == hipSuccess)
    e = f32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nc, k_factor_cols<float, float, float>, kColWaves * kWave, 0)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nc, k_factor_cols<double, double, double>,
                                                           kColWaves * kWave, 0);
  *rows_blocks = std::max(1, nr) * std::max(1, cus);
  *cols_blocks = std::max(1, nc) * std::max(1, cus);
  return e;
}

void launch_begin_step(const DeviceState& d, const StepScalars& s, const float* q, const float* qd, const float* f,
                       const float* dirichlet, hipStream_t stream) {
  hipLaunchKernelGGL(k_begin_step, dim3(d.nvc), dim3(kBlock), 0, stream, d, s, q, qd, f, dirichlet);
  launch_active_prefix(d, false, stream);
}

void launch_active_prefix(const DeviceState& d, bool adjoint, hipStream_t stream) {
  const size_t lds = prefix_lds(d);
  if (adjoint)
    hipLaunchKernelGGL(k_active_prefix, dim3(1), dim3(kBlock), lds, stream, d.aactive, d.vcb, d.pref_av, nullptr,
                       nullptr, d.B, d.tcb, d.t_switch, d.s_switch, d.t_mode);
  else
    hipLaunchKernelGGL(k_active_prefix, dim3(1), dim3(kBlock), lds, stream, d.active, d.vcb, d.pref_fv, d.ecb,
                       d.pref_fe, d.B, d.tcb, d.t_switch, d.s_switch, d.t_mode);
}

void launch_forward_iteration(const DeviceState& d, const StepScalars& s, const double* targets,
                              const FrictionDev* fc, hipStream_t stream) {
  const size_t lds = prefix_lds(d);
  hipLaunchKernelGGL(k_local, dim3(d.g_elem), dim3(kBlock), stage_offset(d.B) + 24 * static_cast<size_t>(d.lv_max),
                     stream, d, targets);
  hipLaunchKernelGGL(k_gather, dim3(d.g_vert), dim3(kBlock), lds, stream, d);
  spmv_iteration(d, d.defect ? d.x : nullptr, d.active, stream);
  if (fc) launch_contact_iteration(*fc, d.x, d.xr, d.q0, d.active, stream);
  if (d.W > 0) hipLaunchKernelGGL(k_push_gram, dim3(d.g_vert), dim3(kBlock), lds, stream, d);
  hipLaunchKernelGGL(k_apply_conv, dim3(d.g_vert), dim3(kBlock), lds, stream, d);
  hipLaunchKernelGGL(k_conv_check, dim3(1), dim3(kBlock), lds, stream, d, s);
}

void launch_any_active(const DeviceState& d, const int* active, const int* iters, hipStream_t stream) {
  hipLaunchKernelGGL(k_any_active, dim3(1), dim3(kBlock), 0, stream, active, iters, d.B, d.any_active);
}

void launch_write_state(const DeviceState& d, const StepScalars& s, float* q, float* qd, double* checkpoint,
                        hipStream_t stream) {
  hipLaunchKernelGGL(k_write_state, dim3(d.nvc), dim3(kBlock), 0, stream, d, s.dt, q, qd, checkpoint);
}

void launch_adjoint_setup(const DeviceState& d, const StepScalars& s, const double* x, const double* targets,
                          const float* gq_out, const float* gqd_out, hipStream_t stream) {
  hipLaunchKernelGGL(k_adj_seed, dim3(d.nvc), dim3(kBlock), 0, stream, d, s.dt, gq_out, gqd_out);
  launch_active_prefix(d, true, stream);
  const int total = d.E * d.Q;
  hipLaunchKernelGGL(k_jacobian, dim3((total + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, d, x, targets);
  const int rows = d.E * d.K;
  if (d.Q == 1)
    hipLaunchKernelGGL(k_element_operator_tet4, dim3((rows + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, d);
  else
    hipLaunchKernelGGL(k_element_operator, dim3((rows + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, d);
  const long long groups = 8LL * ((d.nvc + 7) / 8) * d.kmaxwidth;
  hipLaunchKernelGGL(k_operator_assemble, dim3(static_cast<unsigned>(groups)), dim3(kBlock), 0, stream, d);
}

void launch_adjoint_iteration(const DeviceState& d, const StepScalars& s, hipStream_t stream) {
  const size_t lds = prefix_lds(d);
  hipLaunchKernelGGL(k_adj_rhs, dim3(d.g_rows), dim3(kWave), lds, stream, d);
  spmv_dispatch(d, d.xr, d.defect ? d.sol : nullptr, d.aactive, stream, AdjointEpilogue{d.free, d.t_norm, d.status});
  hipLaunchKernelGGL(k_adj_push, dim3(d.g_vert), dim3(kBlock), lds, stream, d);
  hipLaunchKernelGGL(k_adj_apply, dim3(d.g_vert), dim3(kBlock), lds, stream, d);
  hipLaunchKernelGGL(k_adj_check, dim3(1), dim3(kBlock), lds, stream, d, s);
}

void launch_adjoint_finalize(const DeviceState& d, hipStream_t stream) {
  hipLaunchKernelGGL(k_adj_finalize, dim3(d.nvc), dim3(kBlock), 0, stream, d);
}

void launch_adjoint_route(const DeviceState& d, hipStream_t stream) {
  hipLaunchKernelGGL(k_adj_route, dim3(d.B), dim3(kBlock), 0, stream, d);
}

void launch_adjoint_outputs(const DeviceState& d, const StepScalars& s, const double* x, const float* gqd_out,
                            float* gq_in, float* gqd_in, float* gf_in, double* gmat, double* gact, float* gdir,
                            const double* z, bool any_pinned, hipStream_t stream) {
  if (!z) z = d.sol;
  hipLaunchKernelGGL(k_adj_state_grads, dim3(d.nvc), dim3(kBlock), 0, stream, d, s.dt, gqd_out, gq_in, gqd_in,
                     gf_in);
  if (any_pinned) {
    if (d.npin_chunks) hipLaunchKernelGGL(k_adj_coupling, dim3(d.npin_chunks), dim3(kBlock), 0, stream, d);
    hipLaunchKernelGGL(k_adj_dirichlet, dim3(d.nvc), dim3(kBlock), 0, stream, d, s.dt, gq_in, gqd_in, gdir);
  }
  if (gmat) {
    if (d.Q == 1)
      hipLaunchKernelGGL(k_adj_material_tet4, dim3(d.nec), dim3(kBlock), 0, stream, d, x, z);
    else
      hipLaunchKernelGGL(k_adj_material_hex8, dim3(d.nec), dim3(kBlock), 0, stream, d, x, z);
  }
  if (gact) hipLaunchKernelGGL(k_adj_activation_tet4, dim3(d.nec), dim3(kBlock), 0, stream, d, x, z);
  if (gmat || gact) {
    hipLaunchKernelGGL(k_adj_parameter_guard, dim3(d.B), dim3(kBlock), 0, stream, d, gmat != nullptr ? 1 : 0,
                       gact != nullptr ? 1 : 0);
    hipLaunchKernelGGL(k_adj_parameter_add, dim3((d.E + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, d, gmat,
                       gact);
  }
}

void launch_cols(const DeviceState& d, const int* active, const int* mode, hipStream_t stream) {
  if (!mode) throw std::logic_error("S^T launched without a list choice");
  if (d.f32)
    cols_launch<float, float, float>(d, d.xr, nullptr, active, AdjointEpilogue{}, mode, stream);
  else
    cols_launch<double, double, double>(d, d.xr, nullptr, active, AdjointEpilogue{}, mode, stream);
}

void launch_rows(const DeviceState& d, const int* active, const int* mode, hipStream_t stream) {
  if (!mode) throw std::logic_error("S launched without a list choice");
  if (d.f32)
    rows_launch<float, float, float>(d, active, mode, stream);
  else
    rows_launch<double, double, double>(d, active, mode, stream);
}

void launch_apply_inverse(const DeviceState& d, const double* rhs, double* out, hipStream_t stream) {
  const int grid = (d.N + kBlock - 1) / kBlock;
  if (d.f32)
    hipLaunchKernelGGL(k_pack4<float>, dim3(grid), dim3(kBlock), 0, stream, d.N, d.gpost, rhs,
                       static_cast<float*>(d.rhs4));
  else
    hipLaunchKernelGGL(k_pack4<double>, dim3(grid), dim3(kBlock), 0, stream, d.N, d.gpost, rhs,
                       static_cast<double*>(d.rhs4));
  spmv_dispatch(d, out, nullptr, d.ones, stream, AdjointEpilogue{}, true);
}

void launch_refined_inverse(const DeviceState& d, const double* rhs, double* out, double* r, double* corr, int sweeps,
                            hipStream_t stream) {
  launch_apply_inverse(d, rhs, out, stream);
  if (!d.defect) return;
  const int grid = (d.N + kBlock - 1) / kBlock;
  for (int k = 0; k < sweeps; ++k) {
    hipLaunchKernelGGL(k_residual_a, dim3(grid), dim3(kBlock), 0, stream, d, rhs, out, r);
    launch_apply_inverse(d, r, corr, stream);
    hipLaunchKernelGGL(k_add3, dim3(grid), dim3(kBlock), 0, stream, d.N, corr, out);
  }
}

void launch_hex_determinants(const DeviceState& d, const float* q, double* out, hipStream_t stream) {
  if (d.E == 0) return;
  hipLaunchKernelGGL(k_hex_determinants, dim3((d.E * 8 + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, d, q, out);
}

void launch_eval_projections(int count, const double* f, const double* dirs, double mu, double lame, double penalty,
                             double target, double* sigma, double* recon, double* rotation, double* volume,
                             double* nh, double* fibre, hipStream_t stream) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_eval_projections, dim3((count + 63) / 64), dim3(64), 0, stream, count, f, dirs, mu, lame,
                     penalty, target, sigma, recon, rotation, volume, nh, fibre);
}

void launch_eval_jvps(int count, const double* f, const double* df, const double* dirs, double mu, double lame,
                      double target, double* rotation, double* volume, double* nh, double* fibre,
                      hipStream_t stream) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_eval_jvps, dim3((count + 63) / 64), dim3(64), 0, stream, count, f, df, dirs, mu, lame,
                     target, rotation, volume, nh, fibre);
}

}  // namespace dphd
