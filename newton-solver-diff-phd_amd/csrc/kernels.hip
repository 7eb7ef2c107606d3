// gfx950 kernels of the DiffPhD projective-dynamics step and its adjoint.
//
// Reference kernels (paths under src/newton_solver_diff_phd/_src/solvers/diff_phd/):
//   k_begin_step        kernels.py:155-207 (begin_step, explicit_boundary)
//   k_local_tet4        kernels.py:222-272 (local_step)
//   k_local_hex8        hex.py:122-169 (one lane per Gauss sample, 8 lanes
//                       per cell summed by a lane butterfly)
//   k_gather            kernels.py:75-91 + global_rhs :275-284 + the
//                       projection half of the convergence gate :354-401
//   k_spmv              sparse.py:209-227 (CSR-vector, rows binned by length)
//   k_push_gram         anderson.py:46-94 (push + Gram partials) with
//                       impose_boundary kernels.py:287-291
//   k_gram_solve        anderson.py:150-215 (reduce + pivoted LDL^T solve)
//   k_apply_conv        anderson.py:218-243 + convergence_tiled :354-401
//   k_conv_check        kernels.py:322-347
//   k_write_state       kernels.py:404-414
//   adjoint kernels     adjoint.py:70-321, 474-659
// Every reduction is a fixed-shape butterfly / chunk-ordered sum: results
// are bitwise repeatable run to run. No float atomics.
#include "kernels.h"
#include "pd_math.h"

namespace dphd {
namespace {

constexpr int kWave = 64;

__device__ __forceinline__ V3 ld3(const double* p, int i) { return V3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ void st3(double* p, int i, V3 v) {
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Sum NV values over the 256-thread block. Result in out_smem[0..NV) after
// the final barrier (every thread may read it).
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* smem, double* result) {
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) smem[wave * NV + k] = v[k];
  __syncthreads();
  for (int k = threadIdx.x; k < NV; k += blockDim.x) {
    double t = smem[k];
    for (int w = 1; w < kBlock / kWave; ++w) t += smem[w * NV + k];
    result[k] = t;
  }
  __syncthreads();
}

__device__ __forceinline__ M3 tet_grads(const float* poses, int e, V3 (&g)[4]) {
  const float* p = poses + 9 * e;
  double r[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) r[t] = static_cast<double>(p[t]);
  g[0] = V3{-((r[0] + r[3]) + r[6]), -((r[1] + r[4]) + r[7]), -((r[2] + r[5]) + r[8])};
  g[1] = V3{r[0], r[1], r[2]};
  g[2] = V3{r[3], r[4], r[5]};
  g[3] = V3{r[6], r[7], r[8]};
  return M3{};
}

__device__ __forceinline__ void add_outer(M3& f, V3 x, V3 g) {
  const double xx[3] = {x.x, x.y, x.z}, gg[3] = {g.x, g.y, g.z};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) f.a[i][j] = f.a[i][j] + xx[i] * gg[j];
}

// ---------------------------------------------------------------------------
// Step begin / end.

__global__ __launch_bounds__(kBlock) void k_begin_step(DeviceState d, StepScalars s, const float* q,
                                                        const float* qd, const float* f, const float* dir) {
  const Chunk c = d.vchunks[blockIdx.x];
  const int b = c.body;
  if (threadIdx.x == 0 && static_cast<int>(blockIdx.x) == d.vcb[b]) {
    d.iter[b] = 0;
    d.active[b] = 1;
    d.aa[4 * b] = d.aa[4 * b + 1] = d.aa[4 * b + 2] = d.aa[4 * b + 3] = 0;
    d.accept[b] = 0;
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

__global__ void k_any_active(const int* active, int B, int* out) {
  __shared__ int flag;
  if (threadIdx.x == 0) flag = 0;
  __syncthreads();
  int mine = 0;
  for (int b = threadIdx.x; b < B; b += blockDim.x) mine |= active[b];
  if (mine) flag = 1;  // benign same-value store
  __syncthreads();
  if (threadIdx.x == 0) *out = flag;
}

// ---------------------------------------------------------------------------
// Local step.

__global__ __launch_bounds__(kBlock) void k_local_tet4(DeviceState d, const double* targets) {
  const Chunk c = d.echunks[blockIdx.x];
  if (!d.active[c.body]) return;
  const int e = c.begin + threadIdx.x;
  if (e >= c.end) return;
  const int b = c.body;
  V3 g[4];
  tet_grads(d.poses, e, g);
  const int* nd = d.idx + 4 * e;
  M3 f = zero3();
#pragma unroll
  for (int k = 0; k < 4; ++k) add_outer(f, ld3(d.x, nd[k]), g[k]);
  M3 fd = zero3();
  const unsigned mask = d.pinmask[e];
  if (mask) {
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

// 32 cells per block, 8 lanes per cell (one per Gauss sample).
__global__ __launch_bounds__(kBlock) void k_local_hex8(DeviceState d, const double* targets) {
  const Chunk c = d.echunks[blockIdx.x];
  if (!d.active[c.body]) return;
  const int b = c.body;
  const int s = threadIdx.x & 7;
  const int e = c.begin + (threadIdx.x >> 3);
  const bool valid = e < c.end;
  double out[8][3];
#pragma unroll
  for (int j = 0; j < 8; ++j) out[j][0] = out[j][1] = out[j][2] = 0.0;
  if (valid) {
    const double* G = d.hexg + 192 * b + 24 * s;
    const int* nd = d.idx + 8 * e;
    M3 f = zero3();
#pragma unroll
    for (int j = 0; j < 8; ++j) add_outer(f, ld3(d.x, nd[j]), V3{G[3 * j], G[3 * j + 1], G[3 * j + 2]});
    M3 fd = zero3();
    const unsigned mask = d.pinmask[e];
    if (mask) {
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (mask & (1u << j)) add_outer(fd, ld3(d.xd, nd[j]), V3{G[3 * j], G[3 * j + 1], G[3 * j + 2]});
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
      const V3 r = mul(stress, V3{G[3 * j], G[3 * j + 1], G[3 * j + 2]});
      out[j][0] = r.x;
      out[j][1] = r.y;
      out[j][2] = r.z;
    }
  }
#pragma unroll
  for (int off = 1; off < 8; off <<= 1)
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) out[j][k] += __shfl_xor(out[j][k], off, kWave);
  if (valid) {
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (j == s) st3(d.contrib, 8 * e + j, V3{out[j][0], out[j][1], out[j][2]});
  }
}

// ---------------------------------------------------------------------------
// Gather + right-hand side (+ projection-change partials).

__global__ __launch_bounds__(kBlock) void k_gather(DeviceState d) {
  const Chunk c = d.vchunks[blockIdx.x];
  if (!d.active[c.body]) return;
  __shared__ double smem[2 * kBlock / kWave];
  __shared__ double res[2];
  const int i = c.begin + threadIdx.x;
  double part[2] = {0.0, 0.0};
  if (i < c.end) {
    V3 t{0, 0, 0};
    for (int a = d.inc_off[i]; a < d.inc_off[i + 1]; ++a) t = t + ld3(d.contrib, d.inc_ent[a]);
    const V3 old = ld3(d.p, i);
    V3 r = d.inertia[c.body] * ld3(d.qt, i) + t;
    if (d.defect) {
      V3 ax{0, 0, 0};
      for (int q = d.a_off[i]; q < d.a_off[i + 1]; ++q) ax = ax + d.a_val[q] * ld3(d.x, d.a_col[q]);
      r = r - ax;
    }
    st3(d.p, i, t);
    st3(d.rhs, i, r);
    const V3 dp = t - old;
    part[0] = dot(dp, dp);
    part[1] = dot(old, old);
  }
  block_sum<2>(part, smem, res);
  if (threadIdx.x < 2) d.conv_part[4 * blockIdx.x + 2 + threadIdx.x] = res[threadIdx.x];
}

// ---------------------------------------------------------------------------
// CSR products y = M x for every active body's rows. PRED = false is the
// plain product of every row (SparseInverse.apply), PRED = true skips the
// rows of bodies whose fixed point has converged.

template <typename Real, bool PRED>
__global__ __launch_bounds__(kBlock) void k_spmv(const SpmvTask* tasks, const int* rows, const int* off,
                                                  const int* col, const Real* val, const double* in, double* out,
                                                  const double* base, const int* vbody, const int* active) {
  const SpmvTask t = tasks[blockIdx.x];
  if (t.lanes == kBlock) {
    const int row = rows[t.first];
    if (PRED && !active[vbody[row]]) return;
    __shared__ double smem[3 * kBlock / kWave];
    __shared__ double res[3];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int q = off[row] + threadIdx.x; q < off[row + 1]; q += kBlock) {
      const double v = static_cast<double>(val[q]);
      const double* x = in + 3 * col[q];
      acc[0] += v * x[0];
      acc[1] += v * x[1];
      acc[2] += v * x[2];
    }
    block_sum<3>(acc, smem, res);
    if (threadIdx.x == 0) {
      V3 r{res[0], res[1], res[2]};
      if (base) r = r + ld3(base, row);
      st3(out, row, r);
    }
    return;
  }
  const int L = t.lanes;
  const int group = threadIdx.x / L, lane = threadIdx.x % L;
  int row = -1;
  if (group < t.count) {
    row = rows[t.first + group];
    if (PRED && !active[vbody[row]]) row = -1;
  }
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  if (row >= 0) {
    for (int q = off[row] + lane; q < off[row + 1]; q += L) {
      const double v = static_cast<double>(val[q]);
      const double* x = in + 3 * col[q];
      a0 += v * x[0];
      a1 += v * x[1];
      a2 += v * x[2];
    }
  }
  for (int o = L / 2; o > 0; o >>= 1) {
    a0 += __shfl_xor(a0, o, kWave);
    a1 += __shfl_xor(a1, o, kWave);
    a2 += __shfl_xor(a2, o, kWave);
  }
  if (row >= 0 && lane == 0) {
    V3 r{a0, a1, a2};
    if (base) r = r + ld3(base, row);
    st3(out, row, r);
  }
}

// ---------------------------------------------------------------------------
// Anderson mixing.

__device__ __forceinline__ int gram_index(int a, int b) {  // a <= b < 8
  return a * kMaxWindow - (a * (a - 1)) / 2 + (b - a);
}

// Push (x_prev, g) and emit Gram partials of the body's history for this
// chunk. `prev` is the previous iterate, `raw` the map output; when `zero_pinned`
// Dirichlet entries of raw are zeroed (adjoint) instead of set to xd (forward).
__device__ __forceinline__ void push_and_gram(const DeviceState& d, const Chunk& c, int i, V3 prev, V3 raw,
                                              double (&acc)[kGramEntries]) {
  const int b = c.body;
  const int count = d.aa[4 * b], start = d.aa[4 * b + 1];
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
    acc[kMaxWindow * (kMaxWindow + 1) / 2 + a] += dot(dg[a], g);
  }
}

__global__ __launch_bounds__(kBlock) void k_push_gram(DeviceState d) {
  const Chunk c = d.vchunks[blockIdx.x];
  if (!d.active[c.body]) return;
  __shared__ double smem[kGramEntries * kBlock / kWave];
  __shared__ double res[kGramEntries];
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
    push_and_gram(d, c, i, ld3(d.x, i), raw, acc);
  }
  block_sum<kGramEntries>(acc, smem, res);
  for (int k = threadIdx.x; k < kGramEntries; k += blockDim.x) d.gram_part[kGramEntries * blockIdx.x + k] = res[k];
}

// One block per body: chunk-ordered reduction and the reference _solve.
__global__ void k_gram_solve(DeviceState d, const int* active) {
  const int b = blockIdx.x;
  if (!active[b]) return;
  __shared__ double red[kGramEntries];
  const int c0 = d.vcb[b], c1 = d.vcb[b + 1];
  for (int k = threadIdx.x; k < kGramEntries; k += blockDim.x) {
    double t = 0.0;
    for (int c = c0; c < c1; ++c) t += d.gram_part[kGramEntries * c + k];
    red[k] = t;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const int W = d.W, cap = W + 1;
  int count = d.aa[4 * b], start = d.aa[4 * b + 1];
  d.aa[4 * b + 2] = count;
  d.aa[4 * b + 3] = start;
  int accept = 0;
  if (count >= 1) {
    double gram[kMaxWindow * kMaxWindow], rhs[kMaxWindow], coef[kMaxWindow];
    for (int a = 0; a < count; ++a) {
      rhs[a] = red[kMaxWindow * (kMaxWindow + 1) / 2 + a];
      for (int bb = a; bb < count; ++bb) gram[a * kMaxWindow + bb] = red[gram_index(a, bb)];
    }
    double trace = 0.0;
    for (int a = 0; a < count; ++a) {
      trace += gram[a * kMaxWindow + a];
      for (int bb = a + 1; bb < count; ++bb) gram[bb * kMaxWindow + a] = gram[a * kMaxWindow + bb];
    }
    const double reg = fmax(1.0e-6 * trace / static_cast<double>(count), 1.0e-12);
    for (int a = 0; a < count; ++a) gram[a * kMaxWindow + a] = gram[a * kMaxWindow + a] + reg;
    ldlt_solve<kMaxWindow>(gram, kMaxWindow, rhs, count, coef);
    double norm = 0.0;
    for (int a = 0; a < count; ++a) {
      d.gamma[kMaxWindow * b + a] = coef[a];
      norm += coef[a] * coef[a];
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
  d.aa[4 * b] = count;
  d.aa[4 * b + 1] = start;
  d.accept[b] = accept;
}

__device__ __forceinline__ V3 anderson_value(const DeviceState& d, int b, int i, V3 raw) {
  if (d.W == 0 || !d.accept[b]) return raw;
  const int cap = d.W + 1, N = d.N;
  const int sc = d.aa[4 * b + 2], ss = d.aa[4 * b + 3];
  V3 v = ld3(d.hr, ((ss + sc) % cap) * N + i);
  V3 lo = ld3(d.hr, (ss % cap) * N + i);
  for (int a = 0; a < sc; ++a) {
    const V3 hi = ld3(d.hr, ((ss + a + 1) % cap) * N + i);
    v = v - d.gamma[kMaxWindow * b + a] * (hi - lo);
    lo = hi;
  }
  return v;
}

__global__ __launch_bounds__(kBlock) void k_apply_conv(DeviceState d) {
  const Chunk c = d.vchunks[blockIdx.x];
  const int b = c.body;
  if (!d.active[b]) return;
  __shared__ double smem[2 * kBlock / kWave];
  __shared__ double res[2];
  const int i = c.begin + threadIdx.x;
  double part[2] = {0.0, 0.0};
  if (i < c.end) {
    const V3 prev = ld3(d.x, i);
    V3 v = anderson_value(d, b, i, ld3(d.xr, i));
    if (!d.free[i]) v = ld3(d.xd, i);
    const V3 step = v - prev;
    part[0] = dot(step, step);
    part[1] = dot(prev, prev);
    st3(d.x, i, v);
  }
  block_sum<2>(part, smem, res);
  if (threadIdx.x < 2) d.conv_part[4 * blockIdx.x + threadIdx.x] = res[threadIdx.x];
}

__global__ void k_conv_check(DeviceState d, StepScalars s) {
  const int b = blockIdx.x;
  if (!d.active[b]) return;
  __shared__ double red[4];
  const int c0 = d.vcb[b], c1 = d.vcb[b + 1];
  if (threadIdx.x < 4) {
    double t = 0.0;
    for (int c = c0; c < c1; ++c) t += d.conv_part[4 * c + threadIdx.x];
    red[threadIdx.x] = t;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const int k = d.iter[b];
  const bool step_small = sqrt(red[0]) <= s.rel_tol * sqrt(red[1]) + s.abs_tol;
  const bool rhs_small = sqrt(red[2]) <= s.rel_tol * sqrt(red[3]) + s.abs_tol;
  d.iter[b] = k + 1;
  if (k > 0 && step_small && rhs_small) {
    d.active[b] = 0;
  } else if (k + 1 >= s.max_iterations) {
    d.active[b] = 0;
    atomicOr(d.status, 1);
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
    d.aa[4 * b] = d.aa[4 * b + 1] = d.aa[4 * b + 2] = d.aa[4 * b + 3] = 0;
    d.accept[b] = 0;
  }
  const int i = c.begin + threadIdx.x;
  if (i >= c.end) return;
  for (int k = 0; k < 3; ++k) {
    const double vel = (gqd ? static_cast<double>(gqd[3 * i + k]) : 0.0) / dt;
    d.arhs[3 * i + k] = (gq ? static_cast<double>(gq[3 * i + k]) : 0.0) + vel;
    d.sol[3 * i + k] = 0.0;
  }
}

__device__ __forceinline__ V3 hex_grad(const DeviceState& d, int b, int s, int j) {
  const double* G = d.hexg + 192 * b + 24 * s + 3 * j;
  return V3{G[0], G[1], G[2]};
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
  double jac[81];
  sample_jacobian(f, d.material, d.jwc[e], d.jwv[e], d.jwn[e], nh, wm, m, target, jac);
  const size_t stride = static_cast<size_t>(total);
  for (int t = 0; t < 81; ++t) d.jac[t * stride + sample] = jac[t];
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

// Element rows of K mu (mode 0) or of the Dirichlet coupling (A - K) mu
// (mode 1, elements with a pinned corner only).
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_adj_local_tet4(DeviceState d, const double* mu) {
  const Chunk c = d.echunks[blockIdx.x];
  if (MODE == 0 && !d.aactive[c.body]) return;
  const int e = c.begin + threadIdx.x;
  if (e >= c.end) return;
  if (MODE == 1 && d.pinmask[e] == 0) return;
  V3 g[4];
  tet_grads(d.poses, e, g);
  const int* nd = d.idx + 4 * e;
  M3 df = zero3();
#pragma unroll
  for (int k = 0; k < 4; ++k) add_outer(df, ld3(mu, nd[k]), g[k]);
  M3 ds;
  if (MODE == 0) {
    ds = jac_apply(d, e, df);
  } else {
    if (d.route[2 * c.body + 1] != 0.0) {
#pragma unroll
      for (int k = 0; k < 4; ++k) st3(d.contrib, 4 * e + k, V3{0, 0, 0});
      return;
    }
    ds = scale(d.wc[e] + d.wv[e] + d.wn[e], df);
    if (d.has_muscles) {
      const V3 m = ld3(d.mdir, e);
      ds = add(ds, scale(d.mw[e], outer(mul(df, m), m)));
    }
    ds = sub(ds, jac_apply(d, e, df));
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) st3(d.contrib, 4 * e + k, mul(ds, g[k]));
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_adj_local_hex8(DeviceState d, const double* mu) {
  const Chunk c = d.echunks[blockIdx.x];
  if (MODE == 0 && !d.aactive[c.body]) return;
  const int b = c.body;
  const int s = threadIdx.x & 7;
  const int e = c.begin + (threadIdx.x >> 3);
  bool valid = e < c.end;
  if (MODE == 1 && valid && d.pinmask[e] == 0) valid = false;
  const bool zero = MODE == 1 && d.route[2 * b + 1] != 0.0;
  double out[8][3];
#pragma unroll
  for (int j = 0; j < 8; ++j) out[j][0] = out[j][1] = out[j][2] = 0.0;
  if (valid && !zero) {
    const int* nd = d.idx + 8 * e;
    M3 df = zero3();
#pragma unroll
    for (int j = 0; j < 8; ++j) add_outer(df, ld3(mu, nd[j]), hex_grad(d, b, s, j));
    M3 ds;
    if (MODE == 0) {
      ds = jac_apply(d, 8 * e + s, df);
    } else {
      ds = scale(d.wc[e] + d.wv[e] + d.wn[e], df);
      if (d.has_muscles) {
        const V3 m = ld3(d.mdir, e);
        ds = add(ds, scale(d.mw[e], outer(mul(df, m), m)));
      }
      ds = sub(ds, jac_apply(d, 8 * e + s, df));
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const V3 r = mul(ds, hex_grad(d, b, s, j));
      out[j][0] = r.x;
      out[j][1] = r.y;
      out[j][2] = r.z;
    }
  }
#pragma unroll
  for (int off = 1; off < 8; off <<= 1)
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) out[j][k] += __shfl_xor(out[j][k], off, kWave);
  if (valid) {
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (j == s) st3(d.contrib, 8 * e + j, V3{out[j][0], out[j][1], out[j][2]});
  }
}

// work = sum(contrib) + rhs (gather with accumulate, adjoint.py:939).
__global__ __launch_bounds__(kBlock) void k_adj_gather(DeviceState d) {
  const Chunk c = d.vchunks[blockIdx.x];
  if (!d.aactive[c.body]) return;
  const int i = c.begin + threadIdx.x;
  if (i >= c.end) return;
  V3 t{0, 0, 0};
  for (int a = d.inc_off[i]; a < d.inc_off[i + 1]; ++a) t = t + ld3(d.contrib, d.inc_ent[a]);
  t = t + ld3(d.arhs, i);
  if (d.defect) {
    V3 ax{0, 0, 0};
    for (int q = d.a_off[i]; q < d.a_off[i + 1]; ++q) ax = ax + d.a_val[q] * ld3(d.sol, d.a_col[q]);
    t = t - ax;
  }
  st3(d.rhs, i, t);
}

// Safe-clip partial: Dirichlet entries zeroed, nonfinite components scrubbed
// (adjoint.py:184-204).
__global__ __launch_bounds__(kBlock) void k_adj_clip_partial(DeviceState d) {
  const Chunk c = d.vchunks[blockIdx.x];
  if (!d.aactive[c.body]) return;
  __shared__ double smem[kBlock / kWave];
  __shared__ double res[1];
  const int i = c.begin + threadIdx.x;
  double part[1] = {0.0};
  if (i < c.end) {
    V3 v = ld3(d.xr, i);
    if (!d.free[i]) v = V3{0, 0, 0};
    bool scrubbed = false;
    if (!isfinite(v.x)) { v.x = 0.0; scrubbed = true; }
    if (!isfinite(v.y)) { v.y = 0.0; scrubbed = true; }
    if (!isfinite(v.z)) { v.z = 0.0; scrubbed = true; }
    if (scrubbed) atomicOr(d.status, 16);
    st3(d.xr, i, v);
    part[0] = dot(v, v);
  }
  block_sum<1>(part, smem, res);
  if (threadIdx.x == 0) d.clip_part[blockIdx.x] = res[0];
}

// Clip scale of one body (adjoint.py:207-223): limit / ||v|| when above 1e6.
__global__ void k_adj_clip_scale(DeviceState d, int which) {
  const int b = blockIdx.x;
  if (!d.aactive[b]) return;
  if (threadIdx.x != 0) return;
  double t = 0.0;
  for (int c = d.vcb[b]; c < d.vcb[b + 1]; ++c) t += d.clip_part[c];
  const double len = sqrt(t);
  double sc = 1.0;
  const double limit = 1.0e6;
  if (isfinite(len) && len > limit) {
    sc = limit / len;
    atomicOr(d.status, 16);
  }
  d.ascale[2 * b + which] = sc;
}

__global__ __launch_bounds__(kBlock) void k_adj_push(DeviceState d) {
  const Chunk c = d.vchunks[blockIdx.x];
  const int b = c.body;
  if (!d.aactive[b]) return;
  __shared__ double smem[(kGramEntries + 1) * kBlock / kWave];
  __shared__ double res[kGramEntries + 1];
  const int i = c.begin + threadIdx.x;
  double acc[kGramEntries + 1];
#pragma unroll
  for (int k = 0; k < kGramEntries + 1; ++k) acc[k] = 0.0;
  if (i < c.end) {
    const double sc = d.ascale[2 * b];
    V3 raw = ld3(d.xr, i);
    if (sc != 1.0) raw = sc * raw;
    if (!d.free[i]) raw = V3{0, 0, 0};
    st3(d.xr, i, raw);
    const V3 prev = ld3(d.sol, i);
    const V3 step = raw - prev;
    acc[kGramEntries] = dot(step, step);
    if (d.W > 0) {
      double g[kGramEntries];
#pragma unroll
      for (int k = 0; k < kGramEntries; ++k) g[k] = 0.0;
      push_and_gram(d, c, i, prev, raw, g);
#pragma unroll
      for (int k = 0; k < kGramEntries; ++k) acc[k] = g[k];
    }
  }
  block_sum<kGramEntries + 1>(acc, smem, res);
  for (int k = threadIdx.x; k < kGramEntries; k += blockDim.x) d.gram_part[kGramEntries * blockIdx.x + k] = res[k];
  if (threadIdx.x == 0) d.conv_part[4 * blockIdx.x] = res[kGramEntries];
}

__global__ __launch_bounds__(kBlock) void k_adj_apply(DeviceState d) {
  const Chunk c = d.vchunks[blockIdx.x];
  const int b = c.body;
  if (!d.aactive[b]) return;
  __shared__ double smem[kBlock / kWave];
  __shared__ double res[1];
  const int i = c.begin + threadIdx.x;
  double part[1] = {0.0};
  if (i < c.end) {
    V3 v = anderson_value(d, b, i, ld3(d.xr, i));
    if (!d.free[i]) v = V3{0, 0, 0};
    bool scrubbed = false;
    if (!isfinite(v.x)) { v.x = 0.0; scrubbed = true; }
    if (!isfinite(v.y)) { v.y = 0.0; scrubbed = true; }
    if (!isfinite(v.z)) { v.z = 0.0; scrubbed = true; }
    if (scrubbed) atomicOr(d.status, 16);
    st3(d.xr, i, v);
    part[0] = dot(v, v);
  }
  block_sum<1>(part, smem, res);
  if (threadIdx.x == 0) d.clip_part[blockIdx.x] = res[0];
}

__global__ __launch_bounds__(kBlock) void k_adj_finish(DeviceState d) {
  const Chunk c = d.vchunks[blockIdx.x];
  const int b = c.body;
  if (!d.aactive[b]) return;
  __shared__ double smem[kBlock / kWave];
  __shared__ double res[1];
  const int i = c.begin + threadIdx.x;
  double part[1] = {0.0};
  if (i < c.end) {
    const double sc = d.ascale[2 * b + 1];
    V3 v = ld3(d.xr, i);
    if (sc != 1.0) v = sc * v;
    st3(d.sol, i, v);
    part[0] = dot(v, v);
  }
  block_sum<1>(part, smem, res);
  if (threadIdx.x == 0) d.conv_part[4 * blockIdx.x + 1] = res[0];
}

__global__ void k_adj_check(DeviceState d, StepScalars s) {
  const int b = blockIdx.x;
  if (!d.aactive[b]) return;
  if (threadIdx.x != 0) return;
  double t0 = 0.0, t1 = 0.0;
  for (int c = d.vcb[b]; c < d.vcb[b + 1]; ++c) {
    t0 += d.conv_part[4 * c];
    t1 += d.conv_part[4 * c + 1];
  }
  const int k = d.aiter[b];
  d.aiter[b] = k + 1;
  if (k > 0 && sqrt(t0) < s.rel_tol * sqrt(t1) + s.abs_tol) {
    d.aactive[b] = 0;
  } else if (k + 1 >= s.max_iterations) {
    d.aactive[b] = 0;
    atomicOr(d.status, 8);
  }
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
__device__ __forceinline__ void material_terms(const DeviceState& d, const M3& f, double& first_out,
                                               double& second_out, const V3* g, const V3* z, int K) {
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

__global__ __launch_bounds__(kBlock) void k_adj_material_tet4(DeviceState d, const double* x) {
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
    z[k] = ld3(d.sol, n);
  }
  double fv, sv;
  if (d.material == 0) {
    material_terms(d, f, fv, sv, g, z, 4);
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

__global__ __launch_bounds__(kBlock) void k_adj_material_hex8(DeviceState d, const double* x) {
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
      z[j] = ld3(d.sol, n);
    }
    if (d.material == 0) {
      material_terms(d, f, fv, sv, g, z, 8);
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

// Aggregate guard per body (adjoint.py:622-644), parallel over the body's elements.
__global__ __launch_bounds__(kBlock) void k_adj_material_guard(DeviceState d) {
  const int* ebeg = d.eoff;
  const int b = blockIdx.x;
  __shared__ double smem[2 * kBlock / kWave];
  __shared__ double res[2];
  double part[2] = {0.0, 0.0};
  for (int e = ebeg[b] + threadIdx.x; e < ebeg[b + 1]; e += kBlock) {
    part[0] += d.gmat_tmp[2 * e];
    part[1] += d.gmat_tmp[2 * e + 1];
  }
  block_sum<2>(part, smem, res);
  if (threadIdx.x != 0) return;
  const bool zero_energy = d.material == 0 && d.nh[3 * b + 1] == 0.0;
  const bool ok = !zero_energy && isfinite(res[0]) && isfinite(res[1]);
  if (!ok) atomicOr(d.status, 16);
  d.valid[b] = (ok && d.route[2 * b + 1] == 0.0) ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void k_adj_material_add(DeviceState d, double* gmat) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d.E) return;
  if (!d.valid[d.ebody[e]]) return;
  gmat[2 * e] += d.gmat_tmp[2 * e];
  gmat[2 * e + 1] += d.gmat_tmp[2 * e + 1];
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

template <typename Real, bool PRED>
void spmv_pair(const DeviceState& d, const double* in, double* mid, double* out, const double* base,
               const int* active, hipStream_t stream) {
  if (d.s_ntasks)
    hipLaunchKernelGGL((k_spmv<Real, PRED>), dim3(d.s_ntasks), dim3(kBlock), 0, stream, d.s_tasks, d.s_rows,
                       d.s_off, d.s_col, static_cast<const Real*>(d.s_val), in, mid, nullptr, d.vbody, active);
  if (d.t_ntasks)
    hipLaunchKernelGGL((k_spmv<Real, PRED>), dim3(d.t_ntasks), dim3(kBlock), 0, stream, d.t_tasks, d.t_rows,
                       d.t_off, d.t_col, static_cast<const Real*>(d.t_val), mid, out, base, d.vbody, active);
}

void spmv_pair_any(const DeviceState& d, const double* in, double* mid, double* out, const double* base,
                   const int* active, hipStream_t stream) {
  if (d.f32)
    spmv_pair<float, true>(d, in, mid, out, base, active, stream);
  else
    spmv_pair<double, true>(d, in, mid, out, base, active, stream);
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers.

void launch_begin_step(const DeviceState& d, const StepScalars& s, const float* q, const float* qd, const float* f,
                       const float* dirichlet, hipStream_t stream) {
  hipLaunchKernelGGL(k_begin_step, dim3(d.nvc), dim3(kBlock), 0, stream, d, s, q, qd, f, dirichlet);
}

void launch_forward_iteration(const DeviceState& d, const StepScalars& s, const double* targets, hipStream_t stream) {
  if (d.Q == 1)
    hipLaunchKernelGGL(k_local_tet4, dim3(d.nec), dim3(kBlock), 0, stream, d, targets);
  else
    hipLaunchKernelGGL(k_local_hex8, dim3(d.nec), dim3(kBlock), 0, stream, d, targets);
  hipLaunchKernelGGL(k_gather, dim3(d.nvc), dim3(kBlock), 0, stream, d);
  spmv_pair_any(d, d.rhs, d.y, d.xr, d.defect ? d.x : nullptr, d.active, stream);
  if (d.W > 0) {
    hipLaunchKernelGGL(k_push_gram, dim3(d.nvc), dim3(kBlock), 0, stream, d);
    hipLaunchKernelGGL(k_gram_solve, dim3(d.B), dim3(64), 0, stream, d, static_cast<const int*>(d.active));
  }
  hipLaunchKernelGGL(k_apply_conv, dim3(d.nvc), dim3(kBlock), 0, stream, d);
  hipLaunchKernelGGL(k_conv_check, dim3(d.B), dim3(64), 0, stream, d, s);
}

void launch_any_active(const DeviceState& d, const int* active, hipStream_t stream) {
  hipLaunchKernelGGL(k_any_active, dim3(1), dim3(kBlock), 0, stream, active, d.B, d.any_active);
}

void launch_write_state(const DeviceState& d, const StepScalars& s, float* q, float* qd, double* checkpoint,
                        hipStream_t stream) {
  hipLaunchKernelGGL(k_write_state, dim3(d.nvc), dim3(kBlock), 0, stream, d, s.dt, q, qd, checkpoint);
}

void launch_adjoint_setup(const DeviceState& d, const StepScalars& s, const double* x, const double* targets,
                          const float* gq_out, const float* gqd_out, hipStream_t stream) {
  hipLaunchKernelGGL(k_adj_seed, dim3(d.nvc), dim3(kBlock), 0, stream, d, s.dt, gq_out, gqd_out);
  const int total = d.E * d.Q;
  hipLaunchKernelGGL(k_jacobian, dim3((total + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, d, x, targets);
}

void launch_adjoint_iteration(const DeviceState& d, const StepScalars& s, hipStream_t stream) {
  if (d.Q == 1)
    hipLaunchKernelGGL(k_adj_local_tet4<0>, dim3(d.nec), dim3(kBlock), 0, stream, d,
                       static_cast<const double*>(d.sol));
  else
    hipLaunchKernelGGL(k_adj_local_hex8<0>, dim3(d.nec), dim3(kBlock), 0, stream, d,
                       static_cast<const double*>(d.sol));
  hipLaunchKernelGGL(k_adj_gather, dim3(d.nvc), dim3(kBlock), 0, stream, d);
  spmv_pair_any(d, d.rhs, d.y, d.xr, d.defect ? d.sol : nullptr, d.aactive, stream);
  hipLaunchKernelGGL(k_adj_clip_partial, dim3(d.nvc), dim3(kBlock), 0, stream, d);
  hipLaunchKernelGGL(k_adj_clip_scale, dim3(d.B), dim3(64), 0, stream, d, 0);
  hipLaunchKernelGGL(k_adj_push, dim3(d.nvc), dim3(kBlock), 0, stream, d);
  if (d.W > 0) hipLaunchKernelGGL(k_gram_solve, dim3(d.B), dim3(64), 0, stream, d, static_cast<const int*>(d.aactive));
  hipLaunchKernelGGL(k_adj_apply, dim3(d.nvc), dim3(kBlock), 0, stream, d);
  hipLaunchKernelGGL(k_adj_clip_scale, dim3(d.B), dim3(64), 0, stream, d, 1);
  hipLaunchKernelGGL(k_adj_finish, dim3(d.nvc), dim3(kBlock), 0, stream, d);
  hipLaunchKernelGGL(k_adj_check, dim3(d.B), dim3(64), 0, stream, d, s);
}

void launch_adjoint_outputs(const DeviceState& d, const StepScalars& s, const double* x, const float* gqd_out,
                            float* gq_in, float* gqd_in, float* gf_in, double* gmat, float* gdir, bool any_pinned,
                            hipStream_t stream) {
  hipLaunchKernelGGL(k_adj_route, dim3(d.B), dim3(kBlock), 0, stream, d);
  hipLaunchKernelGGL(k_adj_state_grads, dim3(d.nvc), dim3(kBlock), 0, stream, d, s.dt, gqd_out, gq_in, gqd_in,
                     gf_in);
  if (any_pinned) {
    if (d.Q == 1)
      hipLaunchKernelGGL(k_adj_local_tet4<1>, dim3(d.nec), dim3(kBlock), 0, stream, d,
                         static_cast<const double*>(d.sol));
    else
      hipLaunchKernelGGL(k_adj_local_hex8<1>, dim3(d.nec), dim3(kBlock), 0, stream, d,
                         static_cast<const double*>(d.sol));
    hipLaunchKernelGGL(k_adj_dirichlet, dim3(d.nvc), dim3(kBlock), 0, stream, d, s.dt, gq_in, gqd_in, gdir);
  }
  if (gmat) {
    if (d.Q == 1)
      hipLaunchKernelGGL(k_adj_material_tet4, dim3(d.nec), dim3(kBlock), 0, stream, d, x);
    else
      hipLaunchKernelGGL(k_adj_material_hex8, dim3(d.nec), dim3(kBlock), 0, stream, d, x);
  }
  (void)gmat;
}

void launch_material_finish(const DeviceState& d, double* gmat, hipStream_t stream) {
  hipLaunchKernelGGL(k_adj_material_guard, dim3(d.B), dim3(kBlock), 0, stream, d);
  hipLaunchKernelGGL(k_adj_material_add, dim3((d.E + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, d, gmat);
}

void launch_apply_inverse(const DeviceState& d, const double* rhs, double* out, hipStream_t stream) {
  if (d.f32)
    spmv_pair<float, false>(d, rhs, d.y, out, nullptr, d.ones, stream);
  else
    spmv_pair<double, false>(d, rhs, d.y, out, nullptr, d.ones, stream);
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
