// Pointwise FP64 math of the DiffPhD local step and its differential.
//
// Every function works on one 3x3 deformation gradient held in registers.
// Reference: DP/projections.py (Eigen-convention two-sided Jacobi SVD
// :29-167, projections :170-316, directional derivatives :323-441).
// The functions are __host__ __device__ so the same source is exercised by
// the HIP kernels and by host-side unit checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

namespace dphd {

#define DPHD_HD __host__ __device__ __forceinline__

constexpr double kEps = 2.220446049250313e-16;       // DBL_EPSILON
constexpr double kDblMin = 2.2250738585072014e-308;  // DBL_MIN
constexpr int kJacobiMaxSweeps = 64;

struct M3 {
  double a[3][3];
};
struct V3 {
  double x, y, z;
};

DPHD_HD double vget(const V3& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
DPHD_HD void vset(V3& v, int i, double s) {
  if (i == 0)
    v.x = s;
  else if (i == 1)
    v.y = s;
  else
    v.z = s;
}
DPHD_HD V3 v3(double x, double y, double z) { return V3{x, y, z}; }
DPHD_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
DPHD_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
DPHD_HD V3 operator*(double s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
DPHD_HD double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DPHD_HD double length(V3 a) { return sqrt(dot(a, a)); }

DPHD_HD M3 zero3() {
  M3 m;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m.a[i][j] = 0.0;
  return m;
}
DPHD_HD M3 eye3() {
  M3 m = zero3();
  m.a[0][0] = m.a[1][1] = m.a[2][2] = 1.0;
  return m;
}
DPHD_HD M3 diag3(V3 d) {
  M3 m = zero3();
  m.a[0][0] = d.x;
  m.a[1][1] = d.y;
  m.a[2][2] = d.z;
  return m;
}
DPHD_HD M3 transpose(const M3& m) {
  M3 t;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t.a[i][j] = m.a[j][i];
  return t;
}
DPHD_HD M3 mul(const M3& x, const M3& y) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = x.a[i][0] * y.a[0][j] + x.a[i][1] * y.a[1][j] + x.a[i][2] * y.a[2][j];
  return r;
}
DPHD_HD V3 mul(const M3& m, V3 v) {
  return V3{m.a[0][0] * v.x + m.a[0][1] * v.y + m.a[0][2] * v.z, m.a[1][0] * v.x + m.a[1][1] * v.y + m.a[1][2] * v.z,
            m.a[2][0] * v.x + m.a[2][1] * v.y + m.a[2][2] * v.z};
}
DPHD_HD M3 add(const M3& x, const M3& y) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = x.a[i][j] + y.a[i][j];
  return r;
}
DPHD_HD M3 sub(const M3& x, const M3& y) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = x.a[i][j] - y.a[i][j];
  return r;
}
DPHD_HD M3 scale(double s, const M3& x) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = s * x.a[i][j];
  return r;
}
DPHD_HD M3 outer(V3 u, V3 v) {
  M3 r;
  const double uu[3] = {u.x, u.y, u.z}, vv[3] = {v.x, v.y, v.z};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = uu[i] * vv[j];
  return r;
}
// U diag(d) V^T.
DPHD_HD M3 udvt(const M3& u, V3 d, const M3& v) {
  M3 r;
  const double dd[3] = {d.x, d.y, d.z};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.a[i][j] = u.a[i][0] * dd[0] * v.a[j][0] + u.a[i][1] * dd[1] * v.a[j][1] + u.a[i][2] * dd[2] * v.a[j][2];
  return r;
}

// Jacobi rotation of rows p, q on the left: rows <- [[c, s], [-s, c]] rows.
DPHD_HD void rotate_rows(M3& m, int p, int q, double c, double s) {
  for (int j = 0; j < 3; ++j) {
    const double x = m.a[p][j], y = m.a[q][j];
    m.a[p][j] = c * x + s * y;
    m.a[q][j] = -s * x + c * y;
  }
}
// Columns p, q on the right: cols <- cols [[c, s], [-s, c]].
DPHD_HD void rotate_cols(M3& m, int p, int q, double c, double s) {
  for (int i = 0; i < 3; ++i) {
    const double x = m.a[i][p], y = m.a[i][q];
    m.a[i][p] = c * x - s * y;
    m.a[i][q] = s * x + c * y;
  }
}
DPHD_HD void swap_cols(M3& m, int i, int j) {
  for (int r = 0; r < 3; ++r) {
    const double t = m.a[r][i];
    m.a[r][i] = m.a[r][j];
    m.a[r][j] = t;
  }
}

// Symmetric 2x2 Jacobi rotation zeroing the off-diagonal of [[x, y], [y, z]]
// (Eigen JacobiRotation::makeJacobi for real scalars).
DPHD_HD void make_jacobi(double x, double y, double z, double& c, double& s) {
  c = 1.0;
  s = 0.0;
  const double den = 2.0 * fabs(y);
  if (den >= kDblMin) {
    const double tau = (x - z) / den;
    const double w = sqrt(tau * tau + 1.0);
    const double t = tau > 0.0 ? 1.0 / (tau + w) : 1.0 / (tau - w);
    const double sign = t > 0.0 ? 1.0 : -1.0;
    const double nrm = 1.0 / sqrt(t * t + 1.0);
    s = -sign * (y / fabs(y)) * fabs(t) * nrm;
    c = nrm;
  }
}

// F = U diag(sigma) V^T with sigma >= 0 descending (Eigen JacobiSVD
// conventions: scaled two-sided sweeps over pairs (1,0), (2,0), (2,1),
// negative diagonals flip U columns, the sort stops at a zero maximum).
DPHD_HD void unsigned_svd(const M3& f, M3& u, V3& sigma, M3& v) {
  double sc = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) sc = fmax(sc, fabs(f.a[i][j]));
  if (sc == 0.0) sc = 1.0;
  M3 w;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) w.a[i][j] = f.a[i][j] / sc;
  u = eye3();
  v = eye3();
  double largest = fmax(fmax(fabs(w.a[0][0]), fabs(w.a[1][1])), fabs(w.a[2][2]));
  const double precision = 2.0 * kEps;
  for (int sweep = 0; sweep < kJacobiMaxSweeps; ++sweep) {
    bool finished = true;
    for (int pair = 0; pair < 3; ++pair) {
      const int p = pair == 0 ? 1 : 2;
      const int q = pair == 2 ? 1 : 0;
      const double threshold = fmax(kDblMin, precision * largest);
      if (fabs(w.a[p][q]) > threshold || fabs(w.a[q][p]) > threshold) {
        finished = false;
        const double t = w.a[p][p] + w.a[q][q];
        const double d = w.a[q][p] - w.a[p][q];
        double c1 = 1.0, s1 = 0.0;
        if (fabs(d) >= kDblMin) {
          const double ratio = t / d;
          const double root = sqrt(1.0 + ratio * ratio);
          s1 = 1.0 / root;
          c1 = ratio / root;
        }
        const double m00 = c1 * w.a[p][p] + s1 * w.a[q][p];
        const double m01 = c1 * w.a[p][q] + s1 * w.a[q][q];
        const double m11 = -s1 * w.a[p][q] + c1 * w.a[q][q];
        double cr, sr;
        make_jacobi(m00, m01, m11, cr, sr);
        const double cl = c1 * cr + s1 * sr;
        const double sl = s1 * cr - c1 * sr;
        rotate_rows(w, p, q, cl, sl);
        rotate_cols(u, p, q, cl, -sl);
        rotate_cols(w, p, q, cr, sr);
        rotate_cols(v, p, q, cr, sr);
        largest = fmax(largest, fmax(fabs(w.a[p][p]), fabs(w.a[q][q])));
      }
    }
    if (finished) break;
  }
  double s[3];
  for (int i = 0; i < 3; ++i) {
    const double d = w.a[i][i];
    s[i] = fabs(d) * sc;
    if (d < 0.0)
      for (int r = 0; r < 3; ++r) u.a[r][i] = -u.a[r][i];
  }
  for (int i = 0; i < 3; ++i) {
    int pos = i;
    for (int j = i + 1; j < 3; ++j)
      if (s[j] > s[pos]) pos = j;
    if (s[pos] == 0.0) break;
    if (pos != i) {
      const double t = s[i];
      s[i] = s[pos];
      s[pos] = t;
      swap_cols(u, i, pos);
      swap_cols(v, i, pos);
    }
  }
  sigma = V3{s[0], s[1], s[2]};
}

// Offset d with prod(sigma + d) = 1 by the reference fixed point; flags a
// singular F (smallest stretch below machine epsilon).
DPHD_HD V3 volume_offset(V3 sigma, int& singular) {
  singular = sigma.z < kEps ? 1 : 0;
  const double product = sigma.x * sigma.y * sigma.z;
  const double root = pow(product, 1.0 / 3.0);
  V3 d = V3{sigma.x / root - sigma.x, sigma.y / root - sigma.y, sigma.z / root - sigma.z};
  for (int it = 0; it < 50; ++it) {
    const V3 s = sigma + d;
    const double constraint = s.x * s.y * s.z - 1.0;
    const V3 g = V3{s.y * s.z, s.x * s.z, s.x * s.y};
    const double k = (dot(g, d) - constraint) / dot(g, g);
    const V3 next = k * g;
    const double diff = fmax(fmax(fabs(next.x - d.x), fabs(next.y - d.y)), fabs(next.z - d.z));
    if (diff <= kEps) break;
    d = next;
  }
  return d;
}

// Symmetric indefinite solve of the leading n x n block with Eigen's dense
// LDLT conventions: left-looking, largest remaining original diagonal as
// pivot, pseudo-inverse of D. a is row-major with leading dimension ld.
template <int NMAX>
DPHD_HD void ldlt_solve(const double* a_in, int ld, const double* b, int n, double* x) {
  double m[NMAX][NMAX];
  int perm[NMAX];
  for (int i = 0; i < n; ++i) {
    perm[i] = i;
    for (int j = 0; j < n; ++j) m[i][j] = a_in[i * ld + j];
  }
  for (int k = 0; k < n; ++k) {
    int biggest = k;
    for (int i = k + 1; i < n; ++i)
      if (fabs(m[i][i]) > fabs(m[biggest][biggest])) biggest = i;
    if (biggest != k) {
      for (int j = 0; j < n; ++j) {
        const double t = m[k][j];
        m[k][j] = m[biggest][j];
        m[biggest][j] = t;
      }
      for (int i = 0; i < n; ++i) {
        const double t = m[i][k];
        m[i][k] = m[i][biggest];
        m[i][biggest] = t;
      }
      const int t = perm[k];
      perm[k] = perm[biggest];
      perm[biggest] = t;
    }
    for (int j = 0; j < k; ++j) m[k][k] = m[k][k] - m[k][j] * m[k][j] * m[j][j];
    for (int i = k + 1; i < n; ++i) {
      for (int j = 0; j < k; ++j) m[i][k] = m[i][k] - m[i][j] * m[j][j] * m[k][j];
      if (m[k][k] != 0.0) m[i][k] = m[i][k] / m[k][k];
    }
  }
  double y[NMAX];
  for (int i = 0; i < n; ++i) y[i] = b[perm[i]];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) y[i] = y[i] - m[i][j] * y[j];
  for (int i = 0; i < n; ++i) y[i] = fabs(m[i][i]) > kDblMin ? y[i] / m[i][i] : 0.0;
  for (int r = 0; r < n; ++r) {
    const int i = n - 1 - r;
    for (int j = i + 1; j < n; ++j) y[i] = y[i] - m[j][i] * y[j];
  }
  for (int i = 0; i < n; ++i) x[perm[i]] = y[i];
}

// ldlt_solve for a compile-time size, with the same operations in the same
// order, written so that no local array is indexed by a run-time value (the
// pivot swaps and the permutation are compare-and-select over compile-time
// indices): on the device everything stays in registers.
template <int N>
DPHD_HD void ldlt_solve_regs(const double (&a_in)[N][N], const double (&b)[N], double (&x)[N]) {
  double m[N][N];
  int perm[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    perm[i] = i;
#pragma unroll
    for (int j = 0; j < N; ++j) m[i][j] = a_in[i][j];
  }
#pragma unroll
  for (int k = 0; k < N; ++k) {
    int biggest = k;
    double best = fabs(m[k][k]);
#pragma unroll
    for (int i = k + 1; i < N; ++i) {
      if (fabs(m[i][i]) > best) {
        biggest = i;
        best = fabs(m[i][i]);
      }
    }
#pragma unroll
    for (int i = k + 1; i < N; ++i) {
      if (i == biggest) {
#pragma unroll
        for (int j = 0; j < N; ++j) {
          const double t = m[k][j];
          m[k][j] = m[i][j];
          m[i][j] = t;
        }
#pragma unroll
        for (int r = 0; r < N; ++r) {
          const double t = m[r][k];
          m[r][k] = m[r][i];
          m[r][i] = t;
        }
        const int t = perm[k];
        perm[k] = perm[i];
        perm[i] = t;
      }
    }
#pragma unroll
    for (int j = 0; j < k; ++j) m[k][k] = m[k][k] - m[k][j] * m[k][j] * m[j][j];
#pragma unroll
    for (int i = k + 1; i < N; ++i) {
#pragma unroll
      for (int j = 0; j < k; ++j) m[i][k] = m[i][k] - m[i][j] * m[j][j] * m[k][j];
      if (m[k][k] != 0.0) m[i][k] = m[i][k] / m[k][k];
    }
  }
  double y[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double v = 0.0;
#pragma unroll
    for (int t = 0; t < N; ++t) v = perm[i] == t ? b[t] : v;
    y[i] = v;
  }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < i; ++j) y[i] = y[i] - m[i][j] * y[j];
#pragma unroll
  for (int i = 0; i < N; ++i) y[i] = fabs(m[i][i]) > kDblMin ? y[i] / m[i][i] : 0.0;
#pragma unroll
  for (int r = 0; r < N; ++r) {
    const int i = N - 1 - r;
#pragma unroll
    for (int j = i + 1; j < N; ++j) y[i] = y[i] - m[j][i] * y[j];
  }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int t = 0; t < N; ++t)
      if (perm[i] == t) x[t] = y[i];
}

DPHD_HD M3 nh_prox_hessian(V3 sigma, double mu, double lame, double penalty) {
  const double product = sigma.x * sigma.y * sigma.z;
  const double log_j = log(fmax(product, 1.0e-12));
  const double s[3] = {sigma.x, sigma.y, sigma.z};
  M3 h;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      h.a[i][j] = (i == j) ? mu + (mu + lame - lame * log_j) / (s[i] * s[i]) + penalty : lame / (s[i] * s[j]);
  return h;
}

// Proximal Neo-Hookean stretches by damped Newton (reference SolveProximal).
DPHD_HD V3 nh_prox_stretches(V3 target, double mu, double lame, double penalty) {
  const double floor_ = 1.0e-4;
  const double tol = 1.0e-10;
  const double t[3] = {target.x, target.y, target.z};
  double s[3] = {fmax(t[0], floor_), fmax(t[1], floor_), fmax(t[2], floor_)};
  for (int it = 0; it < 32; ++it) {
    const double product = s[0] * s[1] * s[2];
    const double log_j = log(fmax(product, 1.0e-12));
    double g[3];
    for (int i = 0; i < 3; ++i) g[i] = mu * s[i] - mu / s[i] + lame * log_j / s[i] + penalty * (s[i] - t[i]);
    if (sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]) < tol) break;
    M3 h = nh_prox_hessian(V3{s[0], s[1], s[2]}, mu, lame, penalty);
    for (int i = 0; i < 3; ++i) h.a[i][i] = h.a[i][i] + 1.0e-9;
    double delta[3];
    ldlt_solve<3>(&h.a[0][0], 3, g, 3, delta);
    double alpha = 1.0;
    double trial[3];
    for (int i = 0; i < 3; ++i) trial[i] = s[i] - alpha * delta[i];
    for (int k = 0; k < 20; ++k) {
      if (fmin(fmin(trial[0], trial[1]), trial[2]) > floor_ * 0.5) break;
      alpha *= 0.5;
      for (int i = 0; i < 3; ++i) trial[i] = s[i] - alpha * delta[i];
    }
    const double step = fmax(fmax(fabs(trial[0] - s[0]), fabs(trial[1] - s[1])), fabs(trial[2] - s[2]));
    for (int i = 0; i < 3; ++i) s[i] = fmax(trial[i], floor_);
    if (step < tol) break;
  }
  return V3{s[0], s[1], s[2]};
}

DPHD_HD V3 muscle_projection(const M3& f, V3 dir, double target) {
  const V3 fibre = mul(f, dir);
  const double len = length(fibre);
  return len > kEps ? (target / len) * fibre : V3{0.0, 0.0, 0.0};
}

// ---------------------------------------------------------------------------
// Directional derivatives (reference ProjectToManifoldDifferential).

DPHD_HD double det3(const M3& m) {
  return m.a[0][0] * (m.a[1][1] * m.a[2][2] - m.a[1][2] * m.a[2][1]) -
         m.a[0][1] * (m.a[1][0] * m.a[2][2] - m.a[1][2] * m.a[2][0]) +
         m.a[0][2] * (m.a[1][0] * m.a[2][1] - m.a[1][1] * m.a[2][0]);
}

// Cofactor inverse of a 3x3 matrix.
DPHD_HD M3 inverse3(const M3& m) {
  const double d = det3(m);
  const double r = 1.0 / d;
  M3 o;
  o.a[0][0] = (m.a[1][1] * m.a[2][2] - m.a[1][2] * m.a[2][1]) * r;
  o.a[0][1] = (m.a[0][2] * m.a[2][1] - m.a[0][1] * m.a[2][2]) * r;
  o.a[0][2] = (m.a[0][1] * m.a[1][2] - m.a[0][2] * m.a[1][1]) * r;
  o.a[1][0] = (m.a[1][2] * m.a[2][0] - m.a[1][0] * m.a[2][2]) * r;
  o.a[1][1] = (m.a[0][0] * m.a[2][2] - m.a[0][2] * m.a[2][0]) * r;
  o.a[1][2] = (m.a[0][2] * m.a[1][0] - m.a[0][0] * m.a[1][2]) * r;
  o.a[2][0] = (m.a[1][0] * m.a[2][1] - m.a[1][1] * m.a[2][0]) * r;
  o.a[2][1] = (m.a[0][1] * m.a[2][0] - m.a[0][0] * m.a[2][1]) * r;
  o.a[2][2] = (m.a[0][0] * m.a[1][1] - m.a[0][1] * m.a[1][0]) * r;
  return o;
}

// Solve a 4x4 system by Gaussian elimination with partial pivoting.
DPHD_HD void solve4(double a[4][4], double b[4], double x[4]) {
  for (int k = 0; k < 4; ++k) {
    int p = k;
    for (int i = k + 1; i < 4; ++i)
      if (fabs(a[i][k]) > fabs(a[p][k])) p = i;
    if (p != k) {
      for (int j = 0; j < 4; ++j) {
        const double t = a[k][j];
        a[k][j] = a[p][j];
        a[p][j] = t;
      }
      const double t = b[k];
      b[k] = b[p];
      b[p] = t;
    }
    for (int i = k + 1; i < 4; ++i) {
      const double f = a[i][k] / a[k][k];
      for (int j = k; j < 4; ++j) a[i][j] -= f * a[k][j];
      b[i] -= f * b[k];
    }
  }
  for (int i = 3; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < 4; ++j) s -= a[i][j] * x[j];
    x[i] = s / a[i][i];
  }
}

// dSvd: (dU, dsigma, dV); pairs with sigma_i - sigma_j <= 10 eps are skipped.
DPHD_HD void dsvd(const M3& u, V3 sigma, const M3& v, const M3& df, M3& du, V3& dsigma, M3& dv) {
  const M3 dp = mul(mul(transpose(u), df), v);
  dsigma = V3{dp.a[0][0], dp.a[1][1], dp.a[2][2]};
  const double s[3] = {sigma.x, sigma.y, sigma.z};
  M3 factor = zero3();
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j)
      if (s[i] - s[j] > 10.0 * kEps) {
        factor.a[i][j] = 1.0 / (s[j] * s[j] - s[i] * s[i]);
        factor.a[j][i] = -factor.a[i][j];
      }
  M3 ou, ov;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      // (dp S + S dp^T)_ij = dp_ij s_j + s_i dp_ji ; (S dp + dp^T S)_ij = s_i dp_ij + dp_ji s_j
      ou.a[i][j] = factor.a[i][j] * (dp.a[i][j] * s[j] + s[i] * dp.a[j][i]);
      ov.a[i][j] = factor.a[i][j] * (s[i] * dp.a[i][j] + dp.a[j][i] * s[j]);
    }
  du = mul(u, ou);
  dv = mul(v, ov);
}

// dR/dF of the polar rotation R = U V^T with S = V diag(sigma) V^T.
DPHD_HD M3 corotated_jvp(const M3& u, V3 sigma, const M3& v, const M3& df) {
  const M3 r = mul(u, transpose(v));
  const M3 s = udvt(v, sigma, v);
  const M3 lhs = sub(mul(transpose(r), df), mul(transpose(df), r));
  M3 a;
  a.a[0][0] = s.a[0][0] + s.a[1][1];
  a.a[0][1] = s.a[1][2];
  a.a[0][2] = -s.a[0][2];
  a.a[1][0] = s.a[1][2];
  a.a[1][1] = s.a[0][0] + s.a[2][2];
  a.a[1][2] = s.a[0][1];
  a.a[2][0] = -s.a[0][2];
  a.a[2][1] = s.a[0][1];
  a.a[2][2] = s.a[1][1] + s.a[2][2];
  const V3 xyz = mul(inverse3(a), V3{lhs.a[0][1], lhs.a[0][2], lhs.a[1][2]});
  M3 w = zero3();
  w.a[0][1] = xyz.x;
  w.a[0][2] = xyz.y;
  w.a[1][0] = -xyz.x;
  w.a[1][2] = xyz.z;
  w.a[2][0] = -xyz.y;
  w.a[2][1] = -xyz.z;
  return mul(r, w);
}

// Volume-projection derivative from the reference KKT system.
DPHD_HD M3 volume_jvp(const M3& u, V3 sigma, const M3& v, V3 offset, const M3& df) {
  const V3 s = sigma + offset;
  const V3 g = V3{s.y * s.z, s.x * s.z, s.x * s.y};
  M3 h = zero3();
  h.a[0][1] = s.z;
  h.a[0][2] = s.y;
  h.a[1][0] = s.z;
  h.a[1][2] = s.x;
  h.a[2][0] = s.y;
  h.a[2][1] = s.x;
  M3 du, dv;
  V3 dsig;
  dsvd(u, sigma, v, df, du, dsig, dv);
  const double magnitude = length(offset) / length(g);
  double multiplier = magnitude;
  const V3 plus = offset - magnitude * g;
  const V3 minus = offset + magnitude * g;
  if (dot(plus, plus) < dot(minus, minus)) multiplier = -magnitude;
  double kkt[4][4];
  const double gg[3] = {g.x, g.y, g.z};
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) kkt[i][j] = multiplier * h.a[i][j];
    kkt[i][i] = kkt[i][i] + 1.0;
    kkt[i][3] = gg[i];
    kkt[3][i] = gg[i];
  }
  kkt[3][3] = 0.0;
  const V3 top = -multiplier * mul(h, dsig);
  double rhs[4] = {top.x, top.y, top.z, -dot(g, dsig)};
  double sol[4];
  solve4(kkt, rhs, sol);
  const V3 doff = V3{sol[0], sol[1], sol[2]};
  M3 out = df;
  out = add(out, mul(mul(du, diag3(offset)), transpose(v)));
  out = add(out, udvt(u, doff, v));
  out = add(out, mul(mul(u, diag3(offset)), transpose(dv)));
  return out;
}

DPHD_HD M3 nh_prox_jvp(const M3& u, V3 sigma, const M3& v, V3 stretches, const M3& df, double mu, double lame,
                       double penalty) {
  M3 du, dv;
  V3 dsig;
  dsvd(u, sigma, v, df, du, dsig, dv);
  M3 h = nh_prox_hessian(stretches, mu, lame, penalty);
  for (int i = 0; i < 3; ++i) h.a[i][i] = h.a[i][i] + 1.0e-9;
  const double rhs[3] = {penalty * dsig.x, penalty * dsig.y, penalty * dsig.z};
  double dst[3];
  ldlt_solve<3>(&h.a[0][0], 3, rhs, 3, dst);
  M3 out = mul(mul(du, diag3(stretches)), transpose(v));
  out = add(out, udvt(u, V3{dst[0], dst[1], dst[2]}, v));
  out = add(out, mul(mul(u, diag3(stretches)), transpose(dv)));
  return out;
}

DPHD_HD V3 muscle_jvp(const M3& f, V3 dir, double target, const M3& df) {
  const V3 fibre = mul(f, dir);
  const double len = length(fibre);
  if (!(len > kEps)) return V3{0.0, 0.0, 0.0};
  const V3 dfib = mul(df, dir);
  const double dlen = dot((1.0 / len) * fibre, dfib);
  const double ratio = target / len;
  return ratio * dfib - (target / (len * len) * dlen) * fibre;
}

// True Neo-Hookean first Piola-Kirchhoff stress.
DPHD_HD M3 neo_hookean_stress(const M3& f, double mu, double lame) {
  const double log_j = log(fmax(det3(f), 1.0e-12));
  const M3 it = transpose(inverse3(f));
  return add(scale(mu, sub(f, it)), scale(lame * log_j, it));
}

// ---------------------------------------------------------------------------
// One quadrature sample of the forward local step: the weighted stress
// sum_energies w (P(F) - F_D) (+ fibre) of DP/kernels.py:240-270 and
// DP/hex.py:140-167. Returns status bits (2 on a singular volume projection).
struct SampleWeights {
  double wc, wv, wn, wm;
};
struct NhParams {
  double mu, lame, penalty;
};

DPHD_HD int sample_stress(const M3& f, const M3& fd, SampleWeights w, NhParams nh, V3 mdir, double mtarget,
                          bool always_svd, M3& stress) {
  int status = 0;
  stress = zero3();
  if (always_svd || w.wc != 0.0 || w.wv != 0.0 || w.wn != 0.0) {
    M3 u, v;
    V3 sigma;
    unsigned_svd(f, u, sigma, v);
    if (w.wc != 0.0) stress = add(stress, scale(w.wc, sub(mul(u, transpose(v)), fd)));
    if (w.wv != 0.0) {
      int singular = 0;
      const V3 off = volume_offset(sigma, singular);
      if (singular) status |= 2;
      stress = add(stress, scale(w.wv, sub(udvt(u, sigma + off, v), fd)));
    }
    if (w.wn != 0.0) {
      const V3 st = nh_prox_stretches(sigma, nh.mu, nh.lame, nh.penalty);
      stress = add(stress, scale(w.wn, sub(udvt(u, st, v), fd)));
    }
  }
  if (w.wm != 0.0) {
    const V3 fib = muscle_projection(f, mdir, mtarget);
    stress = add(stress, scale(w.wm, outer(fib - mul(fd, mdir), mdir)));
  }
  return status;
}

// The 9 columns d(w P)/dF . E_c, c = b + 3 n (Eigen column-major flatten),
// of one sample (DP/adjoint.py:92-137), handed to sink(c, column) one at a
// time, so a caller can store each column without holding all 81 values.
template <typename Sink>
DPHD_HD void sample_jacobian_columns(const M3& f, int material, double wc, double wv, double wn, NhParams nh,
                                     double wm, V3 mdir, double mtarget, Sink&& sink) {
  M3 u, v;
  V3 sigma;
  unsigned_svd(f, u, sigma, v);
  V3 off = V3{0, 0, 0}, st = V3{0, 0, 0};
  if (material == 0) {
    int singular = 0;
    off = volume_offset(sigma, singular);
  } else {
    st = nh_prox_stretches(sigma, nh.mu, nh.lame, nh.penalty);
  }
  for (int c = 0; c < 9; ++c) {
    M3 df = zero3();
    df.a[c % 3][c / 3] = 1.0;
    M3 dp = zero3();
    if (material == 0) {
      if (wc != 0.0) dp = add(dp, scale(wc, corotated_jvp(u, sigma, v, df)));
      if (wv != 0.0) dp = add(dp, scale(wv, volume_jvp(u, sigma, v, off, df)));
    } else {
      dp = add(dp, scale(wn, nh_prox_jvp(u, sigma, v, st, df, nh.mu, nh.lame, nh.penalty)));
    }
    if (wm != 0.0) dp = add(dp, scale(wm, outer(muscle_jvp(f, mdir, mtarget, df), mdir)));
    sink(c, dp);
  }
}

// The same columns as jac[c][i][j], row-major 3x3 each.
DPHD_HD void sample_jacobian(const M3& f, int material, double wc, double wv, double wn, NhParams nh, double wm,
                             V3 mdir, double mtarget, double* jac /* 81 */) {
  sample_jacobian_columns(f, material, wc, wv, wn, nh, wm, mdir, mtarget, [&](int c, const M3& dp) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) jac[c * 9 + i * 3 + j] = dp.a[i][j];
  });
}

}  // namespace dphd
