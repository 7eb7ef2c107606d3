/*
 * TEST ORACLE ONLY — plain-C FP64 restatement of the reference DiffPhD
 * projective-dynamics step and backward pass, in the order of the
 * reference's CPU execution path (Warp CPU device). Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it, and
 * only as the checker; the product never calls it.
 *
 * Paths below are under src/newton_solver_diff_phd/_src/solvers/diff_phd/.
 *   svd3, make_jacobi ........ projections.py:29-167 (Eigen JacobiSVD)
 *   volume_offset ............ projections.py:237-260
 *   ldlt_small ............... projections.py:178-228, anderson.py:97-147
 *   nh_stretches ............. projections.py:263-305
 *   muscle ................... projections.py:308-316
 *   planar ................... hex.py:246-260
 *   mesh contact ............. mesh_contact.py:34-177
 *   frictional boundary ...... contact.py (diffphd_contact.c)
 *   dsvd / *_jvp ............. projections.py:323-433
 *   forward step ............. solver.py:689-786 with kernels.py:155-414,
 *                              hex.py:122-187, sparse.py:194-206 (CPU SpMV),
 *                              anderson.py:38-317
 *   backward ................. adjoint.py:70-321, 474-659, 1088-1259
 *                              (activation gradients :595-620)
 * Sums run in the reference's CPU order: sequential CSR rows, incidence
 * entries in stored order, 16-node chunk partials reduced in chunk order.
 * Compiled with -ffp-contract=off.
 *
 * Pinned by: golden vectors from tests/reference_oracle.py (dense NumPy
 * oracle of the same step and backward, LAPACK SVD), closed forms (free
 * fall, rest-shape stationarity), np.linalg.svd, central finite differences.
 */
#include "diffphd_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "contact_oracle.h"

#define EPS 2.220446049250313e-16
#define DMIN 2.2250738585072014e-308
#define CHUNK 16
#define CLIP_NORM 1.0e6

/* ------------------------------------------------------------------------ */
/* 3x3 helpers, row-major double[9].                                        */

static void m_eye(double* m) {
  memset(m, 0, 9 * sizeof(double));
  m[0] = m[4] = m[8] = 1.0;
}
static void m_mul(const double* a, const double* b, double* out) {
  double r[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
  memcpy(out, r, sizeof(r));
}
static void m_t(const double* a, double* out) {
  double r[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[3 * i + j] = a[3 * j + i];
  memcpy(out, r, sizeof(r));
}
static void m_udvt(const double* u, const double* d, const double* v, double* out) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      out[3 * i + j] = u[3 * i] * d[0] * v[3 * j] + u[3 * i + 1] * d[1] * v[3 * j + 1] + u[3 * i + 2] * d[2] * v[3 * j + 2];
}
static void m_vec(const double* m, const double* x, double* out) {
  double r[3];
  for (int i = 0; i < 3; ++i) r[i] = m[3 * i] * x[0] + m[3 * i + 1] * x[1] + m[3 * i + 2] * x[2];
  memcpy(out, r, sizeof(r));
}
static double v_dot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* ------------------------------------------------------------------------ */
/* Eigen-convention two-sided Jacobi SVD.                                   */

static void jac_rows(double* m, int p, int q, double c, double s) {
  for (int j = 0; j < 3; ++j) {
    double x = m[3 * p + j], y = m[3 * q + j];
    m[3 * p + j] = c * x + s * y;
    m[3 * q + j] = -s * x + c * y;
  }
}
static void jac_cols(double* m, int p, int q, double c, double s) {
  for (int i = 0; i < 3; ++i) {
    double x = m[3 * i + p], y = m[3 * i + q];
    m[3 * i + p] = c * x - s * y;
    m[3 * i + q] = s * x + c * y;
  }
}
static void make_jacobi(double x, double y, double z, double* c, double* s) {
  *c = 1.0;
  *s = 0.0;
  double den = 2.0 * fabs(y);
  if (den >= DMIN) {
    double tau = (x - z) / den;
    double w = sqrt(tau * tau + 1.0);
    double t = tau > 0.0 ? 1.0 / (tau + w) : 1.0 / (tau - w);
    double sign = t > 0.0 ? 1.0 : -1.0;
    double n = 1.0 / sqrt(t * t + 1.0);
    *s = -sign * (y / fabs(y)) * fabs(t) * n;
    *c = n;
  }
}
static void svd3(const double* f, double* u, double* s, double* v) {
  double sc = 0.0, w[9];
  for (int t = 0; t < 9; ++t) sc = fmax(sc, fabs(f[t]));
  if (sc == 0.0) sc = 1.0;
  for (int t = 0; t < 9; ++t) w[t] = f[t] / sc;
  m_eye(u);
  m_eye(v);
  double largest = fmax(fmax(fabs(w[0]), fabs(w[4])), fabs(w[8]));
  static const int P[3] = {1, 2, 2}, Q[3] = {0, 0, 1};
  for (int sweep = 0; sweep < 64; ++sweep) {
    int done = 1;
    for (int k = 0; k < 3; ++k) {
      int p = P[k], q = Q[k];
      double thr = fmax(DMIN, 2.0 * EPS * largest);
      if (fabs(w[3 * p + q]) > thr || fabs(w[3 * q + p]) > thr) {
        done = 0;
        double t = w[3 * p + p] + w[3 * q + q], d = w[3 * q + p] - w[3 * p + q];
        double c1 = 1.0, s1 = 0.0;
        if (fabs(d) >= DMIN) {
          double r = t / d, root = sqrt(1.0 + r * r);
          s1 = 1.0 / root;
          c1 = r / root;
        }
        double m00 = c1 * w[3 * p + p] + s1 * w[3 * q + p];
        double m01 = c1 * w[3 * p + q] + s1 * w[3 * q + q];
        double m11 = -s1 * w[3 * p + q] + c1 * w[3 * q + q];
        double cr, sr;
        make_jacobi(m00, m01, m11, &cr, &sr);
        double cl = c1 * cr + s1 * sr, sl = s1 * cr - c1 * sr;
        jac_rows(w, p, q, cl, sl);
        jac_cols(u, p, q, cl, -sl);
        jac_cols(w, p, q, cr, sr);
        jac_cols(v, p, q, cr, sr);
        largest = fmax(largest, fmax(fabs(w[3 * p + p]), fabs(w[3 * q + q])));
      }
    }
    if (done) break;
  }
  for (int i = 0; i < 3; ++i) {
    double d = w[4 * i];
    s[i] = fabs(d) * sc;
    if (d < 0.0)
      for (int r = 0; r < 3; ++r) u[3 * r + i] = -u[3 * r + i];
  }
  for (int i = 0; i < 3; ++i) {
    int pos = i;
    for (int j = i + 1; j < 3; ++j)
      if (s[j] > s[pos]) pos = j;
    if (s[pos] == 0.0) break;
    if (pos != i) {
      double t = s[i];
      s[i] = s[pos];
      s[pos] = t;
      for (int r = 0; r < 3; ++r) {
        t = u[3 * r + i];
        u[3 * r + i] = u[3 * r + pos];
        u[3 * r + pos] = t;
        t = v[3 * r + i];
        v[3 * r + i] = v[3 * r + pos];
        v[3 * r + pos] = t;
      }
    }
  }
}

static int volume_offset(const double* s, double* d) {
  int singular = s[2] < EPS;
  double root = pow(s[0] * s[1] * s[2], 1.0 / 3.0);
  for (int i = 0; i < 3; ++i) d[i] = s[i] / root - s[i];
  for (int it = 0; it < 50; ++it) {
    double t[3] = {s[0] + d[0], s[1] + d[1], s[2] + d[2]};
    double c = t[0] * t[1] * t[2] - 1.0;
    double g[3] = {t[1] * t[2], t[0] * t[2], t[0] * t[1]};
    double k = (v_dot(g, d) - c) / v_dot(g, g);
    double nx[3] = {k * g[0], k * g[1], k * g[2]};
    double diff = fmax(fmax(fabs(nx[0] - d[0]), fabs(nx[1] - d[1])), fabs(nx[2] - d[2]));
    if (diff <= EPS) break;
    d[0] = nx[0];
    d[1] = nx[1];
    d[2] = nx[2];
  }
  return singular;
}

/* Eigen dense LDLT of the leading n x n block (n <= 8). */
static void ldlt_small(const double* a, int ld, const double* b, int n, double* x) {
  double m[8][8];
  int perm[8];
  for (int i = 0; i < n; ++i) {
    perm[i] = i;
    for (int j = 0; j < n; ++j) m[i][j] = a[i * ld + j];
  }
  for (int k = 0; k < n; ++k) {
    int big = k;
    for (int i = k + 1; i < n; ++i)
      if (fabs(m[i][i]) > fabs(m[big][big])) big = i;
    if (big != k) {
      for (int j = 0; j < n; ++j) {
        double t = m[k][j];
        m[k][j] = m[big][j];
        m[big][j] = t;
      }
      for (int i = 0; i < n; ++i) {
        double t = m[i][k];
        m[i][k] = m[i][big];
        m[i][big] = t;
      }
      int t = perm[k];
      perm[k] = perm[big];
      perm[big] = t;
    }
    for (int j = 0; j < k; ++j) m[k][k] = m[k][k] - m[k][j] * m[k][j] * m[j][j];
    for (int i = k + 1; i < n; ++i) {
      for (int j = 0; j < k; ++j) m[i][k] = m[i][k] - m[i][j] * m[j][j] * m[k][j];
      if (m[k][k] != 0.0) m[i][k] = m[i][k] / m[k][k];
    }
  }
  double y[8];
  for (int i = 0; i < n; ++i) y[i] = b[perm[i]];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) y[i] = y[i] - m[i][j] * y[j];
  for (int i = 0; i < n; ++i) y[i] = fabs(m[i][i]) > DMIN ? y[i] / m[i][i] : 0.0;
  for (int r = 0; r < n; ++r) {
    int i = n - 1 - r;
    for (int j = i + 1; j < n; ++j) y[i] = y[i] - m[j][i] * y[j];
  }
  for (int i = 0; i < n; ++i) x[perm[i]] = y[i];
}

static void nh_hessian(const double* s, double mu, double lame, double pen, double* h) {
  double lj = log(fmax(s[0] * s[1] * s[2], 1.0e-12));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      h[3 * i + j] = i == j ? mu + (mu + lame - lame * lj) / (s[i] * s[i]) + pen : lame / (s[i] * s[j]);
}
static void nh_stretches(const double* t, double mu, double lame, double pen, double* s) {
  const double fl = 1.0e-4, tol = 1.0e-10;
  for (int i = 0; i < 3; ++i) s[i] = fmax(t[i], fl);
  for (int it = 0; it < 32; ++it) {
    double lj = log(fmax(s[0] * s[1] * s[2], 1.0e-12)), g[3], h[9], dl[3], tr[3];
    for (int i = 0; i < 3; ++i) g[i] = mu * s[i] - mu / s[i] + lame * lj / s[i] + pen * (s[i] - t[i]);
    if (sqrt(v_dot(g, g)) < tol) break;
    nh_hessian(s, mu, lame, pen, h);
    for (int i = 0; i < 3; ++i) h[4 * i] += 1.0e-9;
    ldlt_small(h, 3, g, 3, dl);
    double alpha = 1.0;
    for (int i = 0; i < 3; ++i) tr[i] = s[i] - alpha * dl[i];
    for (int k = 0; k < 20; ++k) {
      if (fmin(fmin(tr[0], tr[1]), tr[2]) > fl * 0.5) break;
      alpha *= 0.5;
      for (int i = 0; i < 3; ++i) tr[i] = s[i] - alpha * dl[i];
    }
    double step = fmax(fmax(fabs(tr[0] - s[0]), fabs(tr[1] - s[1])), fabs(tr[2] - s[2]));
    for (int i = 0; i < 3; ++i) s[i] = fmax(tr[i], fl);
    if (step < tol) break;
  }
}
static void muscle(const double* f, const double* m, double target, double* out) {
  double fib[3];
  m_vec(f, m, fib);
  double len = sqrt(v_dot(fib, fib));
  for (int i = 0; i < 3; ++i) out[i] = len > EPS ? target / len * fib[i] : 0.0;
}

/* ------------------------------------------------------------------------ */
/* Directional derivatives.                                                 */

static double det3(const double* m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}
static void inv3(const double* m, double* o) {
  double r = 1.0 / det3(m);
  o[0] = (m[4] * m[8] - m[5] * m[7]) * r;
  o[1] = (m[2] * m[7] - m[1] * m[8]) * r;
  o[2] = (m[1] * m[5] - m[2] * m[4]) * r;
  o[3] = (m[5] * m[6] - m[3] * m[8]) * r;
  o[4] = (m[0] * m[8] - m[2] * m[6]) * r;
  o[5] = (m[2] * m[3] - m[0] * m[5]) * r;
  o[6] = (m[3] * m[7] - m[4] * m[6]) * r;
  o[7] = (m[1] * m[6] - m[0] * m[7]) * r;
  o[8] = (m[0] * m[4] - m[1] * m[3]) * r;
}
/* 4x4 solve by Gauss-Jordan with partial pivoting. */
static void solve4(double a[4][4], double* b, double* x) {
  for (int k = 0; k < 4; ++k) {
    int p = k;
    for (int i = k + 1; i < 4; ++i)
      if (fabs(a[i][k]) > fabs(a[p][k])) p = i;
    for (int j = 0; j < 4; ++j) {
      double t = a[k][j];
      a[k][j] = a[p][j];
      a[p][j] = t;
    }
    double t = b[k];
    b[k] = b[p];
    b[p] = t;
    for (int i = 0; i < 4; ++i)
      if (i != k) {
        double f = a[i][k] / a[k][k];
        for (int j = 0; j < 4; ++j) a[i][j] -= f * a[k][j];
        b[i] -= f * b[k];
      }
  }
  for (int i = 0; i < 4; ++i) x[i] = b[i] / a[i][i];
}
static void dsvd(const double* u, const double* s, const double* v, const double* df, double* du, double* ds,
                 double* dv) {
  double ut[9], dp[9], fac[9] = {0}, ou[9], ov[9];
  m_t(u, ut);
  m_mul(ut, df, dp);
  m_mul(dp, v, dp);
  for (int i = 0; i < 3; ++i) ds[i] = dp[4 * i];
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j)
      if (s[i] - s[j] > 10.0 * EPS) {
        fac[3 * i + j] = 1.0 / (s[j] * s[j] - s[i] * s[i]);
        fac[3 * j + i] = -fac[3 * i + j];
      }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      ou[3 * i + j] = fac[3 * i + j] * (dp[3 * i + j] * s[j] + s[i] * dp[3 * j + i]);
      ov[3 * i + j] = fac[3 * i + j] * (s[i] * dp[3 * i + j] + dp[3 * j + i] * s[j]);
    }
  m_mul(u, ou, du);
  m_mul(v, ov, dv);
}
static void corot_jvp(const double* u, const double* sg, const double* v, const double* df, double* out) {
  double vt[9], r[9], s[9], rt[9], dft[9], a1[9], a2[9], lhs[9], a[9], ai[9], w[9] = {0};
  m_t(v, vt);
  m_mul(u, vt, r);
  m_udvt(v, sg, v, s);
  m_t(r, rt);
  m_t(df, dft);
  m_mul(rt, df, a1);
  m_mul(dft, r, a2);
  for (int t = 0; t < 9; ++t) lhs[t] = a1[t] - a2[t];
  a[0] = s[0] + s[4];
  a[1] = s[5];
  a[2] = -s[2];
  a[3] = s[5];
  a[4] = s[0] + s[8];
  a[5] = s[1];
  a[6] = -s[2];
  a[7] = s[1];
  a[8] = s[4] + s[8];
  inv3(a, ai);
  double rhs[3] = {lhs[1], lhs[2], lhs[5]}, xyz[3];
  m_vec(ai, rhs, xyz);
  w[1] = xyz[0];
  w[2] = xyz[1];
  w[3] = -xyz[0];
  w[5] = xyz[2];
  w[6] = -xyz[1];
  w[7] = -xyz[2];
  m_mul(r, w, out);
}
static void vol_jvp(const double* u, const double* sg, const double* v, const double* off, const double* df,
                    double* out) {
  double s[3] = {sg[0] + off[0], sg[1] + off[1], sg[2] + off[2]};
  double g[3] = {s[1] * s[2], s[0] * s[2], s[0] * s[1]};
  double h[9] = {0, s[2], s[1], s[2], 0, s[0], s[1], s[0], 0};
  double du[9], dv[9], dsg[3];
  dsvd(u, sg, v, df, du, dsg, dv);
  double mag = sqrt(v_dot(off, off)) / sqrt(v_dot(g, g));
  double mult = mag, pl[3], mn[3];
  for (int i = 0; i < 3; ++i) {
    pl[i] = off[i] - mag * g[i];
    mn[i] = off[i] + mag * g[i];
  }
  if (v_dot(pl, pl) < v_dot(mn, mn)) mult = -mag;
  double kkt[4][4], hd[3], rhs[4], sol[4];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) kkt[i][j] = mult * h[3 * i + j];
    kkt[i][i] += 1.0;
    kkt[i][3] = g[i];
    kkt[3][i] = g[i];
  }
  kkt[3][3] = 0.0;
  m_vec(h, dsg, hd);
  for (int i = 0; i < 3; ++i) rhs[i] = -mult * hd[i];
  rhs[3] = -v_dot(g, dsg);
  solve4(kkt, rhs, sol);
  double t1[9], t2[9], t3[9], vt[9], dvt[9], dg[9] = {0};
  m_t(v, vt);
  m_t(dv, dvt);
  for (int i = 0; i < 3; ++i) dg[4 * i] = off[i];
  m_mul(du, dg, t1);
  m_mul(t1, vt, t1);
  m_udvt(u, sol, v, t2);
  m_mul(u, dg, t3);
  m_mul(t3, dvt, t3);
  for (int t = 0; t < 9; ++t) out[t] = df[t] + t1[t] + t2[t] + t3[t];
}
static void nh_jvp(const double* u, const double* sg, const double* v, const double* st, const double* df, double mu,
                   double lame, double pen, double* out) {
  double du[9], dv[9], dsg[3], h[9], rhs[3], dst[3];
  dsvd(u, sg, v, df, du, dsg, dv);
  nh_hessian(st, mu, lame, pen, h);
  for (int i = 0; i < 3; ++i) h[4 * i] += 1.0e-9;
  for (int i = 0; i < 3; ++i) rhs[i] = pen * dsg[i];
  ldlt_small(h, 3, rhs, 3, dst);
  double t1[9], t2[9], t3[9], vt[9], dvt[9], dg[9] = {0};
  m_t(v, vt);
  m_t(dv, dvt);
  for (int i = 0; i < 3; ++i) dg[4 * i] = st[i];
  m_mul(du, dg, t1);
  m_mul(t1, vt, t1);
  m_udvt(u, dst, v, t2);
  m_mul(u, dg, t3);
  m_mul(t3, dvt, t3);
  for (int t = 0; t < 9; ++t) out[t] = t1[t] + t2[t] + t3[t];
}
static void muscle_jvp(const double* f, const double* m, double target, const double* df, double* out) {
  double fib[3], dfib[3];
  m_vec(f, m, fib);
  double len = sqrt(v_dot(fib, fib));
  if (!(len > EPS)) {
    out[0] = out[1] = out[2] = 0.0;
    return;
  }
  m_vec(df, m, dfib);
  double unit[3] = {fib[0] / len, fib[1] / len, fib[2] / len};
  double dlen = v_dot(unit, dfib), ratio = target / len;
  for (int i = 0; i < 3; ++i) out[i] = ratio * dfib[i] - target / (len * len) * dlen * fib[i];
}

/* ------------------------------------------------------------------------ */
/* Element-level pieces shared by the step and the backward pass.           */

static const double* sample_grads(const oracle_problem* p, int e, int s) {
  return p->samples == 1 ? p->grads + (size_t)e * p->corners * 3 : p->grads + (size_t)s * p->corners * 3;
}
static void deformation(const oracle_problem* p, const double* x, int e, int s, double* f) {
  const int K = p->corners;
  const int* nodes = p->indices + (size_t)e * K;
  const double* g = sample_grads(p, e, s);
  memset(f, 0, 9 * sizeof(double));
  for (int k = 0; k < K; ++k) {
    const double* xk = x + 3 * (size_t)nodes[k];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) f[3 * i + j] = f[3 * i + j] + xk[i] * g[3 * k + j];
  }
}

/* contrib[(e*Q + s)*K + k] for the forward local step. */
static int local_step(const oracle_problem* p, const double* x, const double* bnd, const double* targets,
                      double* contrib) {
  const int K = p->corners, Q = p->samples;
  int status = 0;
  const int total = p->elements * Q;
#pragma omp parallel for schedule(static) reduction(| : status)
  for (int sample = 0; sample < total; ++sample) {
    const int e = sample / Q, s = sample % Q;
    double f[9], fd[9], st[9] = {0};
    deformation(p, x, e, s, f);
    deformation(p, bnd, e, s, fd);
    const double wc = p->wc[e], wv = p->wv[e], wn = p->wn[e];
    if (Q > 1 || wc != 0.0 || wv != 0.0 || wn != 0.0) {
      double u[9], sg[3], v[9], proj[9];
      svd3(f, u, sg, v);
      if (wc != 0.0) {
        double vt[9];
        m_t(v, vt);
        m_mul(u, vt, proj);
        for (int t = 0; t < 9; ++t) st[t] = st[t] + wc * (proj[t] - fd[t]);
      }
      if (wv != 0.0) {
        double d[3], sd[3];
        if (volume_offset(sg, d)) status |= 2;
        for (int i = 0; i < 3; ++i) sd[i] = sg[i] + d[i];
        m_udvt(u, sd, v, proj);
        for (int t = 0; t < 9; ++t) st[t] = st[t] + wv * (proj[t] - fd[t]);
      }
      if (wn != 0.0) {
        double str[3];
        nh_stretches(sg, p->nh_mu, p->nh_lame, p->nh_penalty, str);
        m_udvt(u, str, v, proj);
        for (int t = 0; t < 9; ++t) st[t] = st[t] + wn * (proj[t] - fd[t]);
      }
    }
    if (p->mw && p->mw[e] != 0.0) {
      const double* m = p->mdir + 3 * (size_t)e;
      double fib[3], fdm[3];
      muscle(f, m, targets ? targets[e] : 1.0, fib);
      m_vec(fd, m, fdm);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) st[3 * i + j] = st[3 * i + j] + p->mw[e] * ((fib[i] - fdm[i]) * m[j]);
    }
    const double* g = sample_grads(p, e, s);
    for (int k = 0; k < K; ++k) m_vec(st, g + 3 * k, contrib + 3 * ((size_t)sample * K + k));
  }
  return status;
}

static void gather(const oracle_problem* p, const double* contrib, int accumulate, double* out) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < p->n; ++i) {
    double t[3] = {0.0, 0.0, 0.0};
    for (int a = p->inc_off[i]; a < p->inc_off[i + 1]; ++a) {
      const double* c = contrib + 3 * (size_t)p->inc_ent[a];
      t[0] = t[0] + c[0];
      t[1] = t[1] + c[1];
      t[2] = t[2] + c[2];
    }
    if (accumulate)
      for (int d = 0; d < 3; ++d) t[d] = t[d] + out[3 * i + d];
    for (int d = 0; d < 3; ++d) out[3 * i + d] = t[d];
  }
}

/* planar_projection (hex.py:246-260), after the element gather. */
static void planar(const oracle_problem* p, const double* x, const double* bnd, double* proj) {
  const double* nrm = p->plane;
  for (int j = 0; j < p->planar_count; ++j) {
    const int i = p->planar_idx[j];
    const double* xi = x + 3 * (size_t)i;
    const double d = v_dot(nrm, xi) + nrm[3];
    const double active = d < 0.0 ? d : 0.0;
    for (int c = 0; c < 3; ++c) {
      const double projected = xi[c] - active * nrm[c];
      proj[3 * (size_t)i + c] = proj[3 * (size_t)i + c] + p->planar_w[j] * (projected - bnd[3 * (size_t)i + c]);
    }
  }
}

static void spmv(int n, const int* off, const int* col, const double* val, const double* in, double* out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int r = 0; r < n; ++r) {
    double t[3] = {0.0, 0.0, 0.0};
    for (int q = off[r]; q < off[r + 1]; ++q) {
      const double* x = in + 3 * (size_t)col[q];
      t[0] = t[0] + val[q] * x[0];
      t[1] = t[1] + val[q] * x[1];
      t[2] = t[2] + val[q] * x[2];
    }
    out[3 * r] = t[0];
    out[3 * r + 1] = t[1];
    out[3 * r + 2] = t[2];
  }
}

void oracle_apply_inverse(const oracle_problem* p, const double* rhs, double* out) {
  double* work = (double*)malloc(sizeof(double) * 3 * (size_t)p->n);
  spmv(p->n, p->s_off, p->s_col, p->s_val, rhs, work);
  spmv(p->n, p->t_off, p->t_col, p->t_val, work, out);
  free(work);
}

/* ------------------------------------------------------------------------ */
/* Type-II Anderson mixing with a ring history (anderson.py).               */

typedef struct {
  int n, window, cap, chunks;
  int count, start, saved_count, saved_start, accept;
  double *hx, *hg, *partials, *reduced, gamma[8];
} anderson;

static void aa_init(anderson* a, int n, int window) {
  memset(a, 0, sizeof(*a));
  a->n = n;
  a->window = window;
  a->cap = window + 1;
  a->chunks = n > 0 ? (n + CHUNK - 1) / CHUNK : 1;
  if (window) {
    a->hx = (double*)calloc((size_t)a->cap * n * 3, sizeof(double));
    a->hg = (double*)calloc((size_t)a->cap * n * 3, sizeof(double));
    a->partials = (double*)calloc((size_t)a->chunks * window * (window + 1), sizeof(double));
    a->reduced = (double*)calloc((size_t)window * (window + 1), sizeof(double));
  }
}
static void aa_free(anderson* a) {
  free(a->hx);
  free(a->hg);
  free(a->partials);
  free(a->reduced);
}
static void aa_reset(anderson* a) { a->count = a->start = a->saved_count = a->saved_start = a->accept = 0; }

static void aa_mix(anderson* a, double* x, const double* prev) {
  const int W = a->window, cap = a->cap, n = a->n;
  if (!W) return;
  const int slot = (a->start + a->count) % cap;
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      a->hx[((size_t)slot * n + i) * 3 + d] = prev[3 * i + d];
      a->hg[((size_t)slot * n + i) * 3 + d] = x[3 * i + d] - prev[3 * i + d];
    }
  const int entries = W * (W + 1), count = a->count, start = a->start;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < a->chunks; ++c) {
    const int lo = c * CHUNK, hi = (c + 1) * CHUNK < n ? (c + 1) * CHUNK : n;
    for (int entry = 0; entry < entries; ++entry) {
      double total = 0.0;
      if (entry >= W * W) {
        const int ia = entry - W * W;
        if (ia < count) {
          const int s_new = (start + count) % cap, s1 = (start + ia + 1) % cap, s0 = (start + ia) % cap;
          for (int i = lo; i < hi; ++i) {
            const double* g1 = a->hg + ((size_t)s1 * n + i) * 3;
            const double* g0 = a->hg + ((size_t)s0 * n + i) * 3;
            const double* gn = a->hg + ((size_t)s_new * n + i) * 3;
            double da[3] = {g1[0] - g0[0], g1[1] - g0[1], g1[2] - g0[2]};
            total += v_dot(da, gn);
          }
        }
      } else {
        const int ia = entry / W, ib = entry % W;
        if (ia <= ib && ib < count) {
          const int a1 = (start + ia + 1) % cap, a0 = (start + ia) % cap;
          const int b1 = (start + ib + 1) % cap, b0 = (start + ib) % cap;
          for (int i = lo; i < hi; ++i) {
            const double* ga1 = a->hg + ((size_t)a1 * n + i) * 3;
            const double* ga0 = a->hg + ((size_t)a0 * n + i) * 3;
            const double* gb1 = a->hg + ((size_t)b1 * n + i) * 3;
            const double* gb0 = a->hg + ((size_t)b0 * n + i) * 3;
            double da[3] = {ga1[0] - ga0[0], ga1[1] - ga0[1], ga1[2] - ga0[2]};
            double db[3] = {gb1[0] - gb0[0], gb1[1] - gb0[1], gb1[2] - gb0[2]};
            total += v_dot(da, db);
          }
        }
      }
      a->partials[(size_t)c * entries + entry] = total;
    }
  }
  for (int entry = 0; entry < entries; ++entry) {
    double total = 0.0;
    for (int c = 0; c < a->chunks; ++c) total += a->partials[(size_t)c * entries + entry];
    a->reduced[entry] = total;
  }
  /* _solve */
  int cnt = a->count, st = a->start;
  a->saved_count = cnt;
  a->saved_start = st;
  a->accept = 0;
  if (cnt >= 1) {
    double gram[64], rhs[8], coef[8];
    for (int i = 0; i < cnt; ++i) {
      rhs[i] = a->reduced[W * W + i];
      for (int j = i; j < cnt; ++j) gram[i * 8 + j] = a->reduced[i * W + j];
    }
    double trace = 0.0;
    for (int i = 0; i < cnt; ++i) {
      trace += gram[i * 8 + i];
      for (int j = i + 1; j < cnt; ++j) gram[j * 8 + i] = gram[i * 8 + j];
    }
    double reg = fmax(1.0e-6 * trace / (double)cnt, 1.0e-12);
    for (int i = 0; i < cnt; ++i) gram[i * 8 + i] = gram[i * 8 + i] + reg;
    ldlt_small(gram, 8, rhs, cnt, coef);
    double norm = 0.0;
    for (int i = 0; i < cnt; ++i) {
      a->gamma[i] = coef[i];
      norm += coef[i] * coef[i];
    }
    if (sqrt(norm) >= 10.0) {
      cnt = 0;
      st = 0;
    } else {
      a->accept = 1;
      cnt += 1;
    }
  } else {
    cnt += 1;
  }
  while (cnt > W) {
    st = (st + 1) % cap;
    cnt -= 1;
  }
  a->count = cnt;
  a->start = st;
  /* _apply */
  if (a->accept) {
    const int sc = a->saved_count, ss = a->saved_start, sl = (ss + sc) % cap;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      double v[3];
      for (int d = 0; d < 3; ++d) v[d] = a->hx[((size_t)sl * n + i) * 3 + d] + a->hg[((size_t)sl * n + i) * 3 + d];
      for (int k = 0; k < sc; ++k) {
        const int lo = (ss + k) % cap, hi = (ss + k + 1) % cap;
        for (int d = 0; d < 3; ++d) {
          double diff = a->hx[((size_t)hi * n + i) * 3 + d] - a->hx[((size_t)lo * n + i) * 3 + d] +
                        a->hg[((size_t)hi * n + i) * 3 + d] - a->hg[((size_t)lo * n + i) * 3 + d];
          v[d] = v[d] - diff * a->gamma[k];
        }
      }
      for (int d = 0; d < 3; ++d) x[3 * i + d] = v[d];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Forward step (solver.py:689-786 on the CPU path).                        */

static void set_threads(const oracle_problem* p) {
#ifdef _OPENMP
  if (p->threads > 0) omp_set_num_threads(p->threads);
#else
  (void)p;
#endif
}

int oracle_forward(const oracle_problem* p, double dt, const float* q, const float* qd, const float* f,
                   const float* dirichlet, const double* targets, double* x_out, float* q_out, float* qd_out,
                   int* iterations, int* status) {
  set_threads(p);
  const int n = p->n;
  const size_t vb = sizeof(double) * 3 * (size_t)n;
  double *pos = malloc(vb), *tgt = malloc(vb), *bnd = malloc(vb), *x = malloc(vb), *prev = malloc(vb);
  double *proj = malloc(vb), *pprev = malloc(vb), *rhs = malloc(vb), *work = malloc(vb);
  double* contrib = malloc(sizeof(double) * 3 * (size_t)p->elements * p->samples * p->corners);
  const double inertia = p->mass / (dt * dt);
  for (int i = 0; i < n; ++i) {
    for (int d = 0; d < 3; ++d) {
      const double start = (double)q[3 * i + d], vel = (double)qd[3 * i + d];
      const double acc = (double)f[3 * i + d] / p->mass + p->gravity[d];
      pos[3 * i + d] = start;
      tgt[3 * i + d] = start + dt * vel + dt * dt * acc;
      bnd[3 * i + d] = 0.0;
      if (!p->free[i]) bnd[3 * i + d] = p->integrate_pinned ? start + dt * vel : start;
      if (!p->free[i] && dirichlet) bnd[3 * i + d] = (double)dirichlet[3 * i + d];
      x[3 * i + d] = start;
      proj[3 * i + d] = 0.0;
    }
  }
  if (p->state_force)
    for (int t = 0; t < 3 * n; ++t) tgt[t] = tgt[t] + dt * dt / p->mass * p->state_force[t];
  contact_work cw;
  double* vel = NULL;
  double* solved = NULL;
  if (p->contact) {
    contact_prepare(p, &cw);
    contact_begin(p, &cw, pos, dt);
    vel = malloc(vb);
    solved = malloc(vb);
    for (int t = 0; t < 3 * n; ++t) vel[t] = (double)qd[t];
  }
  anderson aa;
  aa_init(&aa, n, p->window);
  aa_reset(&aa);
  int st = 0, iteration = 0, active = 1;
  const int chunks = n > 0 ? (n + CHUNK - 1) / CHUNK : 1;
  double* part = malloc(sizeof(double) * 4 * chunks);
  for (int it = 0; it < p->max_iterations && active; ++it) {
    memcpy(prev, x, vb);
    memcpy(pprev, proj, vb);
    st |= local_step(p, x, bnd, targets, contrib);
    gather(p, contrib, 0, proj);
    planar(p, x, bnd, proj);
    for (int t = 0; t < 3 * n; ++t) rhs[t] = inertia * tgt[t] + proj[t];
    spmv(n, p->s_off, p->s_col, p->s_val, rhs, work);
    if (p->contact) {
      /* x = A^{-1}(b + J^T lambda) after one Newton step on the multipliers (solver.py:628-630). */
      spmv(n, p->t_off, p->t_col, p->t_val, work, solved);
      contact_iterate(p, &cw, x, solved, pos, x);
    } else {
      spmv(n, p->t_off, p->t_col, p->t_val, work, x);
    }
    for (int i = 0; i < n; ++i)
      if (!p->free[i])
        for (int d = 0; d < 3; ++d) x[3 * i + d] = bnd[3 * i + d];
    aa_mix(&aa, x, prev);
    for (int c = 0; c < chunks; ++c) {
      double tot[4] = {0, 0, 0, 0};
      const int hi = (c + 1) * CHUNK < n ? (c + 1) * CHUNK : n;
      for (int i = c * CHUNK; i < hi; ++i) {
        if (!p->free[i])
          for (int d = 0; d < 3; ++d) x[3 * i + d] = bnd[3 * i + d];
        double s[3], ch[3];
        for (int d = 0; d < 3; ++d) {
          s[d] = x[3 * i + d] - prev[3 * i + d];
          ch[d] = proj[3 * i + d] - pprev[3 * i + d];
        }
        tot[0] += v_dot(s, s);
        tot[1] += v_dot(prev + 3 * i, prev + 3 * i);
        tot[2] += v_dot(ch, ch);
        tot[3] += v_dot(pprev + 3 * i, pprev + 3 * i);
      }
      memcpy(part + 4 * c, tot, sizeof(tot));
    }
    double total[4] = {0, 0, 0, 0};
    for (int c = 0; c < chunks; ++c)
      for (int k = 0; k < 4; ++k) total[k] += part[4 * c + k];
    const int k = iteration;
    const int step_small = sqrt(total[0]) <= p->rel_tol * sqrt(total[1]) + p->abs_tol;
    const int rhs_small = sqrt(total[2]) <= p->rel_tol * sqrt(total[3]) + p->abs_tol;
    iteration = k + 1;
    if (k > 0 && step_small && rhs_small) {
      active = 0;
    } else if (k + 1 >= p->max_iterations) {
      active = 0;
      st |= 1;
    }
  }
  if (p->contact) {
    contact_finish(p, &cw, pos, vel, dt, x);
    contact_free(&cw);
    free(vel);
    free(solved);
  }
  for (int t = 0; t < 3 * n; ++t) {
    if (x_out) x_out[t] = x[t];
    if (q_out) q_out[t] = (float)x[t];
    if (qd_out) qd_out[t] = (float)((x[t] - pos[t]) / dt);
  }
  *iterations = iteration;
  *status = st;
  aa_free(&aa);
  free(pos);
  free(tgt);
  free(bnd);
  free(x);
  free(prev);
  free(proj);
  free(pprev);
  free(rhs);
  free(work);
  free(contrib);
  free(part);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Backward pass (adjoint.py).                                              */

/* Safe clip of an adjoint vector (adjoint.py:187-225, _clip_partials and
 * _clip_apply): Dirichlet rows zeroed, non-finite components replaced by 0
 * (status 16), then, when limit > 0 and the norm is finite and above it, the
 * whole vector scaled to the limit (status 16). The squared norm is summed in
 * CHUNK-row partials, in chunk order. */
static int clip_vector(int n, const unsigned char* free_rows, double* v, int chunks, double* part, double limit) {
  int st = 0;
  for (int c = 0; c < chunks; ++c) {
    double tot = 0.0;
    const int hi = (c + 1) * CHUNK < n ? (c + 1) * CHUNK : n;
    for (int i = c * CHUNK; i < hi; ++i) {
      double* x = v + 3 * i;
      if (!free_rows[i]) x[0] = x[1] = x[2] = 0.0;
      for (int d = 0; d < 3; ++d)
        if (!isfinite(x[d])) {
          x[d] = 0.0;
          st |= 16;
        }
      tot += v_dot(x, x);
    }
    part[c] = tot;
  }
  double total = 0.0;
  for (int c = 0; c < chunks; ++c) total += part[c];
  const double len = sqrt(total);
  if (limit > 0.0 && isfinite(len) && len > limit) {
    for (int t = 0; t < 3 * n; ++t) v[t] = (limit / len) * v[t];
    st |= 16;
  }
  return st;
}

static int safe_clip(const oracle_problem* p, double* v, int chunks, double* part) {
  return clip_vector(p->n, p->free, v, chunks, part, CLIP_NORM);
}

/* Routing guard of a finished adjoint solve (adjoint.py:226-261) from the
 * squared norms of its right-hand side and solution: diverged when
 * |mu| >= 0.999 clip_norm (scale 0), else scale = min(1, (|g| + 1) / (m/h^2 |mu|));
 * either intervention sets status 16. */
static int routing_guard(double rhs_sq, double mu_sq, double inertia, double clip_norm, double* scale,
                         int* diverged) {
  const double mu_norm = sqrt(mu_sq);
  *diverged = clip_norm > 0.0 && mu_norm >= 0.999 * clip_norm;
  *scale = 1.0;
  if (*diverged) {
    *scale = 0.0;
    return 16;
  }
  const double cap = sqrt(rhs_sq) + 1.0, inj = inertia * mu_norm;
  if (inj > cap) {
    *scale = cap / inj;
    return 16;
  }
  return 0;
}

int oracle_safe_clip(int n, const unsigned char* free_rows, double* v, double limit) {
  const int chunks = n > 0 ? (n + CHUNK - 1) / CHUNK : 1;
  double* part = malloc(sizeof(double) * chunks);
  const int st = clip_vector(n, free_rows, v, chunks, part, limit);
  free(part);
  return st;
}

int oracle_routing_guard(double rhs_sq, double mu_sq, double inertia, double clip_norm, double* scale,
                         int* diverged) {
  return routing_guard(rhs_sq, mu_sq, inertia, clip_norm, scale, diverged);
}

/* Fixed point (A - K) sol = rhs from a zero start with at most `limit`
 * iterations (adjoint.py:927-964, 1054-1086): sol <- A^{-1}(rhs + K sol),
 * safe clips, Dirichlet rows zeroed, Anderson, gate
 * ||raw - prev|| < rel ||raw|| + abs. Returns the status bits. */
static int fixed_point(const oracle_problem* p, const double* jac, const double* rhs, double* sol, int limit,
                       int report, int* iterations) {
  const int n = p->n, K = p->corners, Q = p->samples, E = p->elements;
  const size_t vb = sizeof(double) * 3 * (size_t)n;
  double *prev = malloc(vb), *work = malloc(vb), *raw = malloc(vb), *tmp = malloc(vb);
  double* contrib = malloc(sizeof(double) * 3 * (size_t)E * Q * K);
  const int chunks = n > 0 ? (n + CHUNK - 1) / CHUNK : 1;
  double* part = malloc(sizeof(double) * 2 * chunks);
  double* cpart = malloc(sizeof(double) * chunks);
  int st = 0;
  memset(sol, 0, vb);
  anderson aa;
  aa_init(&aa, n, p->window);
  aa_reset(&aa);
  int iteration = 0, active = 1;
  for (int it = 0; it < limit && active; ++it) {
    memcpy(prev, sol, vb);
    memcpy(work, rhs, vb);
#pragma omp parallel for schedule(static)
    for (int sample = 0; sample < E * Q; ++sample) {
      const int e = sample / Q, s = sample % Q;
      double df[9], ds[9] = {0};
      deformation(p, sol, e, s, df);
      const double* J = jac + (size_t)sample * 81;
      for (int c = 0; c < 9; ++c) {
        const double w = df[3 * (c % 3) + c / 3];
        for (int t = 0; t < 9; ++t) ds[t] = ds[t] + w * J[c * 9 + t];
      }
      const double* g = sample_grads(p, e, s);
      for (int k = 0; k < K; ++k) m_vec(ds, g + 3 * k, contrib + 3 * ((size_t)sample * K + k));
    }
    gather(p, contrib, 1, work);
    spmv(n, p->s_off, p->s_col, p->s_val, work, tmp);
    spmv(n, p->t_off, p->t_col, p->t_val, tmp, raw);
    st |= safe_clip(p, raw, chunks, cpart);
    for (int c = 0; c < chunks; ++c) {
      double tot = 0.0;
      const int hi = (c + 1) * CHUNK < n ? (c + 1) * CHUNK : n;
      for (int i = c * CHUNK; i < hi; ++i) {
        if (!p->free[i]) raw[3 * i] = raw[3 * i + 1] = raw[3 * i + 2] = 0.0;
        double s[3] = {raw[3 * i] - prev[3 * i], raw[3 * i + 1] - prev[3 * i + 1], raw[3 * i + 2] - prev[3 * i + 2]};
        tot += v_dot(s, s);
      }
      part[2 * c] = tot;
      part[2 * c + 1] = 0.0;
    }
    aa_mix(&aa, raw, prev);
    st |= safe_clip(p, raw, chunks, cpart);
    for (int c = 0; c < chunks; ++c) {
      double tot = 0.0;
      const int hi = (c + 1) * CHUNK < n ? (c + 1) * CHUNK : n;
      for (int i = c * CHUNK; i < hi; ++i) {
        for (int d = 0; d < 3; ++d) sol[3 * i + d] = raw[3 * i + d];
        tot += v_dot(raw + 3 * i, raw + 3 * i);
      }
      part[2 * c + 1] = tot;
    }
    double t0 = 0.0, t1 = 0.0;
    for (int c = 0; c < chunks; ++c) {
      t0 += part[2 * c];
      t1 += part[2 * c + 1];
    }
    const int k = iteration;
    iteration = k + 1;
    if (k > 0 && sqrt(t0) < p->rel_tol * sqrt(t1) + p->abs_tol) {
      active = 0;
    } else if (k + 1 >= limit) {
      active = 0;
      /* Woodbury columns are truncated by design; only the backbone reports. */
      if (report) st |= 8;
    }
  }
  *iterations = iteration;
  aa_free(&aa);
  free(prev);
  free(work);
  free(raw);
  free(tmp);
  free(contrib);
  free(part);
  free(cpart);
  return st;
}

/* Woodbury projection of the backbone adjoint off the active contact rows
 * (Adjoint._woodbury, adjoint.py:324-471, 663-806, 1272-1354): rows omega n
 * of contacts with lambda > 0 and omega >= 0.5 (and, with friction, two
 * tangent rows where lambda > 1e-8); rows with |J mu| below
 * max(1e-12, 1e-3 max |J mu|) dropped; columns W_T = (A - K)^{-1} J^T by
 * fixed points of at most min(max_iterations, 120) iterations; then
 * z = mu - W_T G^{-1} (J mu) with G = J W_T (lower triangle, Eigen LDLT). */
static int woodbury(const oracle_problem* p, const double* jac, const double* mu, int diverged, double* z) {
  const oracle_contact* c = p->contact;
  const int K = c->count, n = p->n;
  const int tangents = c->friction > 0.0;
  const int R = tangents ? 3 * K : K;
  const size_t vb = sizeof(double) * 3 * (size_t)n;
  int* row_node = calloc(R, sizeof(int));
  int* row_valid = calloc(R, sizeof(int));
  int* order = calloc(R, sizeof(int));
  double* row_vec = calloc(3 * (size_t)R, sizeof(double));
  double* value = calloc(R, sizeof(double));
  for (int k = 0; k < K; ++k) {
    const int node = c->nodes[k];
    const double* nk = c->normal + 3 * (size_t)k;
    const int active = !diverged && c->multiplier[k] > 0.0 && c->omega[k] >= 0.5;
    row_node[k] = node;
    for (int d = 0; d < 3; ++d) row_vec[3 * k + d] = c->omega[k] * nk[d];
    row_valid[k] = active;
    if (tangents) {
      int axis = 0;
      for (int d = 1; d < 3; ++d)
        if (fabs(nk[d]) < fabs(nk[axis])) axis = d;
      double unit[3] = {0.0, 0.0, 0.0}, t1[3], t2[3];
      unit[axis] = 1.0;
      const double nu = v_dot(nk, unit);
      for (int d = 0; d < 3; ++d) t1[d] = unit[d] - nu * nk[d];
      double length = sqrt(v_dot(t1, t1));
      if (length > 1.0e-12) {
        for (int d = 0; d < 3; ++d) t1[d] = t1[d] / length;
      } else {
        t1[0] = 1.0;
        t1[1] = t1[2] = 0.0;
      }
      t2[0] = nk[1] * t1[2] - nk[2] * t1[1];
      t2[1] = nk[2] * t1[0] - nk[0] * t1[2];
      t2[2] = nk[0] * t1[1] - nk[1] * t1[0];
      length = sqrt(v_dot(t2, t2));
      if (length > 1.0e-12) {
        for (int d = 0; d < 3; ++d) t2[d] = t2[d] / length;
      } else {
        t2[1] = 1.0;
        t2[0] = t2[2] = 0.0;
      }
      const int valid = active && c->multiplier[k] > 1.0e-8;
      for (int r = 0; r < 2; ++r) {
        const int row = K + 2 * k + r;
        row_node[row] = node;
        memcpy(row_vec + 3 * row, r == 0 ? t1 : t2, sizeof(t1));
        row_valid[row] = valid;
      }
    }
  }
  double largest = 0.0;
  for (int r = 0; r < R; ++r) {
    value[r] = row_valid[r] ? v_dot(row_vec + 3 * r, mu + 3 * (size_t)row_node[r]) : 0.0;
    if (row_valid[r] && fabs(value[r]) > largest) largest = fabs(value[r]);
  }
  const double gate = 1.0e-3 * largest > 1.0e-12 ? 1.0e-3 * largest : 1.0e-12;
  int count = 0;
  for (int r = 0; r < R; ++r)
    if (row_valid[r] && fabs(value[r]) >= gate) order[count++] = r;
  int st = 0;
  memcpy(z, mu, vb);
  if (count > 0) {
    const int limit = p->max_iterations < 120 ? p->max_iterations : 120;
    double* columns = calloc((size_t)count * 3 * n, sizeof(double));
    double* rhs = calloc(3 * (size_t)n, sizeof(double));
    for (int k = 0; k < count; ++k) {
      const int r = order[k];
      memset(rhs, 0, vb);
      memcpy(rhs + 3 * (size_t)row_node[r], row_vec + 3 * r, sizeof(double) * 3);
      int iterations;
      st |= fixed_point(p, jac, rhs, columns + (size_t)k * 3 * n, limit, 0, &iterations);
    }
    double* gram = calloc((size_t)count * count, sizeof(double));
    double* coef = calloc(count, sizeof(double));
    for (int i = 0; i < count; ++i) {
      for (int j = 0; j < count; ++j) {
        const int hi = i > j ? i : j, lo = i > j ? j : i;
        const int rh = order[hi];
        gram[(size_t)i * count + j] =
            v_dot(row_vec + 3 * rh, columns + (size_t)lo * 3 * n + 3 * (size_t)row_node[rh]);
      }
      coef[i] = value[order[i]];
    }
    oracle_ldlt_solve(gram, count, coef);
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < count; ++k)
        for (int d = 0; d < 3; ++d)
          z[3 * (size_t)i + d] = z[3 * (size_t)i + d] - coef[k] * columns[(size_t)k * 3 * n + 3 * (size_t)i + d];
    free(columns);
    free(rhs);
    free(gram);
    free(coef);
  }
  const int chunks = n > 0 ? (n + CHUNK - 1) / CHUNK : 1;
  double* cpart = malloc(sizeof(double) * chunks);
  st |= safe_clip(p, z, chunks, cpart);
  free(cpart);
  free(row_node);
  free(row_valid);
  free(order);
  free(row_vec);
  free(value);
  return st;
}

int oracle_backward(const oracle_problem* p, double dt, const double* x, const double* targets, const float* gq_out,
                    const float* gqd_out, float* gq_in, float* gqd_in, float* gf_in, double* gmaterials,
                    float* gdirichlet, double* gactivations, double* mu_out, int* iterations, int* status) {
  /* The reference records no adjoint with planar energies (Hex8 only). */
  if (p->planar_count > 0) return 1;
  set_threads(p);
  const int n = p->n, K = p->corners, Q = p->samples, E = p->elements;
  const size_t vb = sizeof(double) * 3 * (size_t)n;
  const double inertia = p->mass / (dt * dt);
  double *rhs = malloc(vb), *base = malloc(vb), *sol = malloc(vb), *prev = malloc(vb), *work = malloc(vb);
  double *raw = malloc(vb), *tmp = malloc(vb), *coupling = malloc(vb);
  double* contrib = malloc(sizeof(double) * 3 * (size_t)E * Q * K);
  double* jac = malloc(sizeof(double) * 81 * (size_t)E * Q);
  const int chunks = n > 0 ? (n + CHUNK - 1) / CHUNK : 1;
  double* part = malloc(sizeof(double) * 2 * chunks);
  double* cpart = malloc(sizeof(double) * chunks);
  int st = 0;
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      const double vel = (gqd_out ? (double)gqd_out[3 * i + d] : 0.0) / dt;
      rhs[3 * i + d] = (gq_out ? (double)gq_out[3 * i + d] : 0.0) + vel;
      base[3 * i + d] = -vel;
      sol[3 * i + d] = 0.0;
    }
  /* Local Jacobian at x (adjoint.py:92-137). */
#pragma omp parallel for schedule(static)
  for (int sample = 0; sample < E * Q; ++sample) {
    const int e = sample / Q, s = sample % Q;
    double f[9], u[9], sg[3], v[9], off[3] = {0, 0, 0}, str[3] = {0, 0, 0};
    deformation(p, x, e, s, f);
    svd3(f, u, sg, v);
    if (p->material == 0)
      volume_offset(sg, off);
    else
      nh_stretches(sg, p->nh_mu, p->nh_lame, p->nh_penalty, str);
    const double wm = p->mw ? p->mw[e] : 0.0;
    const double* m = p->mdir ? p->mdir + 3 * (size_t)e : NULL;
    for (int c = 0; c < 9; ++c) {
      double df[9] = {0}, dp[9] = {0}, t[9];
      df[3 * (c % 3) + c / 3] = 1.0;
      if (p->material == 0) {
        if (p->jwc[e] != 0.0) {
          corot_jvp(u, sg, v, df, t);
          for (int k = 0; k < 9; ++k) dp[k] = dp[k] + p->jwc[e] * t[k];
        }
        if (p->jwv[e] != 0.0) {
          vol_jvp(u, sg, v, off, df, t);
          for (int k = 0; k < 9; ++k) dp[k] = dp[k] + p->jwv[e] * t[k];
        }
      } else {
        nh_jvp(u, sg, v, str, df, p->nh_mu, p->nh_lame, p->nh_penalty, t);
        for (int k = 0; k < 9; ++k) dp[k] = dp[k] + p->jwn[e] * t[k];
      }
      if (wm != 0.0) {
        double mj[3];
        muscle_jvp(f, m, targets ? targets[e] : 1.0, df, mj);
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) dp[3 * i + j] = dp[3 * i + j] + wm * (mj[i] * m[j]);
      }
      memcpy(jac + ((size_t)sample * 9 + c) * 9, dp, sizeof(dp));
    }
  }
  /* Fixed point (A - K) mu = rhs. */
  int iteration = 0;
  st |= fixed_point(p, jac, rhs, sol, p->max_iterations, 1, &iteration);
  /* Routing guard (adjoint.py:226-261). */
  double r0 = 0.0, r1 = 0.0;
  for (int c = 0; c < chunks; ++c) {
    double a0 = 0.0, a1 = 0.0;
    const int hi = (c + 1) * CHUNK < n ? (c + 1) * CHUNK : n;
    for (int i = c * CHUNK; i < hi; ++i) {
      a0 += v_dot(rhs + 3 * i, rhs + 3 * i);
      a1 += v_dot(sol + 3 * i, sol + 3 * i);
    }
    r0 += a0;
    r1 += a1;
  }
  double scale = 1.0;
  int diverged = 0;
  st |= routing_guard(r0, r1, inertia, CLIP_NORM, &scale, &diverged);
  /* Parameter gradients use z, the backbone projected off the active
   * contact rows; state gradients use mu (adjoint.py:1145-1148). */
  double* z = sol;
  double* zbuf = NULL;
  if ((gmaterials || gactivations) && p->contact && p->contact->count > 0) {
    zbuf = malloc(vb);
    st |= woodbury(p, jac, sol, diverged, zbuf);
    z = zbuf;
  }
  /* State gradients from mu (adjoint.py:474-490). */
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      const double inj = scale * inertia * sol[3 * i + d];
      gq_in[3 * i + d] = gq_in[3 * i + d] + (float)(base[3 * i + d] + inj);
      gqd_in[3 * i + d] = gqd_in[3 * i + d] + (float)(dt * inj);
      gf_in[3 * i + d] = gf_in[3 * i + d] + (float)(scale * sol[3 * i + d]);
    }
  /* Dirichlet values (adjoint.py:493-553). */
  int any_pinned = 0;
  for (int i = 0; i < n; ++i) any_pinned |= !p->free[i];
  if (any_pinned) {
    for (int sample = 0; sample < E * Q; ++sample) {
      const int e = sample / Q, s = sample % Q;
      double* out = contrib + 3 * (size_t)sample * K;
      if (diverged) {
        memset(out, 0, sizeof(double) * 3 * K);
        continue;
      }
      double df[9], ds[9];
      deformation(p, sol, e, s, df);
      const double wsum = p->wc[e] + p->wv[e] + p->wn[e];
      for (int t = 0; t < 9; ++t) ds[t] = wsum * df[t];
      if (p->mw && p->mw[e] != 0.0) {
        const double* m = p->mdir + 3 * (size_t)e;
        double dm[3];
        m_vec(df, m, dm);
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) ds[3 * i + j] = ds[3 * i + j] + p->mw[e] * (dm[i] * m[j]);
      }
      const double* J = jac + (size_t)sample * 81;
      for (int c = 0; c < 9; ++c) {
        const double w = df[3 * (c % 3) + c / 3];
        for (int t = 0; t < 9; ++t) ds[t] = ds[t] - w * J[c * 9 + t];
      }
      const double* g = sample_grads(p, e, s);
      for (int k = 0; k < K; ++k) m_vec(ds, g + 3 * k, out + 3 * k);
    }
    gather(p, contrib, 0, coupling);
    for (int i = 0; i < n; ++i) {
      if (p->free[i]) continue;
      for (int d = 0; d < 3; ++d) {
        const double value = rhs[3 * i + d] - coupling[3 * i + d];
        if (gdirichlet) {
          gdirichlet[3 * i + d] = gdirichlet[3 * i + d] + (float)value;
        } else {
          gq_in[3 * i + d] = gq_in[3 * i + d] + (float)value;
          if (p->integrate_pinned) gqd_in[3 * i + d] = gqd_in[3 * i + d] + (float)(dt * value);
        }
      }
    }
  }
  /* Material and activation gradients (adjoint.py:556-620) and their
   * aggregate guard (adjoint.py:622-659). */
  if (gmaterials || gactivations) {
    double* g = calloc((size_t)E * 2, sizeof(double));
    double* ga = calloc((size_t)E, sizeof(double));
    const double svol = p->sample_volume;
    for (int e = 0; gmaterials && e < E; ++e) {
      double first_total = 0.0, second_total = 0.0;
      for (int s = 0; s < Q; ++s) {
        double f[9], first[9], second[9] = {0};
        deformation(p, x, e, s, f);
        if (p->material == 0) {
          double u[9], sg[3], v[9], vt[9], r[9], off[3], sd[3], pv[9];
          svd3(f, u, sg, v);
          volume_offset(sg, off);
          m_t(v, vt);
          m_mul(u, vt, r);
          for (int i = 0; i < 3; ++i) sd[i] = sg[i] + off[i];
          m_udvt(u, sd, v, pv);
          for (int t = 0; t < 9; ++t) {
            first[t] = f[t] - r[t];
            second[t] = f[t] - pv[t];
          }
        } else {
          double fi[9], it[9];
          const double lj = log(fmax(det3(f), 1.0e-12));
          inv3(f, fi);
          m_t(fi, it);
          for (int t = 0; t < 9; ++t)
            first[t] = (p->nh_mu * (f[t] - it[t]) + p->nh_lame * lj * it[t]) / (2.0 * p->nh_mu);
        }
        const double* gr = sample_grads(p, e, s);
        double fv = 0.0, sv = 0.0;
        for (int k = 0; k < K; ++k) {
          const double* zk = z + 3 * (size_t)p->indices[(size_t)e * K + k];
          double a[3], b[3];
          m_vec(first, gr + 3 * k, a);
          m_vec(second, gr + 3 * k, b);
          fv -= v_dot(zk, a);
          sv -= v_dot(zk, b);
        }
        first_total += fv;
        second_total += sv;
      }
      /* dL/dmu = 2 V dL/dw_c and dL/dlambda = V dL/dw_v with V the sample
       * volume (Vbar for Tet4, adjoint.py:588-591). */
      g[2 * e] = 2.0 * svol * first_total;
      g[2 * e + 1] = p->material == 0 ? svol * second_total : 0.0;
    }
    /* d/du of the fibre energy with target length a = 1 - u (Tet4 only):
     * -sum_k z_k . (w_m / |F m|) (m . g_k) F m, zero for a collapsed fibre. */
    for (int e = 0; gactivations && p->mw && Q == 1 && e < E; ++e) {
      const double wm = p->mw[e];
      if (wm == 0.0) continue;
      const double* m = p->mdir + 3 * (size_t)e;
      double f[9], fibre[3];
      deformation(p, x, e, 0, f);
      m_vec(f, m, fibre);
      const double length = sqrt(v_dot(fibre, fibre));
      if (length <= EPS) continue;
      const double* gr = sample_grads(p, e, 0);
      double value = 0.0;
      for (int k = 0; k < K; ++k) {
        const double* zk = z + 3 * (size_t)p->indices[(size_t)e * K + k];
        const double c = wm / length * v_dot(m, gr + 3 * k);
        const double t[3] = {c * fibre[0], c * fibre[1], c * fibre[2]};
        value += v_dot(zk, t);
      }
      ga[e] = -value;
    }
    double s0 = 0.0, s1 = 0.0;
    int activation_ok = 1;
    for (int e = 0; e < E; ++e) {
      s0 += g[2 * e];
      s1 += g[2 * e + 1];
      activation_ok = activation_ok && isfinite(ga[e]);
    }
    const int zero_energy = gmaterials && p->material == 0 && p->nh_lame == 0.0;
    const int material_ok = !zero_energy && isfinite(s0) && isfinite(s1);
    if (!material_ok || !activation_ok) st |= 16;
    if (gmaterials && !diverged && material_ok)
      for (int t = 0; t < 2 * E; ++t) gmaterials[t] += g[t];
    if (gactivations && !diverged && activation_ok)
      for (int e = 0; e < E; ++e) gactivations[e] += ga[e];
    free(g);
    free(ga);
  }
  if (mu_out) memcpy(mu_out, sol, vb);
  *iterations = iteration;
  *status = st;
  free(zbuf);
  free(rhs);
  free(base);
  free(sol);
  free(prev);
  free(work);
  free(raw);
  free(tmp);
  free(coupling);
  free(contrib);
  free(jac);
  free(part);
  free(cpart);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Pointwise evaluation for projection/JVP checks.                          */

void oracle_projections(int count, const double* f, const double* dirs, double mu, double lame, double penalty,
                        double fibre_target, double* sigma, double* recon, double* rotation, double* volume,
                        double* neo_hookean, double* fibre) {
  for (int i = 0; i < count; ++i) {
    const double* F = f + 9 * (size_t)i;
    double u[9], s[3], v[9], vt[9], d[3], sd[3], st[3];
    svd3(F, u, s, v);
    memcpy(sigma + 3 * i, s, sizeof(s));
    m_udvt(u, s, v, recon + 9 * i);
    m_t(v, vt);
    m_mul(u, vt, rotation + 9 * i);
    volume_offset(s, d);
    for (int k = 0; k < 3; ++k) sd[k] = s[k] + d[k];
    m_udvt(u, sd, v, volume + 9 * i);
    nh_stretches(s, mu, lame, penalty, st);
    m_udvt(u, st, v, neo_hookean + 9 * i);
    muscle(F, dirs + 3 * i, fibre_target, fibre + 3 * i);
  }
}

void oracle_jvps(int count, const double* f, const double* df, const double* dirs, double mu, double lame,
                 double fibre_target, double* rotation, double* volume, double* neo_hookean, double* fibre) {
  const double pen = 2.0 * mu;
  for (int i = 0; i < count; ++i) {
    const double* F = f + 9 * (size_t)i;
    const double* D = df + 9 * (size_t)i;
    double u[9], s[3], v[9], d[3], st[3];
    svd3(F, u, s, v);
    corot_jvp(u, s, v, D, rotation + 9 * i);
    volume_offset(s, d);
    vol_jvp(u, s, v, d, D, volume + 9 * i);
    nh_stretches(s, mu, lame, pen, st);
    nh_jvp(u, s, v, st, D, mu, lame, pen, neo_hookean + 9 * i);
    muscle_jvp(F, dirs + 3 * i, fibre_target, D, fibre + 3 * i);
  }
}

/* ------------------------------------------------------------------------ */
/* Mesh contact state force (mesh_contact.py).                              */

/* closest_point_triangle (mesh_contact.py:34-87). */
static void closest_point(const double* p, const double* a, const double* b, const double* c, double* point,
                          double* bary, double* dist2) {
  double ab[3], ac[3], ap[3], bp[3], cp[3], d[3];
  for (int k = 0; k < 3; ++k) {
    ab[k] = b[k] - a[k];
    ac[k] = c[k] - a[k];
    ap[k] = p[k] - a[k];
    bp[k] = p[k] - b[k];
    cp[k] = p[k] - c[k];
  }
  const double d1 = v_dot(ab, ap), d2 = v_dot(ac, ap);
  const double d3 = v_dot(ab, bp), d4 = v_dot(ac, bp);
  const double d5 = v_dot(ab, cp), d6 = v_dot(ac, cp);
  double u0 = 0, u1 = 0, u2 = 0;
  if (d1 <= 0.0 && d2 <= 0.0) {
    u0 = 1.0;
    memcpy(point, a, sizeof(double) * 3);
  } else if (d3 >= 0.0 && d4 <= d3) {
    u1 = 1.0;
    memcpy(point, b, sizeof(double) * 3);
  } else if (d1 * d4 - d3 * d2 <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
    const double w = d1 / (d1 - d3);
    u0 = 1.0 - w;
    u1 = w;
    for (int k = 0; k < 3; ++k) point[k] = a[k] + w * ab[k];
  } else if (d6 >= 0.0 && d5 <= d6) {
    u2 = 1.0;
    memcpy(point, c, sizeof(double) * 3);
  } else if (d5 * d2 - d1 * d6 <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
    const double w = d2 / (d2 - d6);
    u0 = 1.0 - w;
    u2 = w;
    for (int k = 0; k < 3; ++k) point[k] = a[k] + w * ac[k];
  } else if (d3 * d6 - d5 * d4 <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    u1 = 1.0 - w;
    u2 = w;
    for (int k = 0; k < 3; ++k) point[k] = b[k] + w * (c[k] - b[k]);
  } else {
    const double va = d3 * d6 - d5 * d4, vb = d5 * d2 - d1 * d6, vc = d1 * d4 - d3 * d2;
    const double denom = 1.0 / (va + vb + vc);
    const double v = vb * denom, w = vc * denom;
    u0 = (1.0 - v) - w;
    u1 = v;
    u2 = w;
    for (int k = 0; k < 3; ++k) point[k] = (a[k] + ab[k] * v) + ac[k] * w;
  }
  bary[0] = u0;
  bary[1] = u1;
  bary[2] = u2;
  for (int k = 0; k < 3; ++k) d[k] = p[k] - point[k];
  *dist2 = v_dot(d, d);
}

/* _pair_forces (mesh_contact.py:90-137); returns 1 for a pair in contact. */
static int pair_force(const double* q, const double* v, int vid, const int* face, double radius, double kn, double kf,
                      double mu, double* force, double* bary) {
  const double *p = q + 3 * (size_t)vid, *a = q + 3 * (size_t)face[0], *b = q + 3 * (size_t)face[1],
               *c = q + 3 * (size_t)face[2];
  double closest[3], dist2, n[3], d[3];
  closest_point(p, a, b, c, closest, bary, &dist2);
  if (dist2 >= radius * radius) return 0;
  const double dist = sqrt(dist2);
  for (int k = 0; k < 3; ++k) d[k] = p[k] - closest[k];
  if (dist > EPS) {
    for (int k = 0; k < 3; ++k) n[k] = d[k] / dist;
  } else {
    double e1[3], e2[3];
    for (int k = 0; k < 3; ++k) {
      e1[k] = b[k] - a[k];
      e2[k] = c[k] - a[k];
    }
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    const double nn = sqrt(v_dot(n, n));
    if (nn <= EPS) return 0;
    for (int k = 0; k < 3; ++k) n[k] = n[k] / nn;
  }
  if (v_dot(n, d) < 0.0)
    for (int k = 0; k < 3; ++k) n[k] = -n[k];
  const double fn_mag = kn * (radius - dist);
  double vrel[3], vt[3];
  for (int k = 0; k < 3; ++k) {
    const double vtri = (bary[0] * v[3 * (size_t)face[0] + k] + bary[1] * v[3 * (size_t)face[1] + k]) +
                        bary[2] * v[3 * (size_t)face[2] + k];
    vrel[k] = v[3 * (size_t)vid + k] - vtri;
  }
  const double vn = v_dot(vrel, n);
  for (int k = 0; k < 3; ++k) vt[k] = vrel[k] - vn * n[k];
  const double vt_mag = sqrt(v_dot(vt, vt));
  double scale = 0.0;
  if (vt_mag > EPS) {
    const double bound = mu * fn_mag / (vt_mag + EPS);
    scale = -(kf < bound ? kf : bound);
  }
  for (int k = 0; k < 3; ++k) force[k] = fn_mag * n[k] + scale * vt[k];
  return 1;
}

static int unique_sorted(const int* faces, int count, int n, int* slot, int* verts) {
  for (int i = 0; i < n; ++i) slot[i] = -1;
  for (int t = 0; t < 3 * count; ++t) slot[faces[t]] = 0;
  int m = 0;
  for (int i = 0; i < n; ++i)
    if (slot[i] == 0) {
      slot[i] = m;
      verts[m++] = i;
    }
  return m;
}

int oracle_mesh_contact_force(int n, const double* q, const double* v, int fa, const int* faces_a, int fb,
                              const int* faces_b, double radius, double kn_left, double kn_right, double kf,
                              double mu, double* force) {
  int *slot_a = malloc(sizeof(int) * (size_t)n), *slot_b = malloc(sizeof(int) * (size_t)n);
  int *verts_a = malloc(sizeof(int) * (size_t)n), *verts_b = malloc(sizeof(int) * (size_t)n);
  const int na = unique_sorted(faces_a, fa, n, slot_a, verts_a);
  const int nb = unique_sorted(faces_b, fb, n, slot_b, verts_b);
  /* Per-pair forces and barycentric weights, zero outside the radius. */
  double* f_ab = calloc((size_t)na * fb * 6, sizeof(double));
  double* f_ba = calloc((size_t)nb * fa * 6, sizeof(double));
  int pairs = 0;
  for (int s = 0; s < na; ++s)
    for (int f = 0; f < fb; ++f) {
      double* e = f_ab + 6 * ((size_t)s * fb + f);
      if (!pair_force(q, v, verts_a[s], faces_b + 3 * f, radius, kn_left, kf, mu, e, e + 3)) {
        memset(e, 0, sizeof(double) * 6);
      } else {
        ++pairs;
      }
    }
  for (int s = 0; s < nb; ++s)
    for (int f = 0; f < fa; ++f) {
      double* e = f_ba + 6 * ((size_t)s * fa + f);
      if (!pair_force(q, v, verts_b[s], faces_a + 3 * f, radius, kn_right, kf, mu, e, e + 3)) {
        memset(e, 0, sizeof(double) * 6);
      } else {
        ++pairs;
      }
    }
  /* _gather_forces (mesh_contact.py:140-170): own pairs in face order, then
   * the reactions on the faces the particle is a corner of, in face and
   * vertex order. */
  for (int i = 0; i < n; ++i) {
    double t[3] = {0.0, 0.0, 0.0};
    if (slot_a[i] >= 0)
      for (int f = 0; f < fb; ++f)
        for (int k = 0; k < 3; ++k) t[k] = t[k] + f_ab[6 * ((size_t)slot_a[i] * fb + f) + k];
    if (slot_b[i] >= 0)
      for (int f = 0; f < fa; ++f)
        for (int k = 0; k < 3; ++k) t[k] = t[k] + f_ba[6 * ((size_t)slot_b[i] * fa + f) + k];
    for (int f = 0; f < fb; ++f)
      for (int corner = 0; corner < 3; ++corner)
        if (faces_b[3 * f + corner] == i)
          for (int s = 0; s < na; ++s) {
            const double* e = f_ab + 6 * ((size_t)s * fb + f);
            for (int k = 0; k < 3; ++k) t[k] = t[k] - e[k] * e[3 + corner];
          }
    for (int f = 0; f < fa; ++f)
      for (int corner = 0; corner < 3; ++corner)
        if (faces_a[3 * f + corner] == i)
          for (int s = 0; s < nb; ++s) {
            const double* e = f_ba + 6 * ((size_t)s * fa + f);
            for (int k = 0; k < 3; ++k) t[k] = t[k] - e[k] * e[3 + corner];
          }
    memcpy(force + 3 * (size_t)i, t, sizeof(t));
  }
  free(f_ab);
  free(f_ba);
  free(slot_a);
  free(slot_b);
  free(verts_a);
  free(verts_b);
  return pairs;
}
