/*
 * TEST ORACLE ONLY — plain-C FP64 restatement of the reference frictional
 * boundary (FrictionalContact) used by oracle_forward. Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it.
 *
 * Paths below are under src/newton_solver_diff_phd/_src/solvers/diff_phd/.
 *   boundary gap / normal .... contact.py:72-95
 *   mesh closest point ....... contact.py:98-147 (first minimal face wins)
 *   complementarity .......... contact.py:150-156
 *   step start (frames) ...... contact.py:171-244 (planar/spherical, mesh filter)
 *   Delassus ................. contact.py:247-263
 *   scalars .................. contact.py:266-316
 *   Newton step .............. contact.py:865-928 (+ :373-420, 432-543)
 *   friction ................. contact.py:546-584
 *   restitution .............. contact.py:587-660
 *   dense solves ............. dense.py:24-106 (Cholesky up to 32 rows, LDL^T above)
 * A^{-1} is formed column by column from the factor products, as the
 * reference's panel of the registered vertices (contact.py:782-800).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "contact_oracle.h"

#define EPS 2.220446049250313e-16
#define DMIN 2.2250738585072014e-308
#define TILE 32

static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double len3(const double* a) { return sqrt(dot3(a, a)); }
static void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

/* closest_point_triangle (mesh_contact.py:34-87). */
static void closest_point(const double* p, const double* a, const double* b, const double* c, double* point,
                          double* dist2) {
  double ab[3], ac[3], ap[3], bp[3], cp[3], d[3];
  for (int k = 0; k < 3; ++k) {
    ab[k] = b[k] - a[k];
    ac[k] = c[k] - a[k];
    ap[k] = p[k] - a[k];
    bp[k] = p[k] - b[k];
    cp[k] = p[k] - c[k];
  }
  const double d1 = dot3(ab, ap), d2 = dot3(ac, ap), d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  const double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
  if (d1 <= 0.0 && d2 <= 0.0) {
    memcpy(point, a, sizeof(double) * 3);
  } else if (d3 >= 0.0 && d4 <= d3) {
    memcpy(point, b, sizeof(double) * 3);
  } else if (d1 * d4 - d3 * d2 <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
    const double w = d1 / (d1 - d3);
    for (int k = 0; k < 3; ++k) point[k] = a[k] + w * ab[k];
  } else if (d6 >= 0.0 && d5 <= d6) {
    memcpy(point, c, sizeof(double) * 3);
  } else if (d5 * d2 - d1 * d6 <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
    const double w = d2 / (d2 - d6);
    for (int k = 0; k < 3; ++k) point[k] = a[k] + w * ac[k];
  } else if (d3 * d6 - d5 * d4 <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    for (int k = 0; k < 3; ++k) point[k] = b[k] + w * (c[k] - b[k]);
  } else {
    const double va = d3 * d6 - d5 * d4, vb = d5 * d2 - d1 * d6, vc = d1 * d4 - d3 * d2;
    const double denom = 1.0 / (va + vb + vc);
    const double v = vb * denom, w = vc * denom;
    for (int k = 0; k < 3; ++k) point[k] = (a[k] + ab[k] * v) + ac[k] * w;
  }
  for (int k = 0; k < 3; ++k) d[k] = p[k] - point[k];
  *dist2 = dot3(d, d);
}

/* The surface opposite `node` at positions `x`: closest point, unit face
 * normal, squared distance (infinite for an empty surface). */
static void mesh_closest(const oracle_contact* c, const double* p, int node, const double* x, double* closest,
                         double* normal, double* best) {
  int first = c->left_count, last = c->face_count;
  if (node >= c->split) {
    first = 0;
    last = c->left_count;
  }
  *best = INFINITY;
  closest[0] = closest[1] = closest[2] = 0.0;
  normal[0] = normal[1] = 0.0;
  normal[2] = 1.0;
  for (int f = first; f < last; ++f) {
    const int* face = c->faces + 3 * (size_t)f;
    const double *a = x + 3 * (size_t)face[0], *b = x + 3 * (size_t)face[1], *cc = x + 3 * (size_t)face[2];
    double point[3], dist2;
    closest_point(p, a, b, cc, point, &dist2);
    if (dist2 < *best) {
      *best = dist2;
      memcpy(closest, point, sizeof(point));
      double e1[3], e2[3], n[3];
      for (int k = 0; k < 3; ++k) {
        e1[k] = b[k] - a[k];
        e2[k] = cc[k] - a[k];
      }
      cross3(e1, e2, n);
      const double length = len3(n);
      if (length > EPS) {
        for (int k = 0; k < 3; ++k) normal[k] = n[k] / length;
      } else {
        normal[0] = normal[1] = 0.0;
        normal[2] = 1.0;
      }
    }
  }
}

static void mesh_direction(const double* p, const double* closest, const double* face_normal, double* n) {
  double offset[3];
  for (int k = 0; k < 3; ++k) offset[k] = p[k] - closest[k];
  const double length = len3(offset);
  if (length > 1.0e-12) {
    for (int k = 0; k < 3; ++k) n[k] = offset[k] / length;
  } else {
    memcpy(n, face_normal, sizeof(double) * 3);
  }
}

static double boundary_gap(const oracle_contact* c, const double* x) {
  const double* q = c->params;
  if (c->kind == 0) {
    const double norm = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    const double axis[3] = {q[0] / norm, q[1] / norm, q[2] / norm};
    return dot3(x, axis) + q[3] / norm;
  }
  const double d[3] = {x[0] - q[0], x[1] - q[1], x[2] - q[2]};
  return len3(d) - q[3];
}

static void boundary_normal(const oracle_contact* c, const double* x, double* n) {
  const double* q = c->params;
  if (c->kind == 0) {
    const double norm = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    for (int k = 0; k < 3; ++k) n[k] = q[k] / norm;
    return;
  }
  const double d[3] = {x[0] - q[0], x[1] - q[1], x[2] - q[2]};
  const double length = len3(d);
  for (int k = 0; k < 3; ++k) n[k] = d[k] / length;
}

static void complementarity(double delta, double lam, double r, double* omega, double* e, double* phi) {
  double s = sqrt(delta * delta + r * r * lam * lam);
  if (s < 1.0e-14) s = 1.0e-14;
  *omega = 1.0 - delta / s;
  *e = (1.0 - r * lam / s) * r;
  *phi = delta + r * lam - s;
}

/* Cholesky solve of an SPD block (the tile path of dense.py:24-37). */
static void cholesky_solve(double* m, int n, double* b) {
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

/* Eigen's LDLT with diagonal pivoting on the lower triangle (dense.py:40-106). */
void oracle_ldlt_solve(double* m, int n, double* b) {
  int* perm = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  double* work = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) perm[i] = i;
  for (int k = 0; k < n; ++k) {
    int big = k;
    for (int i = k + 1; i < n; ++i)
      if (fabs(m[i * n + i]) > fabs(m[big * n + big])) big = i;
    if (big != k) {
      double t;
      for (int j = 0; j < k; ++j) {
        t = m[k * n + j];
        m[k * n + j] = m[big * n + j];
        m[big * n + j] = t;
      }
      for (int i = big + 1; i < n; ++i) {
        t = m[i * n + k];
        m[i * n + k] = m[i * n + big];
        m[i * n + big] = t;
      }
      t = m[k * n + k];
      m[k * n + k] = m[big * n + big];
      m[big * n + big] = t;
      for (int i = k + 1; i < big; ++i) {
        t = m[i * n + k];
        m[i * n + k] = m[big * n + i];
        m[big * n + i] = t;
      }
      const int pi = perm[k];
      perm[k] = perm[big];
      perm[big] = pi;
    }
    for (int j = 0; j < k; ++j) m[k * n + k] = m[k * n + k] - m[k * n + j] * m[k * n + j] * m[j * n + j];
    for (int i = k + 1; i < n; ++i) {
      for (int j = 0; j < k; ++j) m[i * n + k] = m[i * n + k] - m[i * n + j] * m[j * n + j] * m[k * n + j];
      if (m[k * n + k] != 0.0) m[i * n + k] = m[i * n + k] / m[k * n + k];
    }
  }
  for (int i = 0; i < n; ++i) work[i] = b[perm[i]];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) work[i] = work[i] - m[i * n + j] * work[j];
  for (int i = 0; i < n; ++i) work[i] = fabs(m[i * n + i]) > DMIN ? work[i] / m[i * n + i] : 0.0;
  for (int i = n - 1; i >= 0; --i)
    for (int j = i + 1; j < n; ++j) work[i] = work[i] - m[j * n + i] * work[j];
  for (int i = 0; i < n; ++i) b[perm[i]] = work[i];
  free(perm);
  free(work);
}

int contact_prepare(const oracle_problem* p, contact_work* w) {
  const oracle_contact* c = p->contact;
  const int K = c->count, n = p->n;
  memset(w, 0, sizeof(*w));
  w->K = K;
  const int k1 = K > 0 ? K : 1;
  w->panel = calloc((size_t)k1 * n, sizeof(double));
  w->block = calloc((size_t)k1 * k1, sizeof(double));
  w->delassus = calloc((size_t)k1 * k1, sizeof(double));
  w->reg = calloc(k1, sizeof(double));
  w->t1 = calloc(3 * (size_t)k1, sizeof(double));
  w->t2 = calloc(3 * (size_t)k1, sizeof(double));
  w->fl = calloc(2 * (size_t)k1, sizeof(double));
  w->gap = calloc(k1, sizeof(double));
  w->diag = calloc(k1, sizeof(double));
  w->resid = calloc(k1, sizeof(double));
  w->rhs = calloc(k1, sizeof(double));
  w->force = calloc(3 * (size_t)k1, sizeof(double));
  w->response = calloc(3 * (size_t)k1, sizeof(double));
  w->order = calloc(k1, sizeof(int));
  w->slot = calloc(k1, sizeof(int));
  w->matrix = calloc((size_t)k1 * k1, sizeof(double));
  w->compact = calloc(k1, sizeof(double));
  w->surface = calloc(3 * (size_t)n, sizeof(double));
  /* Columns A^{-1} e_j of the registered vertices (the same scalar column
   * for every coordinate). */
  double* e = calloc(3 * (size_t)n, sizeof(double));
  double* col = calloc(3 * (size_t)n, sizeof(double));
  for (int j = 0; j < K; ++j) {
    const int node = c->nodes[j];
    e[3 * (size_t)node] = 1.0;
    oracle_apply_inverse(p, e, col);
    e[3 * (size_t)node] = 0.0;
    for (int i = 0; i < n; ++i) w->panel[(size_t)j * n + i] = col[3 * (size_t)i];
  }
  free(e);
  free(col);
  for (int i = 0; i < K; ++i)
    for (int j = 0; j < K; ++j) w->block[(size_t)i * K + j] = w->panel[(size_t)j * n + c->nodes[i]];
  return 0;
}

void contact_free(contact_work* w) {
  free(w->panel);
  free(w->block);
  free(w->delassus);
  free(w->reg);
  free(w->t1);
  free(w->t2);
  free(w->fl);
  free(w->gap);
  free(w->diag);
  free(w->resid);
  free(w->rhs);
  free(w->force);
  free(w->response);
  free(w->order);
  free(w->slot);
  free(w->matrix);
  free(w->compact);
  free(w->surface);
}

/* begin_contact_step / begin_mesh_step and contact_delassus. */
void contact_begin(const oracle_problem* p, contact_work* w, const double* pos, double dt) {
  oracle_contact* c = p->contact;
  const int K = c->count;
  for (int k = 0; k < K; ++k) {
    const int node = c->nodes[k];
    const double* x = pos + 3 * (size_t)node;
    double* n = c->normal + 3 * (size_t)k;
    double* t1 = w->t1 + 3 * (size_t)k;
    double* t2 = w->t2 + 3 * (size_t)k;
    double u[3];
    if (c->kind == 2) {
      double closest[3], face_normal[3], dist2;
      mesh_closest(c, x, node, pos, closest, face_normal, &dist2);
      const int active = sqrt(dist2) - c->radius <= 5.0 * c->radius;
      mesh_direction(x, closest, face_normal, n);
      int axis = 0;
      for (int d = 1; d < 3; ++d)
        if (fabs(n[d]) < fabs(n[axis])) axis = d;
      double unit[3] = {0.0, 0.0, 0.0};
      unit[axis] = 1.0;
      const double nu = dot3(n, unit);
      for (int d = 0; d < 3; ++d) t1[d] = unit[d] - nu * n[d];
      const double l1 = len3(t1);
      for (int d = 0; d < 3; ++d) t1[d] = t1[d] / l1;
      c->candidate[k] = active;
      if (!active) c->multiplier[k] = 0.0;
    } else {
      boundary_normal(c, x, n);
      double best[3] = {0.0, 0.0, 0.0};
      for (int i = 0; i < 3; ++i) {
        double axis[3] = {0.0, 0.0, 0.0}, cand[3];
        axis[i] = 1.0;
        cross3(n, axis, cand);
        if (dot3(cand, cand) > dot3(best, best)) memcpy(best, cand, sizeof(best));
      }
      const double lb = len3(best);
      for (int d = 0; d < 3; ++d) t1[d] = best[d] / lb;
      c->candidate[k] = 1;
    }
    cross3(n, t1, u);
    const double l2 = len3(u);
    for (int d = 0; d < 3; ++d) t2[d] = u[d] / l2;
    w->fl[2 * k] = w->fl[2 * k + 1] = 0.0;
  }
  for (int i = 0; i < K; ++i)
    for (int j = 0; j < K; ++j) {
      const int lo = i < j ? i : j, hi = i < j ? j : i;
      w->delassus[(size_t)i * K + j] =
          dot3(c->normal + 3 * (size_t)lo, c->normal + 3 * (size_t)hi) * w->block[(size_t)hi * K + lo];
    }
  for (int i = 0; i < K; ++i) {
    const double r = dt * dt * w->delassus[(size_t)i * K + i];
    w->reg[i] = r > 1.0e-10 ? r : 1.0e-10;
  }
}

/* contact_forces: omega lambda n (+ the tangential multipliers). */
static void forces(const oracle_contact* c, contact_work* w, int tangential) {
  for (int k = 0; k < c->count; ++k) {
    const double s = c->omega[k] * c->multiplier[k];
    for (int d = 0; d < 3; ++d) {
      double v = s * c->normal[3 * (size_t)k + d];
      if (tangential) v = v + (w->fl[2 * k] * w->t1[3 * (size_t)k + d] + w->fl[2 * k + 1] * w->t2[3 * (size_t)k + d]);
      w->force[3 * (size_t)k + d] = v;
    }
  }
}

/* contact_response: B f over the coupled candidates. */
static void response(const oracle_contact* c, contact_work* w) {
  const int K = c->count;
  for (int i = 0; i < K; ++i) {
    double v[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < w->active; ++a) {
      const int j = w->order[a];
      const double b = w->block[(size_t)i * K + j];
      for (int d = 0; d < 3; ++d) v[d] = v[d] + b * w->force[3 * (size_t)j + d];
    }
    memcpy(w->response + 3 * (size_t)i, v, sizeof(v));
  }
}

void contact_iterate(const oracle_problem* p, contact_work* w, const double* x, const double* solved,
                     const double* pos, double* out) {
  oracle_contact* c = p->contact;
  const int K = c->count, n = p->n;
  if (c->kind == 2 && c->restitution > 0.0) memcpy(w->surface, x, sizeof(double) * 3 * (size_t)n);
  /* contact_scalars / mesh_contact_scalars */
  for (int k = 0; k < K; ++k) {
    const int node = c->nodes[k];
    const double* pt = x + 3 * (size_t)node;
    double delta;
    if (c->kind == 2) {
      if (!c->candidate[k]) {
        w->gap[k] = 0.0;
        c->omega[k] = 0.0;
        w->diag[k] = 1.0;
        w->resid[k] = 0.0;
        continue;
      }
      double closest[3], fn[3], dist2, d[3];
      mesh_closest(c, pt, node, x, closest, fn, &dist2);
      for (int t = 0; t < 3; ++t) d[t] = pt[t] - closest[t];
      delta = dot3(d, c->normal + 3 * (size_t)k) - c->radius;
    } else {
      delta = boundary_gap(c, pt);
    }
    w->gap[k] = delta;
    complementarity(delta, c->multiplier[k], w->reg[k], &c->omega[k], &w->diag[k], &w->resid[k]);
  }
  /* contact_active */
  w->active = 0;
  for (int k = 0; k < K; ++k) {
    w->slot[k] = -1;
    if (c->omega[k] != 0.0) {
      w->order[w->active] = k;
      w->slot[k] = w->active++;
    }
  }
  forces(c, w, 1);
  response(c, w);
  /* contact_newton_rhs */
  for (int k = 0; k < K; ++k) {
    const int node = c->nodes[k];
    const double* nk = c->normal + 3 * (size_t)k;
    double ps[3];
    for (int d = 0; d < 3; ++d) ps[d] = solved[3 * (size_t)node + d] + w->response[3 * (size_t)k + d];
    w->rhs[k] = -w->resid[k] + c->omega[k] * dot3(nk, x + 3 * (size_t)node) - c->omega[k] * dot3(nk, ps);
  }
  /* contact_system + the dense solve */
  const int A = w->active;
  for (int i = 0; i < A; ++i) {
    const int a = w->order[i];
    for (int j = 0; j < A; ++j) {
      const int b = w->order[j];
      double v = c->omega[a] * w->delassus[(size_t)a * K + b] * c->omega[b];
      if (i == j) v += w->diag[a];
      w->matrix[(size_t)i * A + j] = v;
    }
    w->compact[i] = w->rhs[a];
  }
  if (A > 0) {
    if (A <= TILE)
      cholesky_solve(w->matrix, A, w->compact);
    else
      oracle_ldlt_solve(w->matrix, A, w->compact);
  }
  /* contact_update */
  for (int k = 0; k < K; ++k) {
    const double step = w->slot[k] >= 0 ? w->compact[w->slot[k]] : w->rhs[k] / w->diag[k];
    double v = c->multiplier[k] + step;
    v = v > 0.0 ? v : 0.0;
    c->multiplier[k] = v < 1.0e8 ? v : 1.0e8;
  }
  /* contact_friction with the normal-only response of the new multipliers */
  if (c->friction > 0.0) {
    forces(c, w, 0);
    response(c, w);
    for (int k = 0; k < K; ++k) {
      if (c->multiplier[k] <= 0.0) {
        w->fl[2 * k] = w->fl[2 * k + 1] = 0.0;
        continue;
      }
      const int node = c->nodes[k];
      const double *nk = c->normal + 3 * (size_t)k, *t1 = w->t1 + 3 * (size_t)k, *t2 = w->t2 + 3 * (size_t)k;
      double step[3], tan[3];
      for (int d = 0; d < 3; ++d)
        step[d] = solved[3 * (size_t)node + d] + w->response[3 * (size_t)k + d] - pos[3 * (size_t)node + d];
      const double sn = dot3(step, nk);
      for (int d = 0; d < 3; ++d) tan[d] = step[d] - sn * nk[d];
      const double u1 = dot3(tan, t1), u2 = dot3(tan, t2);
      const double bd = w->block[(size_t)k * K + k];
      const double w11 = dot3(t1, t1) * bd, w12 = dot3(t1, t2) * bd, w22 = dot3(t2, t2) * bd;
      const double det = w11 * w22 - w12 * w12;
      double s1 = 0.0, s2 = 0.0;
      if (fabs(det) > 1.0e-20) {
        s1 = -(w22 * u1 - w12 * u2) / det;
        s2 = -(-w12 * u1 + w11 * u2) / det;
      }
      const double radius = c->friction * c->multiplier[k];
      if (sqrt(s1 * s1 + s2 * s2) <= radius + 1.0e-14) {
        w->fl[2 * k] = s1;
        w->fl[2 * k + 1] = s2;
      } else {
        const double mag = sqrt(u1 * u1 + u2 * u2);
        if (mag > 1.0e-14) {
          w->fl[2 * k] = -radius * u1 / mag;
          w->fl[2 * k + 1] = -radius * u2 / mag;
        } else {
          w->fl[2 * k] = w->fl[2 * k + 1] = 0.0;
        }
      }
    }
  }
  /* contact_forces + contact_solution */
  forces(c, w, 1);
  for (int i = 0; i < n; ++i) {
    double v[3] = {solved[3 * (size_t)i], solved[3 * (size_t)i + 1], solved[3 * (size_t)i + 2]};
    for (int a = 0; a < A; ++a) {
      const int j = w->order[a];
      const double pv = w->panel[(size_t)j * n + i];
      for (int d = 0; d < 3; ++d) v[d] = v[d] + pv * w->force[3 * (size_t)j + d];
    }
    memcpy(out + 3 * (size_t)i, v, sizeof(v));
  }
}

/* apply_restitution / apply_mesh_restitution after convergence. */
void contact_finish(const oracle_problem* p, contact_work* w, const double* pos, const double* vel, double dt,
                    double* x) {
  const oracle_contact* c = p->contact;
  if (c->restitution <= 0.0) return;
  for (int k = 0; k < c->count; ++k) {
    if (c->multiplier[k] <= 0.0) continue;
    const int node = c->nodes[k];
    double* pt = x + 3 * (size_t)node;
    double n[3], distance;
    if (c->kind == 2) {
      double closest[3], fn[3], dist2, d[3];
      mesh_closest(c, pt, node, w->surface, closest, fn, &dist2);
      const double* ref = c->normal + 3 * (size_t)k;
      for (int t = 0; t < 3; ++t) d[t] = pt[t] - closest[t];
      distance = dot3(d, ref) - c->radius;
      mesh_direction(pt, closest, fn, n);
      if (dot3(n, ref) < 0.0)
        for (int t = 0; t < 3; ++t) n[t] = -n[t];
    } else {
      distance = boundary_gap(c, pt);
      boundary_normal(c, pt, n);
    }
    const double norm = len3(n);
    if (norm <= 1.0e-12) continue;
    for (int t = 0; t < 3; ++t) n[t] = n[t] / norm;
    if (distance > (c->kind == 2 ? 0.1 * c->radius : 1.0e-8)) continue;
    const double incoming = dot3(vel + 3 * (size_t)node, n);
    if (incoming < 0.0) {
      double move[3];
      for (int t = 0; t < 3; ++t) move[t] = (pt[t] - pos[3 * (size_t)node + t]) / dt;
      const double change = -c->restitution * incoming - dot3(move, n);
      if (change > 0.0)
        for (int t = 0; t < 3; ++t) pt[t] = pt[t] + dt * change * n[t];
    }
  }
}
