/*
 * TEST ORACLE ONLY — plain-C FP64 restatement of the reference DiffPhD step
 * and its backward pass, following the reference's CPU execution path.
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load it, and only as the checker. See diffphd_oracle.c for citations.
 */
#ifndef DIFFPHD_ORACLE_H
#define DIFFPHD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Frictional boundary of the step (contact.py FrictionalContact): planar
 * (n, offset), spherical (center, radius) or mesh (MeshFrictionalBoundary)
 * against `count` registered vertices. `multiplier` carries the normal
 * multipliers between steps (warm start); omega/normal/candidate receive
 * the last step's values for the adjoint's Woodbury rows. */
typedef struct {
  int kind;               /* 0 planar, 1 spherical, 2 mesh */
  double params[4];
  int count;
  const int* nodes;
  double friction, restitution;
  double* multiplier;     /* [count] in/out */
  double* omega;          /* [count] out */
  double* normal;         /* [count,3] out */
  int* candidate;         /* [count] out */
  /* mesh: faces = left faces then right faces; a vertex below `split`
   * touches the right faces, the others the left ones */
  double radius;
  int split, left_count, face_count;
  const int* faces;
} oracle_contact;

typedef struct {
  int n;          /* vertices */
  int elements;   /* elements */
  int corners;    /* 4 (Tet4) or 8 (Hex8) */
  int samples;    /* 1 (Tet4) or 8 (Hex8) */
  const int* indices;     /* [elements, corners] */
  const double* grads;    /* Tet4: [elements, 4, 3]; Hex8: [8, 8, 3] shared */
  const double* wc;       /* [elements] per-sample corotated weight */
  const double* wv;       /* [elements] per-sample volume weight */
  const double* wn;       /* [elements] per-sample neo-Hookean weight */
  const double* jwc;      /* [elements] Jacobian weights (mean or element) */
  const double* jwv;
  const double* jwn;
  const double* mdir;     /* [elements, 3] or NULL */
  const double* mw;       /* [elements] per-sample fibre weight or NULL */
  int material;           /* 0 pd, 1 neo-Hookean */
  double nh_mu, nh_lame, nh_penalty;
  const unsigned char* free; /* [n] */
  double mass;            /* particle mass */
  double sample_volume;   /* Vbar / samples: the volume each quadrature sample carries */
  double gravity[3];
  /* Reference incidence: entries index [elements * samples, corners]. */
  const int* inc_off;
  const int* inc_ent;
  /* S and S^T (DP/sparse.py). */
  const int* s_off;
  const int* s_col;
  const double* s_val;
  const int* t_off;
  const int* t_col;
  const double* t_val;
  int window;
  int max_iterations;
  double abs_tol, rel_tol;
  int integrate_pinned;
  int threads;            /* OpenMP threads, <= 0 = default */
  /* Planar vertex energies (hex.py:204-260): k_j (x - min(0, n.x + d) n - x_D)
   * added to the projection of vertex idx[j]; plane = (n, d) with |n| = 1. */
  int planar_count;
  const int* planar_idx;
  const double* planar_w;
  double plane[4];
  /* State force [n,3] or NULL, added to the inertial target as
   * q~ += h^2/m f (mesh_contact.py:173-177). */
  const double* state_force;
  /* Frictional boundary or NULL. */
  oracle_contact* contact;
} oracle_problem;

int oracle_forward(const oracle_problem* p, double dt, const float* q, const float* qd, const float* f,
                   const float* dirichlet, const double* targets, double* x_out, float* q_out, float* qd_out,
                   int* iterations, int* status);

int oracle_backward(const oracle_problem* p, double dt, const double* x, const double* targets, const float* gq_out,
                    const float* gqd_out, float* gq_in, float* gqd_in, float* gf_in, double* gmaterials,
                    float* gdirichlet, double* gactivations, double* mu_out, int* iterations, int* status);

/* Safe clip of one adjoint vector v [n,3] in place with norm limit `limit`
 * (adjoint.py _clip_partials / _clip_apply); returns the status bits (16 when
 * a component was scrubbed or the norm clipped). */
int oracle_safe_clip(int n, const unsigned char* free_rows, double* v, double limit);

/* Routing guard (adjoint.py _routing_guard) from |rhs|^2 and |mu|^2: writes
 * the state-gradient scale and the divergence flag; returns the status bits. */
int oracle_routing_guard(double rhs_sq, double mu_sq, double inertia, double clip_norm, double* scale,
                         int* diverged);

void oracle_projections(int count, const double* f, const double* dirs, double mu, double lame, double penalty,
                        double fibre_target, double* sigma, double* recon, double* rotation, double* volume,
                        double* neo_hookean, double* fibre);

void oracle_jvps(int count, const double* f, const double* df, const double* dirs, double mu, double lame,
                 double fibre_target, double* rotation, double* volume, double* neo_hookean, double* fibre);

void oracle_apply_inverse(const oracle_problem* p, const double* rhs, double* out);

/* Mesh contact state force of positions q and velocities v [n,3]
 * (MeshContact.evaluate, mesh_contact.py:90-170) into force [n,3]; returns
 * the number of vertex-face pairs within the radius. */
int oracle_mesh_contact_force(int n, const double* q, const double* v, int fa, const int* faces_a, int fb,
                              const int* faces_b, double radius, double kn_left, double kn_right, double kf,
                              double mu, double* force);

#ifdef __cplusplus
}
#endif

#endif
