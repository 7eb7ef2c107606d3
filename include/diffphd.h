/*
 * diffphd.h — C-ABI of the MI355X-native DiffPhD projective-dynamics path.
 *
 * Drop-in boundary for the reference solver plugin SolverDiffPhD
 * (src/newton_solver_diff_phd/_src/solvers/diff_phd/solver.py:47). The
 * reference binds nothing through an FFI (it is Python + Warp); these entry
 * points are what its Python class would call through ctypes in place of its
 * Warp launches. Every pointer is either a host array (setup calls) or a raw
 * device pointer (step calls: torch.Tensor.data_ptr()); streams are
 * hipStream_t passed as void*. All calls return DPHD_OK or an error code;
 * dphd_last_error() describes the last failure on the calling thread.
 * Numerical failures never abort a step: they set status bits (see below),
 * read with dphd_read_status.
 */
#ifndef DIFFPHD_H
#define DIFFPHD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DPHD_ABI_VERSION 6

enum {
  DPHD_OK = 0,
  DPHD_EINVAL = 1,   /* invalid argument (reference: ValueError/TypeError) */
  DPHD_ESTATE = 2,   /* call out of order (reference: RuntimeError) */
  DPHD_EHIP = 3,     /* HIP runtime failure */
  DPHD_ENOTIMPL = 4, /* unsupported option (reference: NotImplementedError) */
  DPHD_ENOMEM = 5
};

enum { DPHD_TET4 = 0, DPHD_HEX8 = 1 };
enum { DPHD_PD = 0, DPHD_NEO_HOOKEAN = 1 };

/* Status bits, identical to DP/kernels.py:29-31 and DP/adjoint.py:64-65. */
enum {
  DPHD_STATUS_NOT_CONVERGED = 1,
  DPHD_STATUS_SINGULAR = 2,
  DPHD_STATUS_ACTIVATION = 4,
  DPHD_STATUS_ADJOINT_NOT_CONVERGED = 8,
  DPHD_STATUS_ADJOINT_SAFEGUARD = 16
};

typedef struct dphd_factor dphd_factor;
typedef struct dphd_solver dphd_solver;

const char* dphd_last_error(void);
int dphd_version(void);

/* ---- Exact sparse inverse factor A^{-1} = S^T S --------------------------
 * Replaces SparseInverse.__init__ (DP/sparse.py:291-320): nested dissection,
 * LDL^T, inverse rows, CSR of S and S^T. Input: full symmetric CSR (both
 * halves, ascending columns). Runs on the host. */
int dphd_factor_create(int n, const int* offsets, const int* columns, const double* values, dphd_factor** out);
int dphd_factor_sizes(const dphd_factor* f, int64_t* nnz, int64_t* factorization_nnz, int64_t* matrix_nnz);
int dphd_factor_export(const dphd_factor* f, int* permutation, int* s_offsets, int* s_columns, double* s_values,
                       int* t_offsets, int* t_columns, double* t_values);
void dphd_factor_destroy(dphd_factor* f);

/* ---- Topology and geometry helpers ----------------------------------------
 * Vertex incidence sorted by (element, corner) (DP/kernels.py:61-72,
 * DP/hex.py:83-92); entries = element * corners + corner. */
int dphd_build_incidence(int vertices, int elements, int corners, const int* indices, int* offsets, int* entries);
/* Tet4 gradients [E,4,3] and volumes from float32 inverse rest edge matrices
 * (DP/solver.py:214-223). */
int dphd_tet_geometry(int elements, const float* poses, double* gradients, double* volumes);
/* Hex8 shared Gauss gradients [8 samples, 8 corners, 3] (DP/hex.py:63-80). */
int dphd_hex_gradients(double dx, double* out);
/* Constant matrix of one body (DP/solver.py:521-545). Writes nnz; copies the
 * CSR when capacity >= nnz. */
int dphd_assemble(int vertices, int elements, int corners, int samples, const int* indices, const double* grads,
                  int shared_grads, const double* weights, const double* fibre_dir, const double* fibre_w,
                  const unsigned char* free, double inertia, int plane_count, const int* plane_vertices,
                  const double* plane_stiffness, int* nnz, int* offsets, int* columns, double* values, int capacity);

/* ---- Solver ----------------------------------------------------------------
 * One solver holds B >= 1 independent bodies packed by vertex and element
 * offsets (B = 1 is the reference's single-model solver). Constructor
 * options mirror SolverDiffPhD.__init__ (DP/solver.py:157-179). */
typedef struct {
  int element_kind;      /* DPHD_TET4 or DPHD_HEX8 */
  int material;          /* DPHD_PD or DPHD_NEO_HOOKEAN */
  int anderson_window;   /* 0..8, reference default 5 */
  int max_iterations;    /* >= 1 */
  double abs_tol;        /* >= 0 */
  double rel_tol;        /* >= 0 */
  int factor_precision;  /* 64: FP64 S values; 32: FP32 S values with an FP64 defect correction */
  int integrate_pinned_velocities;
  int jacobian_mode;     /* 0: reference mean weights in K; 1: per-element weights (exact derivative) */
  int iteration_chunk;   /* iterations per captured graph launch; 0 picks the defaults (4 forward, 8 adjoint) */
  int slice_rows;        /* rows per slice of the sliced S^T product; 0 = tuned at dphd_build on this solver's
                            factor (the choice is in dphd_info), > 0 fixes it (bitwise repeatable builds) */
} dphd_config;

typedef struct {
  int bodies, vertices, elements, samples_per_element, corners;
  int64_t factor_nnz;       /* nnz(S) summed over bodies (S^T has the same) */
  int64_t matrix_nnz;       /* nnz(A) per coordinate summed over bodies */
  int64_t jacobian_bytes;   /* device bytes of the adjoint's Jacobian, element blocks and assembled operator */
  double setup_seconds;     /* host time of the last dphd_build */
  int64_t factor_values;    /* values stored on the device for S (dense supernode blocks, zeros included) */
  int64_t factor_slots;     /* postorder slots of the device vectors (vertices plus alignment gaps) */
  int slice_rows;           /* rows per slice of the sliced S^T product (tuned or fixed) */
  int slice_switch;         /* the sliced S^T runs while the active bodies hold fewer column groups */
  double tune_seconds;      /* time of the product tuning within setup_seconds (0 when fixed) */
  int split_quads;          /* quads per column chunk of the split S product (tuned or the default) */
  int split_switch;         /* the split S runs while the active bodies hold fewer column groups */
} dphd_info;

int dphd_create(const dphd_config* config, int device, dphd_solver** out);
void dphd_destroy(dphd_solver* s);

/* Packed bodies: vertex_offsets[B+1], element_offsets[B+1]. */
int dphd_set_bodies(dphd_solver* s, int bodies, const int* vertex_offsets, const int* element_offsets);
/* Tet4: indices [E,4] global vertex ids, poses [E,3,3] float32 Dm^{-1}
 * (Newton model.tet_indices / model.tet_poses). */
int dphd_set_tet4(dphd_solver* s, const int* indices, const float* poses);
/* Hex8: indices [E,8] binary corner order, rest vertices [N,3] (DP/hex.py:27-92). */
int dphd_set_hex8(dphd_solver* s, const int* indices, const double* rest_vertices);
/* inverse_mass [N] (0 = Dirichlet), gravity [B,3], density [B]. */
int dphd_set_particles(dphd_solver* s, const float* inverse_mass, const double* gravity, const double* density);
/* materials [E,2]: (mu, lambda) per element (model.tet_materials[:, :2]). */
int dphd_set_materials(dphd_solver* s, const double* materials);
/* Optional fibres (DP/solver.py:429-462): directions [E,3], stiffness [E]. */
int dphd_set_muscles(dphd_solver* s, const double* directions, const double* stiffness);
/* Optional planar vertex energies (PlanarEnergies, DP/hex.py:204-260, and
 * their matrix diagonal, DP/solver.py:542-545): `count` distinct vertices
 * [count] int32 with stiffness [count] float64 >= 0 against one plane
 * (nx, ny, nz, offset), scaled to a unit normal. count = 0 removes them.
 * Forward only: dphd_backward rejects a solver that has them. */
int dphd_set_planar(dphd_solver* s, int count, const int* vertices, const double* stiffness, const double* plane);
/* Optional mesh contact state force between two bodies (MeshContact,
 * DP/mesh_contact.py:180-272; solver.py:321-325, 762-770): contact radius [m],
 * normal stiffnesses of A's vertices against B's faces (kn_left) and of B's
 * against A's (kn_right) [N/m], friction bound kf [N s/m], Coulomb mu, and
 * the surface triangles of A and B [count,3] int32 (host). Evaluated on the
 * start-of-step state into q~ += h^2/m f; no gradient. All-zero/NULL faces
 * remove it. */
int dphd_set_mesh_contact(dphd_solver* s, double contact_radius, double kn_left, double kn_right, double kf,
                          double mu, int faces_a_count, const int* faces_a, int faces_b_count, const int* faces_b);
/* Optional frictional boundary (FrictionalContact, DP/contact.py:686-1015;
 * solver.py:302-319, 628-630, 772-774): kind 0 planar (nx, ny, nz, offset),
 * 1 spherical (cx, cy, cz, radius), 2 mesh (parameters[0] = contact radius;
 * vertices below `split` touch faces [left_count, face_count), the others
 * faces [0, left_count) of `faces` [face_count,3]); `count` registered
 * vertices [count] int32; Coulomb friction and restitution coefficients.
 * Host arrays. kind < 0 removes it. Needs a single-body solver
 * (DPHD_ENOTIMPL otherwise, at dphd_build); works with either factor
 * precision. */
int dphd_set_frictional_boundary(dphd_solver* s, int kind, const double* parameters, int count, const int* vertices,
                                 double friction, double restitution, int split, int left_count, int face_count,
                                 const int* faces);
/* Read (get) and/or replace (set) the four boundary parameters between steps
 * (SolverDiffPhD.boundary_parameters, DP/solver.py:340-343); host arrays. */
int dphd_boundary_parameters(dphd_solver* s, const double* set, double* get);
/* Contact state after the last step (SolverDiffPhD.contact_multipliers,
 * DP/solver.py:345-348): optional copies of the normal multipliers [K], the
 * Fischer-Burmeister weights omega [K], the frozen normals [K,3] and the mesh
 * candidate flags [K]; set_multiplier replaces the multipliers (the warm
 * start). Host or device pointers; synchronizes. */
int dphd_contact_state(dphd_solver* s, const double* set_multiplier, double* multiplier, double* omega, double* normal,
                       int* candidate);
/* The contact state a following dphd_backward projects its parameter
 * gradients with (StepRecord.contact, DP/adjoint.py:809-841): multipliers
 * [K], omega [K], normals [K,3] of the recorded step, e.g. copies taken with
 * dphd_contact_state after it. Every dphd_step records its own; this
 * replaces it for the backward of an earlier step. Host or device pointers;
 * waits for the solver's queued work and synchronizes. */
int dphd_contact_record(dphd_solver* s, const double* multiplier, const double* omega, const double* normal);
/* Mesh contact force [N,3] float64 of float64 positions/velocities [N,3]
 * (MeshContact.evaluate, DP/mesh_contact.py:245-272); device pointers. Runs
 * on `stream` after the solver's queued work, and the solver's later work
 * runs after it. */
int dphd_mesh_contact_force(dphd_solver* s, const double* position, const double* velocity, double* force,
                            void* stream);
/* Vertex-face pairs within the contact radius in the last evaluation
 * (SolverDiffPhD.mesh_contact_pairs, DP/solver.py:350-353); -1 without mesh
 * contact. Synchronizes. */
int dphd_mesh_contact_pairs(dphd_solver* s, int* pairs);
/* The same pairs by direction: counts[0] vertices of faces_a against faces_b,
 * counts[1] vertices of faces_b against faces_a (the two per-step pair counts
 * the reference's parity harness records from MeshContact.bary_ab / bary_ba,
 * tools/parity/port_cases.py:444-447); -1 each without mesh contact.
 * Synchronizes. */
int dphd_mesh_contact_pair_counts(dphd_solver* s, int* counts);
/* Mutable between steps (DP/solver.py:206-209). */
int dphd_set_tolerances(dphd_solver* s, int max_iterations, double abs_tol, double rel_tol);
/* Element weights, assembly and factor per body (DP/solver.py:491-552);
 * with a frictional boundary also the solver of the backward's Woodbury
 * columns. Returns with the device idle. */
int dphd_build(dphd_solver* s, double dt);
int dphd_get_info(const dphd_solver* s, dphd_info* info);
/* nnz(S) of every body after dphd_build (per_body [B]); sizes the byte model. */
int dphd_body_nnz(const dphd_solver* s, int64_t* per_body);
/* Loop counters since the last dphd_build (diagnostics; no reference
 * counterpart): stats[6] = for the forward, adjoint and Woodbury loops in
 * turn, the iteration chunks launched and the host waits between them. */
int dphd_loop_stats(const dphd_solver* s, int64_t* stats);

/* One forward step of all bodies (DP/solver.py:689-786). q/qd/f are [N,3]
 * float32 device arrays; dirichlet is [N,3] float32 or NULL
 * (DP/solver.py:755-761); muscle_targets [E] float64 device or NULL (target
 * fibre lengths a = 1 - u); checkpoint [N,3] float64 device or NULL
 * receives the converged iterate for the adjoint (StepRecord.x). */
int dphd_step(dphd_solver* s, const float* q_in, const float* qd_in, const float* f_in, const float* dirichlet,
              const double* muscle_targets, float* q_out, float* qd_out, double* checkpoint, void* stream);

/* Backward of one recorded step (Adjoint.backward, DP/adjoint.py:1088-1259).
 * Seeds from gq_out/gqd_out (NULL = zero); accumulates (+=) into gq_in,
 * gqd_in, gf_in [N,3] float32, gmaterials [E,2] float64 (mu, lambda; NULL
 * skips material gradients), gdirichlet [N,3] float32 (non-NULL when the
 * step used explicit Dirichlet targets) and gactivations [E] float64
 * (d/du of the Tet4 fibre contraction u = 1 - target, DP/adjoint.py:595-620;
 * NULL skips them, non-NULL needs Tet4 fibres). Material and activation
 * gradients pass the reference's aggregate guard (DP/adjoint.py:622-659);
 * with a frictional boundary they use the adjoint projected off the
 * recorded step's active contact rows (Woodbury, DP/adjoint.py:1272-1354). */
int dphd_backward(dphd_solver* s, const double* checkpoint, const double* muscle_targets, const float* gq_out,
                  const float* gqd_out, float* gq_in, float* gqd_in, float* gf_in, double* gmaterials,
                  float* gdirichlet, double* gactivations, void* stream);

/* dphd_backward without device allocation. No counterpart in the reference:
 * for callers that cannot call hipMalloc (tight serving loops, harnesses
 * that forbid allocation). Same arguments, results and status bits as
 * dphd_backward, whose first call allocates the adjoint's Jacobian, element
 * blocks and assembled operator; here they live in a static 64 MiB device
 * buffer per GPU (up to roughly 30,000 Tet4 or 6,000 Hex8 elements per
 * solver; larger solvers get DPHD_EINVAL with the bytes they need), and the
 * adjoint iterations are launched directly instead of through a graph.
 * The buffer is shared by every solver on a GPU: static backward calls
 * that overlap on one GPU race. Parameter gradients with a frictional
 * boundary (whose Woodbury columns are allocated) give DPHD_ENOTIMPL. */
int dphd_backward_static(dphd_solver* s, const double* checkpoint, const double* muscle_targets, const float* gq_out,
                         const float* gqd_out, float* gq_in, float* gqd_in, float* gf_in, double* gmaterials,
                         float* gdirichlet, double* gactivations, void* stream);

/* Status word (DP/solver.py:380-390): OR of all bodies since the last clear. */
int dphd_read_status(dphd_solver* s, int clear, int* bits);
/* Iterations of the last forward (adjoint = 0) or backward (adjoint = 1)
 * solve per body; host copy, synchronizes. */
int dphd_iterations(dphd_solver* s, int adjoint, int* per_body);
int dphd_synchronize(dphd_solver* s);

/* Device-side application of A^{-1} = S^T S of every body to [N,3] float64
 * device arrays (SparseInverse.apply, DP/sparse.py:350-352). Runs on
 * `stream` after the solver's queued work (it shares the solver's product
 * buffers), and the solver's later work runs after it. */
int dphd_apply_inverse(dphd_solver* s, const double* rhs, double* out, void* stream);

/* det(F) at the eight Gauss samples of every Hex8 cell, out [E,8] float64,
 * from float32 positions q [N,3] (SolverDiffPhD.deformation_determinants,
 * DP/solver.py:360-378; DP/hex.py:190-201). Device pointers; Hex8 only. */
int dphd_deformation_determinants(dphd_solver* s, const float* q, double* out, void* stream);

/* Evaluate the pointwise projections of `count` float64 3x3 matrices on the
 * device (the reference's _evaluate_projections test kernel,
 * tests/test_diff_phd_solver.py:34-56): sigma [count,3], rotation, volume and
 * neo-Hookean projections [count,3,3], fibre projection [count,3] with
 * target length fibre_target. All device pointers. */
int dphd_eval_projections(int count, const double* f, const double* directions, double mu, double lame,
                          double penalty, double fibre_target, double* sigma, double* reconstruction,
                          double* rotation, double* volume, double* neo_hookean, double* fibre, void* stream);
/* Directional derivatives of the same projections along df (the reference's
 * _jvps kernel, tests/test_diff_phd_adjoint.py:54-73). */
int dphd_eval_jvps(int count, const double* f, const double* df, const double* directions, double mu, double lame,
                   double fibre_target, double* rotation, double* volume, double* neo_hookean, double* fibre,
                   void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DIFFPHD_H */
