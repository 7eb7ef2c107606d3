// Contact between bodies on gfx950.
//
// Mesh contact (the reference's MeshContactStateForce,
// src/newton_solver_diff_phd/_src/solvers/diff_phd/mesh_contact.py): a
// penalty force between the surface triangles of two bodies, evaluated once
// per step on the start-of-step state and added to the inertial target
// q~ += h^2/m f; it has no gradient.
//
// Layout: the vertices of each face set (sorted, unique) and the faces are
// device arrays. Broad phase: per direction, a uniform grid over the other
// side's faces, rebuilt per evaluation, whose cells are at least as large as
// every face's box widened by the contact radius (so a face overlaps at most
// 2 x 2 x 2 cells); each cell lists the faces whose widened box overlaps it
// and the vertices inside it, both ascending. A vertex within the radius of
// a face lies in the face's widened box, hence in one of its cells: a
// particle's own pairs are tested over its cell's faces in face order, the
// reactions on a face over its cells' vertices merged in vertex order, so
// every sum keeps the reference's order and nothing of size V x F is stored
// (pair forces are recomputed where they are summed; no float atomics).
// A non-finite position leaves the grid off and every pair is tested.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dphd {

constexpr int kGridDim = 32;                               // cells per axis at most
constexpr int kGridCells = kGridDim * kGridDim * kGridDim;

struct MeshContactSide {
  int nv = 0;                 // vertices of this side's faces
  int nf = 0;                 // faces of the other side
  const int* verts = nullptr; // [nv] particle ids (sorted)
  const int* faces = nullptr; // [nf*3] the other side's faces
  double kn = 0.0;            // normal stiffness of this side's vertices
  double* grid = nullptr;     // [11]: origin (3), cell size, cells per axis (3), 1 = on, upper bounds (3)
  int* fcell_off = nullptr;   // [kGridCells + 1] faces per cell (CSR over cells)
  int* fcell = nullptr;       // [8 * nf]
  int* vcell_off = nullptr;   // [kGridCells + 1] vertices per cell
  int* vcell = nullptr;       // [nv] vertex slots
  int* cursor = nullptr;      // [2 * kGridCells] fill positions (faces, vertices)
  int* unsorted = nullptr;    // [8 * nf + nv] cell entries in arrival order (faces, then vertices)
  int* vcell_of = nullptr;    // [nv] cell of each vertex, -1 outside every widened face box
};

struct MeshContactDev {
  int n = 0;                  // particles
  double radius = 0.0, kf = 0.0, mu = 0.0;
  MeshContactSide ab, ba;     // A's vertices against B's faces, and B's against A's
  const int* slot_a = nullptr;  // [n] row of a particle among A's vertices, -1 if none
  const int* slot_b = nullptr;
  // Per particle, the (face, corner) entries of the faces of A (inc_a) and of
  // B (inc_b) it is a corner of, in face order (CSR).
  const int* inc_a_off = nullptr;
  const int* inc_a = nullptr;  // [2 * entries]
  const int* inc_b_off = nullptr;
  const int* inc_b = nullptr;
  double* force = nullptr;    // [n*3]
  int* pairs = nullptr;       // [2] pairs within the contact radius in the last evaluation (A-B, B-A)
};

// Evaluate the force of positions/velocities given as float32 (state arrays,
// pos64/vel64 null) or float64 (pos32/vel32 null) into mc.force; with
// `target` non-null, also target[i] += dt^2 / mass[vbody[i]] * force[i].
void launch_mesh_contact(const MeshContactDev& mc, const float* pos32, const float* vel32, const double* pos64,
                         const double* vel64, double* target, const int* vbody, const double* mass, double dt,
                         hipStream_t stream);

}  // namespace dphd

namespace dphd {

// Frictional boundary (the reference's FrictionalContact,
// src/newton_solver_diff_phd/_src/solvers/diff_phd/contact.py) of a
// single-body solver: planar or spherical obstacle, or the mesh boundary
// between two bodies. K registered vertices; A^{-1} restricted to them and
// its columns (the panel) are formed once per factor.
//
// Layout: panel [K][n] (column j = A^{-1} e_nodes[j], one scalar column per
// vertex, the same for every coordinate, so a thread over vertices reads it
// coalesced); blockT [K][K] with blockT[j*K + i] = A^{-1}(nodes[i], nodes[j]);
// per-candidate scalars and frames [K] / [K*3]; the coupled Newton system
// [A*A] row-major. The Newton step runs in one workgroup.
struct FrictionDev {
  int kind = 0;  // 0 planar, 1 spherical, 2 mesh
  int K = 0, n = 0;
  double friction = 0.0, restitution = 0.0;
  double* params = nullptr;  // [4] (mutable between steps)
  const int* nodes = nullptr;
  double* panel = nullptr;
  double* blockT = nullptr;
  double* delassus = nullptr;
  double* reg = nullptr;
  double *normal = nullptr, *t1 = nullptr, *t2 = nullptr;
  double* fl = nullptr;  // [K*2] tangential multipliers
  double* mult = nullptr;  // [K] normal multipliers (warm start across steps)
  double *gap = nullptr, *omega = nullptr, *diag = nullptr, *resid = nullptr, *rhs = nullptr;
  double *force = nullptr, *response = nullptr;
  int *order = nullptr, *slot = nullptr, *nactive = nullptr, *perm = nullptr;
  // The previous iteration's coupled candidates (ascending), kept by the
  // Newton kernel before it selects this iteration's: in the FP32-factor mode
  // only these and the current ones carry a nonzero f - f_prev.
  int *prev_order = nullptr, *nprev = nullptr;
  double *matrix = nullptr, *compact = nullptr, *work = nullptr;
  int* candidate = nullptr;
  // mesh boundary: faces = left then right; vertices below `split` touch the right faces
  double radius = 0.0;
  int split = 0, left_count = 0, face_count = 0;
  const int* faces = nullptr;
  double* surface = nullptr;  // [n*3] the iterate the boundary last saw (restitution)
  // Nearest-face search bounds: each side's faces in words of 64 (left words
  // [0, left_words), right words [left_words, words)), box [words][6] = the
  // corners' min xyz and max xyz at the positions the search reads.
  int left_words = 0, words = 0;
  double* face_box = nullptr;
  // FP32-factor mode: the solve is x + S^T S (b - A x), so the coupled forces
  // of the previous iteration (fprev [K*3]) enter b, only their change goes
  // through the panel, and freec [K*3] holds the free solution A^{-1} b at
  // the candidates (the iterate's value minus the previous forces' response).
  int defect = 0;
  double *fprev = nullptr, *freec = nullptr;
};

// Columns A^{-1} e_j for registered vertices [j0, j0 + 3) from out [n*3]
// (lane l of out = column j0 + l), then the K x K block once all are stored.
void launch_contact_units(const FrictionDev& fc, int j0, double* rhs, hipStream_t stream);
void launch_contact_store(const FrictionDev& fc, int j0, const double* out, hipStream_t stream);
void launch_contact_block(const FrictionDev& fc, hipStream_t stream);
// Step start: frames (and the mesh candidate filter), Delassus and scaling.
void launch_contact_begin(const FrictionDev& fc, const double* position, double dt, hipStream_t stream);
// One Fischer-Burmeister Newton step and friction update: x is the current
// iterate, xr holds A^{-1} b on entry and A^{-1}(b + J^T lambda) on exit.
// Skipped when *active is zero.
void launch_contact_iteration(const FrictionDev& fc, const double* x, double* xr, const double* position,
                              const int* active, hipStream_t stream);
// Post-solve restitution on x (positions/velocities of the step start).
void launch_contact_finish(const FrictionDev& fc, const double* position, const float* velocity, double dt, double* x,
                           hipStream_t stream);

}  // namespace dphd

namespace dphd {

// Woodbury projection of the adjoint off the active contact rows for the
// parameter gradients (the reference's Adjoint._woodbury,
// src/newton_solver_diff_phd/_src/solvers/diff_phd/adjoint.py:324-471,
// 1272-1354). Rows: omega n of each registered vertex (and, with friction,
// two tangent rows), valid where the recorded multiplier is positive and
// omega >= 0.5; selected where |J mu| >= max(1e-12, 1e-3 max |J mu|). Each
// selected row's column (A - K)^{-1} J^T is an adjoint fixed point run by the
// solver; then z = mu - W_T G^{-1} (J mu).
struct WoodburyDev {
  int K = 0, R = 0, n = 0;
  int tangents = 0;
  const int* nodes = nullptr;
  const double *mult = nullptr, *omega = nullptr, *normal = nullptr;  // the recorded step's contact state
  int* row_node = nullptr;     // [R]
  double* row_vec = nullptr;   // [R*3]
  int* row_valid = nullptr;    // [R]
  double* value = nullptr;     // [R] J mu
  int* order = nullptr;        // [R] selected rows, ascending
  int* count = nullptr;        // [1]
  double* columns = nullptr;   // [count][n*3]
  double* gram = nullptr;      // [R*R]
  double* coef = nullptr;      // [R]
  int* perm = nullptr;         // [R]
  double* work = nullptr;      // [R]
};

// Rows, values and the selection (one workgroup); `route` holds the routing
// guard's (scale, diverged) of body 0.
void launch_woodbury_rows(const WoodburyDev& w, const double* mu, const double* route, hipStream_t stream);
// Seed the batch of columns first .. first + C - 1 on the C bodies (n
// vertices each) of the batch solver: body k's right-hand side is J_r^T of
// selected row first + k, or off (inactive) past the selected count; its
// solution, mixed iterate and iteration state reset.
void launch_woodbury_batch_seed(const WoodburyDev& w, int first, int bodies, int n, double* rhs, double* sol,
                                double* xr, double* fscale, int* aiter, int* aactive, int* aa, hipStream_t stream);
// to |= from, then from = 0 (one status word into another).
void launch_status_merge(int* from, int* to, hipStream_t stream);
// G alpha = J mu and z = mu - sum_k alpha_k column_k, then the safe clip of z.
void launch_woodbury_project(const WoodburyDev& w, const double* mu, const uint8_t* free, double* z, int* status,
                             hipStream_t stream);

}  // namespace dphd
