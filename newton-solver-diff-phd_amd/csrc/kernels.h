// Device data layout and kernel launchers of the DiffPhD path on gfx950.
//
// Layout in HBM (all bodies packed; B = 1 is the single-model solver):
//   vertices   [N]      double3 AoS vectors (x, targets, boundary, p, rhs, ...)
//   elements   [E]      indices [E,K] int32, Tet4 poses [E,9] float32 (exact
//                       source of the gradients), Hex8 gradient table per body
//                       [B,8,8,3], per-element weights float64
//   factor     S        dense supernode blocks over each body's elimination-tree
//                       postorder (a row of S is one dense column range); both
//                       products read the same blocks; float64 or float32
//   history    [W+1,N]  Anderson ring: g = x_raw - x_prev and raw = x_prev + g
//   chunks     256-vertex / 256-thread element chunks never straddle bodies;
//              per-chunk partial sums are reduced per body in chunk order, so
//              every reduction is deterministic and float-atomic free.
//   work lists every iteration kernel runs a fixed grid whose workgroups loop
//              over the chunks / SpMV tasks of the bodies that are still
//              iterating (an LDS prefix over the per-body active flags), so a
//              converged body costs nothing in later iterations.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dphd {

constexpr int kBlock = 256;
constexpr int kWave = 64;  // wavefront width of gfx950
constexpr int kMaxWindow = 8;
constexpr int kGramEntries = kMaxWindow * (kMaxWindow + 1) / 2 + kMaxWindow;  // upper triangle + rhs

// The inverse factor on the device: dense supernode blocks over postorder
// slots. A block holds rows [row0, row0 + rows) by stored columns [col0,
// col0 + 4*quads), row-major from its value offset (zero where the factor is
// zero).

// Rows of one S-product work item (one wavefront) and columns of one
// S^T-product work item (one workgroup: 64 lanes x 4 columns).
constexpr int kRowsPerGroup = 8;
constexpr int kColsPerGroup = 256;
constexpr int kColWaves = 4;  // wavefronts of an S^T workgroup

// Blocks of at most kPackedQuads quads per row are packed: their items take
// kPackedPasses x (64 >> s) rows, a row over 2^s lanes (2^s >= quads).
constexpr int kPackedQuads = 16;
constexpr int kPackedPasses = 4;

// S product work item (one wavefront): rows out_row .. out_row + count - 1
// (count <= kRowsPerGroup, or the packed count) of one block, whose values
// start at `offset` with 4*quads values per row from column col0 (rows
// 4*stride values apart). Self-contained, so a wavefront needs one load per
// item.
struct RowItem {
  long long offset;
  int col0;
  int quads;
  int out_row;
  int count;
  int body;
  int part;    // -1: the item stores its rows; else its partial-sum slot (a chunk of a split item)
  int stride;  // quads per stored row of the item's block (= quads unless a chunk of a split item)
  int packed;  // 0: rows of kRowsPerGroup over all 64 lanes; s + 1: a packed narrow item, 2^s lanes per row
};

// Combine item of an S item split into column chunks: rows out_row ..
// out_row + count - 1 of body `body`, the chunks' partial sums in slots
// part0 .. part0 + parts - 1 (column order), each [32] = 8 rows x 4 values.
struct RowCombine {
  int out_row;
  int count;
  int body;
  int part0;
  int parts;
  int pad[3];
};

// Workgroups (one wavefront each) per vertex chunk of the adjoint operator
// product.
constexpr int kRowParts = 4;

// Quads per column chunk of the split S list; the default of the build.
constexpr int kSplitQuads = 256;
// Workgroups of the two combine kernels (each strides over its items).
constexpr int kCombineGrid = 256;

// One row of one block as an S^T column group sees it: the row's values
// start at `rowoff` with `quads` quads from column col0, and its S-product
// value is y[yrow].
struct ColEntry {
  long long rowoff;
  int col0;
  int quads;
  int yrow;
  int pad;
};

// S^T product work item (one workgroup): the 256 columns [c0, c0 + 256) of
// body `body` (columns from `end` on belong to the next body), over the rows
// entries[e0 .. e1) of the blocks that cover them (block by block, rows
// ascending). `gid` is the column group (the index of its whole-group item).
// A whole group, or a group's only slice, writes its columns (part < 0); a
// slice of a group cut into several writes its sums to partial slot `part`,
// and the group's combine item adds the slices in order.
struct ColGroup {
  int c0;
  int e0;
  int e1;
  int end;
  int body;
  int part;
  int gid;
  int pad;
};

// Combine item of a group cut into slices: columns [c0, c0 + 256) of body
// `body` (up to `end`), partial slots part0 .. part0 + parts - 1 in row order.
struct ColCombine {
  int c0;
  int end;
  int body;
  int part0;
  int parts;
  int gid;
  int pad[2];
};

// Rows of a slice of a column group cut for the sliced S^T list (entries);
// the default of the build.
constexpr int kSliceRows = 256;

struct Chunk {
  int body;
  int begin;
  int end;
};

// Pointers of one solver's device state (all owned by the solver).
struct DeviceState {
  // topology
  int B = 0, N = 0, E = 0, K = 4, Q = 1;
  int* idx = nullptr;          // [E*K]
  float* poses = nullptr;      // Tet4 [E*9]
  double* hexg = nullptr;      // Hex8 [B*192]
  uint8_t* pinmask = nullptr;  // [E] bit k: corner k is a Dirichlet vertex
  // Local-step staging: element chunk c's distinct vertices are
  // lv_list[lv_off[c] .. lv_off[c + 1]) (ascending); corner k of element e is
  // entry lidx[e*K + k] of its chunk's list; lv_max is the longest list.
  int* lv_off = nullptr;       // [nec + 1]
  int* pin_chunks = nullptr;   // element chunks with a pinned corner, ascending
  int npin_chunks = 0;
  int* lv_list = nullptr;
  uint16_t* lidx = nullptr;    // [E*K]
  int lv_max = 0;
  int* ebody = nullptr;        // [E]
  int* vbody = nullptr;        // [N]
  uint8_t* free = nullptr;     // [N]
  int* inc_off = nullptr;      // [N+1]
  int* inc_ent = nullptr;      // [E*K] entry = e*K + k
  // weights
  double *wc = nullptr, *wv = nullptr, *wn = nullptr;     // [E] per sample
  double *jwc = nullptr, *jwv = nullptr, *jwn = nullptr;  // [E] Jacobian weights
  double* mdir = nullptr;                                 // [E*3] or null
  double* mw = nullptr;                                   // [E] or null
  // per body
  double* inertia = nullptr;  // [B]
  double* mass = nullptr;     // [B]
  double* gravity = nullptr;  // [B*3]
  double* nh = nullptr;       // [B*3] mean mu, mean lambda, penalty
  double* svol = nullptr;     // [B] sample volume
  // inverse factor as supernode blocks over postorder slots (each body's
  // slots start at a multiple of 4; gap slots hold zeros)
  int slots = 0;              // postorder slots of all bodies
  int* gpost = nullptr;       // [N] original vertex -> slot
  int* gorig = nullptr;       // [slots] slot -> original vertex, -1 for a gap
  void* fval = nullptr;       // block values, double or float
  // Work lists: wavefront w of the S product takes the items
  // s_list[s_woff[w] .. s_woff[w + 1]), workgroup g of the S^T product the
  // groups t_list[t_woff[g] .. t_woff[g + 1]) (cost-balanced on the host).
  RowItem* s_items = nullptr;
  int s_nitems = 0;
  int *s_woff = nullptr, *s_list = nullptr;
  ColGroup* t_items = nullptr;
  int t_nitems = 0;
  int *t_woff = nullptr, *t_list = nullptr;
  ColEntry* t_entries = nullptr;  // block rows of the column groups
  int g_s = 0, g_t = 0;       // grids of the two products
  // The sliced S^T list (the same columns and rows, groups cut into slices of
  // at most `slice_rows` entries; items ts_items, per-workgroup lists
  // ts_woff/ts_list over a grid of g_ts), the combine items of the cut groups
  // (tc_items, g_tc workgroups of four) and their partial sums (t_part,
  // [slots][12][64] of the accumulation type).
  ColGroup* ts_items = nullptr;
  int ts_nitems = 0;
  int *ts_woff = nullptr, *ts_list = nullptr;
  ColCombine* tc_items = nullptr;
  int tc_nitems = 0;
  void* t_part = nullptr;
  int g_ts = 0, g_tc = 0;
  int slice_rows = kSliceRows;
  // The split S list: items wider than split_quads quads cut into column
  // chunks that write partial sums (ss_items, per-wavefront lists
  // ss_woff/ss_list over g_ss workgroups), the combine items of the split
  // ones (sc_items, one thread per value) and their partial sums (s_part,
  // [parts][32] of the accumulation type).
  RowItem* ss_items = nullptr;
  int ss_nitems = 0;
  int *ss_woff = nullptr, *ss_list = nullptr;
  RowCombine* sc_items = nullptr;
  int sc_nitems = 0;
  void* s_part = nullptr;
  int g_ss = 0;
  int split_quads = kSplitQuads;
  // Which lists run: t_mode[0] (S^T) and t_mode[2] (S) are written with
  // each active-body prefix (1 = sliced / split, when the active bodies hold
  // fewer than t_switch / s_switch column groups); t_mode[1] and t_mode[3]
  // are the constant choices for every body at once.
  int* t_mode = nullptr;
  int t_switch = 0;
  int s_switch = 0;
  bool f32 = false;           // FP32 factor values
  // constant matrix for the defect correction (per coordinate)
  int *a_off = nullptr, *a_col = nullptr;
  double* a_val = nullptr;
  // vectors [N*3]
  double *q0 = nullptr, *qt = nullptr, *xd = nullptr, *x = nullptr, *xr = nullptr, *p = nullptr;
  double* contrib = nullptr;  // [E*K*3]
  // Right-hand side of the S product and the S product itself over the
  // postorder slots, 4-wide ([(slots+4)*4], fourth lane zero) so that a quad
  // of columns reads 64 (float) or 128 (double) contiguous bytes.
  void *rhs4 = nullptr, *y4 = nullptr;
  // Anderson
  int W = 0;
  double *hg = nullptr, *hr = nullptr;  // [(W+1)*N*3]
  int* aa = nullptr;                    // [2*B*2] (iteration parity, body): count, start
  // chunks
  Chunk* vchunks = nullptr;
  int nvc = 0;
  int* vcb = nullptr;  // [B+1] vertex-chunk range of each body
  Chunk* echunks = nullptr;
  int nec = 0;
  int* ecb = nullptr;  // [B+1] element-chunk range of each body
  // Fixed grid sizes of the work-list kernels (each workgroup, or each
  // wavefront for the factor products, loops over the items of the
  // still-active bodies); multiples of 8 once above 8, so that every XCD
  // holds the same number of workgroups.
  int g_vert = 0, g_elem = 0;
  int g_rows = 0;  // the adjoint operator product's grid (one-wavefront workgroups, kRowParts per chunk)
  double* gram_part = nullptr;  // [nvc*kGramEntries]
  double* conv_part = nullptr;  // [nvc*4]
  double* clip_part2 = nullptr; // [nvc] squared norm of the adjoint's mixed iterate per chunk
  // Squared norm of the adjoint's raw map output per S^T column group
  // [t_nitems] (written by the S^T product's adjoint epilogue); the groups
  // of body b are tcb[b] .. tcb[b + 1] - 1.
  double* t_norm = nullptr;
  int* tcb = nullptr;
  // iteration control
  int* iter = nullptr;      // [B]
  int* active = nullptr;    // [B]
  // Work-list prefixes of the active bodies, [B + 1] each: pref[b] = the
  // chunks of the active bodies before b, pref[B] = their total; vertex
  // chunks and element chunks of the forward loop, vertex chunks of the
  // adjoint loop. Written by the seed and check kernels.
  int *pref_fv = nullptr, *pref_fe = nullptr, *pref_av = nullptr;
  int* status = nullptr;    // [1]
  int* any_active = nullptr;  // mapped host ints [2]: any body active, most iterations run
  // adjoint
  double* jac = nullptr;     // [E*Q*81]
  // Assembled adjoint operator K (minus A in the defect-correction mode):
  // 3x3 blocks in a sliced-ELL layout whose slices are the vertex chunks.
  // Slot s of row i of chunk c is block kbase[c] + s*256 + (i - begin); its
  // nine values are slot-major, kval[9*kbase[c] + (9*s + t)*256 + lane], so
  // a wavefront reads 64 consecutive values per (slot, entry).
  int* kwidth = nullptr;     // [nvc] slots per row (the chunk's longest row)
  int* kbase = nullptr;      // [nvc] first block of the chunk
  int* kcol = nullptr;       // [kblocks] column vertex, -1 = padding (rows end with padding)
  double* kdiag = nullptr;   // [kblocks] A_ij (defect mode only)
  double* ke = nullptr;      // [E*K*K*9] element blocks sum_s G_s^T J_s G_s
  // Element blocks summed into operator block b: kmap_src[kmap_off[b] ..
  // kmap_off[b + 1]) (indices into ke's blocks, in the order they are added).
  int *kmap_off = nullptr, *kmap_src = nullptr;
  double* kval = nullptr;    // [kblocks*9]
  // FP32-factor mode with a frictional boundary: each vertex's candidate
  // (-1 if none) and the previous iteration's coupled forces, which the
  // right-hand side b - A x includes.
  const int* fc_vslot = nullptr;
  const double* fc_fprev = nullptr;
  int kblocks = 0;
  int kmaxwidth = 0;         // the widest chunk's slots
  double *arhs = nullptr, *sol = nullptr;  // [N*3]
  int* aiter = nullptr;      // [B]
  int* aactive = nullptr;    // [B]
  // [B] safe-clip scale of the adjoint's last mixed iterate: mu = fscale * xr
  // until the next iteration (or the loop's end) stores it in sol.
  double* fscale = nullptr;
  double* route = nullptr;   // [B*2] routing scale, diverged flag
  double* gmat_tmp = nullptr;  // [E*2]
  double* gact_tmp = nullptr;  // [E]
  int* valid = nullptr;        // [2*B] material, activation gradients pass the guard
  int* ones = nullptr;         // [B] all ones (every body active)
  int* eoff = nullptr;         // [B+1] element offsets
  // options
  int material = 0;
  int has_muscles = 0;
  int integrate = 0;
  int defect = 0;  // FP32 factor: x_raw = x + S^T S (b - A x)
  // planar vertex energies: stiffness per vertex (0 = none) or null, and the
  // plane (unit normal, offset)
  double* planar_w = nullptr;
  double plane[4] = {0, 0, 1, 0};
};

struct StepScalars {
  double dt;
  double abs_tol, rel_tol;
  int max_iterations;
  int report = 1;  // an adjoint solve that hits max_iterations sets status 8
};

// Launchers (all asynchronous on `stream`).
void launch_begin_step(const DeviceState& d, const StepScalars& s, const float* q, const float* qd, const float* f,
                       const float* dirichlet, hipStream_t stream);
struct FrictionDev;
// One local/global iteration of every active body; with a frictional
// boundary `fc`, its Newton step turns the factor products' A^{-1} b into
// A^{-1}(b + J^T lambda) before the Anderson step.
void launch_forward_iteration(const DeviceState& d, const StepScalars& s, const double* targets,
                              const FrictionDev* fc, hipStream_t stream);
void launch_any_active(const DeviceState& d, const int* active, const int* iters, hipStream_t stream);
// Work-list prefixes after the active flags were set from outside a check
// kernel (adjoint: pref_av from aactive, else pref_fv and pref_fe from active).
void launch_active_prefix(const DeviceState& d, bool adjoint, hipStream_t stream);
void launch_write_state(const DeviceState& d, const StepScalars& s, float* q, float* qd, double* checkpoint,
                        hipStream_t stream);

void launch_adjoint_setup(const DeviceState& d, const StepScalars& s, const double* x, const double* targets,
                          const float* gq_out, const float* gqd_out, hipStream_t stream);
void launch_adjoint_iteration(const DeviceState& d, const StepScalars& s, hipStream_t stream);
// After an adjoint loop: mu = fscale * xr of every body into sol.
void launch_adjoint_finalize(const DeviceState& d, hipStream_t stream);
// Routing guard of a finished adjoint solve (scale and divergence flag per body).
void launch_adjoint_route(const DeviceState& d, hipStream_t stream);
// State, Dirichlet, material (gmat) and activation (gact, Tet4 fibres)
// gradients of a finished, routed adjoint solve; gmat/gact may be null. The
// parameter gradients read z (the Woodbury-projected adjoint with contact)
// or, when z is null, the adjoint itself.
void launch_adjoint_outputs(const DeviceState& d, const StepScalars& s, const double* x, const float* gqd_out,
                            float* gq_in, float* gqd_in, float* gf_in, double* gmat, double* gact, float* gdir,
                            const double* z, bool any_pinned, hipStream_t stream);
void launch_apply_inverse(const DeviceState& d, const double* rhs, double* out, hipStream_t stream);
// The S^T product alone (of y4 into xr) over the bodies flagged in `active`,
// from the list *mode selects (0 whole groups, 1 slices and combine): the
// launches the product tuning times.
void launch_cols(const DeviceState& d, const int* active, const int* mode, hipStream_t stream);
// The S product alone (of rhs4 into y4), from the list *mode selects (0
// whole items, 1 split items and their combine): timed by the tuning too.
void launch_rows(const DeviceState& d, const int* active, const int* mode, hipStream_t stream);
// A^{-1} rhs to FP64 accuracy in the FP32-factor mode: S^T S rhs, then
// `sweeps` refinements out += S^T S (rhs - A out) with the FP64 A (r, corr
// are [N*3] work vectors); the plain products in the FP64 mode.
void launch_refined_inverse(const DeviceState& d, const double* rhs, double* out, double* r, double* corr, int sweeps,
                            hipStream_t stream);
// Workgroups of the S and S^T product kernels the current device holds at
// once (the kernel's occupancy per compute unit times the compute units).
hipError_t product_capacity(bool f32, int* rows_blocks, int* cols_blocks);
void launch_hex_determinants(const DeviceState& d, const float* q, double* out, hipStream_t stream);

void launch_eval_projections(int count, const double* f, const double* dirs, double mu, double lame, double penalty,
                             double target, double* sigma, double* recon, double* rotation, double* volume,
                             double* nh, double* fibre, hipStream_t stream);
void launch_eval_jvps(int count, const double* f, const double* df, const double* dirs, double mu, double lame,
                      double target, double* rotation, double* volume, double* nh, double* fibre,
                      hipStream_t stream);

}  // namespace dphd
