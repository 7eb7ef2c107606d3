// Host orchestration of the DiffPhD path and the solver half of the C-ABI.
//
// A solver owns packed bodies, their factors and every device buffer. The
// local/global/Anderson loop has no host synchronization inside an
// iteration: `iteration_chunk` iterations are captured once into a hipGraph
// whose kernels are predicated on each body's device-side `active` flag, and
// the host polls one mapped flag between graph launches (HIP 7.2 has no
// conditional graph nodes). Reference orchestration: DP/solver.py:582-786
// (forward) and DP/adjoint.py:1054-1259 (backward).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/diffphd.h"
#include "contact.h"
#include "error.h"
#include "kernels.h"
#include "setup.h"

namespace dphd {
namespace {

#define HIP_CHECK(expr)                                                                            \
  do {                                                                                             \
    hipError_t _e = (expr);                                                                        \
    if (_e != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

template <typename T>
T* dalloc(std::vector<void*>& owned, size_t count) {
  void* p = nullptr;
  if (count == 0) count = 1;
  HIP_CHECK(hipMalloc(&p, count * sizeof(T)));
  owned.push_back(p);
  return static_cast<T*>(p);
}

template <typename T>
T* dupload(std::vector<void*>& owned, const std::vector<T>& host) {
  T* p = dalloc<T>(owned, host.size());
  if (!host.empty()) HIP_CHECK(hipMemcpy(p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

template <typename T>
T* dzeros(std::vector<void*>& owned, size_t count) {
  T* p = dalloc<T>(owned, count);
  HIP_CHECK(hipMemset(p, 0, std::max<size_t>(count, 1) * sizeof(T)));
  return p;
}

// Workgroups of a work-list kernel: enough to fill the device, never more
// than there are items, a multiple of the 8 XCDs once there are 8 or more.
int grid_for(int items) {
  int g = std::max(1, std::min(items, 2048));
  if (g >= 8) g &= ~7;
  return g;
}

// Workgroups of a factor-product grid per resident workgroup.
constexpr int kProductOversubscription = 3;

// The factor products' grids: up to 8192 workgroups (a multiple of 8).
int product_grid(int items) {
  int g = std::max(1, std::min(items, 8192));
  if (g >= 8) g &= ~7;
  return g;
}

// Device scratch of dphd_backward_static, for callers that cannot allocate:
// the adjoint's Jacobian, element blocks and assembled operator of one
// backward (every value is written before it is read, so nothing needs
// clearing). One array per GPU, held until the process ends, shared by every
// solver on that GPU: overlapping static backward calls on one GPU race.
constexpr size_t kScratchBytes = size_t{64} << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

// Device address of each GPU's copy of `scratch`, looked up once per device
// with that device current (the host-side name is not a device address).
constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
double* scratch_ptr[kMaxDevices] = {};
hipError_t scratch_status[kMaxDevices] = {};
bool scratch_looked_up[kMaxDevices] = {};

void scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (scratch_looked_up[device]) return;
  void* ptr = nullptr;
  scratch_status[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  scratch_ptr[device] = scratch_status[device] == hipSuccess ? static_cast<double*>(ptr) : nullptr;
  scratch_looked_up[device] = true;
}

double* scratch_address(int device) {
  if (device < 0 || device >= kMaxDevices) throw std::runtime_error("dphd_backward_static: device index out of range");
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_looked_up[device] || !scratch_ptr[device])
    throw std::runtime_error(std::string("dphd_backward_static: no device address of the static scratch: ") +
                             (scratch_looked_up[device] ? hipGetErrorString(scratch_status[device]) : "not looked up"));
  return scratch_ptr[device];
}

}  // namespace
}  // namespace dphd

struct dphd_solver {
  dphd_config cfg{};
  int device = 0;
  // topology / inputs (host)
  int B = 0, N = 0, E = 0, K = 4, Q = 1;
  std::vector<int> voff, eoff;
  std::vector<int> indices;
  std::vector<float> poses;
  std::vector<double> rest;
  std::vector<float> inv_mass;
  std::vector<double> gravity, density;
  std::vector<double> materials;
  std::vector<double> mdir_in, mstiff_in;
  std::vector<int> planar_idx;    // planar vertex energies (hex.py:204-243)
  std::vector<double> planar_w;
  double plane[4] = {0, 0, 1, 0};  // unit normal, offset
  // mesh contact state force (mesh_contact.py): radius, kn_left, kn_right, kf, mu
  bool has_mesh_contact = false;
  double mc_params[5] = {0, 0, 0, 0, 0};
  std::vector<int> mc_faces_a, mc_faces_b;
  dphd::MeshContactDev mc;
  // frictional boundary (contact.py FrictionalContact)
  bool has_friction = false;
  int fc_kind = 0;
  double fc_params[4] = {0, 0, 0, 0};
  std::vector<int> fc_nodes, fc_faces;
  double fc_friction = 0.0, fc_restitution = 0.0, fc_radius = 0.0;
  int fc_split = 0, fc_left = 0;
  dphd::FrictionDev fc;
  // Woodbury rows of the backward: the recorded step's contact state, and
  // buffers of the projection
  dphd::WoodburyDev wb;
  double *rec_mult = nullptr, *rec_omega = nullptr, *rec_normal = nullptr;
  double *wb_mu = nullptr, *wb_z = nullptr;
  int wb_capacity = 0;
  bool have_topology = false, have_particles = false, have_materials = false;
  // build products
  bool built = false;
  double dt = 0.0;
  dphd_info info{};
  std::vector<int64_t> body_nnz;
  bool any_pinned = false;
  std::vector<void*> owned;
  dphd::DeviceState d;
  int* any_host = nullptr;
  double* targets = nullptr;  // device copy of muscle targets [E]
  hipStream_t stream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  // A caller stream whose queued work (apply_inverse, mesh_contact_force)
  // shares this solver's buffers; the solver's next stream work waits for it.
  hipStream_t user_pending = nullptr;
  hipEvent_t ev_user = nullptr;
  // Iteration graphs of the forward, the adjoint and the Woodbury columns,
  // keyed on the tolerances they bake in.
  hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
  int key_iter[3] = {-1, -1, -1};
  double key_abs[3] = {-1, -1, -1}, key_rel[3] = {-1, -1, -1};
  int last_iters[3] = {0, 0, 0};
  // Per loop mode since the last build: iteration chunks launched, host waits.
  int64_t loop_stats[6] = {0, 0, 0, 0, 0, 0};
  // Every body a copy of body 0 (the factor is computed once).
  bool identical_bodies = false;
  // The Woodbury columns' solver: copies of this (single) body, one column
  // per copy, solved together; its checkpoint and fibre targets replicated.
  dphd_solver* wb_batch = nullptr;
  double* wb_ckpt = nullptr;
  double* wb_targets = nullptr;

  ~dphd_solver() { release(); }

  void release() {
    drop_woodbury_batch();
    for (auto& e : exec) {
      if (e) (void)hipGraphExecDestroy(e);
      e = nullptr;
    }
    wb_capacity = 0;
    for (void* p : owned) (void)hipFree(p);
    owned.clear();
    if (any_host) (void)hipHostFree(any_host);
    any_host = nullptr;
    built = false;
  }

  void drop_woodbury_batch();
};

namespace dphd {
namespace {

// Iterations per captured graph: the configured count, or by default 4 for
// the forward loop (a few iterations per step, so a long chunk runs mostly
// predicated off) and 8 for the adjoint loops (where shorter chunks cost
// more host waits than they save).
int chunk_size(const dphd_solver* s, int mode) {
  if (s->cfg.iteration_chunk > 0) return s->cfg.iteration_chunk;
  return mode == 0 ? 4 : 8;
}

void build_mesh_contact(dphd_solver* s);
void build_friction(dphd_solver* s);

// The sliced S^T list of `t_items` with slices of at most `rows` entries into
// d (items, per-workgroup lists, combine items, partial buffer, grids);
// device buffers go to `owned`.
void build_sliced(DeviceState& d, const std::vector<ColGroup>& t_items, int rows, int bodies, int cols_cap,
                  bool defect, std::vector<void*>& owned) {
  std::vector<ColGroup> ts_items;
  std::vector<ColCombine> tc_items;
  int parts = 0;
  for (int g = 0; g < static_cast<int>(t_items.size()); ++g) {
    const ColGroup& G = t_items[g];
    const int count = G.e1 - G.e0;
    const int pieces = std::max(1, (count + rows - 1) / rows);
    if (pieces == 1) {
      ts_items.push_back(G);
      continue;
    }
    tc_items.push_back(ColCombine{G.c0, G.end, G.body, parts, pieces, g, {0, 0}});
    for (int k = 0; k < pieces; ++k) {
      ColGroup it = G;
      it.e0 = G.e0 + static_cast<int>(static_cast<long long>(count) * k / pieces);
      it.e1 = G.e0 + static_cast<int>(static_cast<long long>(count) * (k + 1) / pieces);
      it.part = parts + k;
      ts_items.push_back(it);
    }
    parts += pieces;
  }
  d.slice_rows = rows;
  d.ts_items = dupload(owned, ts_items);
  d.ts_nitems = static_cast<int>(ts_items.size());
  d.tc_items = dupload(owned, tc_items);
  d.tc_nitems = static_cast<int>(tc_items.size());
  d.t_part = dalloc<unsigned char>(owned, static_cast<size_t>(std::max(parts, 1)) * 12 * kWave *
                                              (defect ? sizeof(float) : sizeof(double)));
  d.g_ts = std::min(product_grid(d.ts_nitems), cols_cap >= 8 ? cols_cap & ~7 : cols_cap);
  d.g_tc = std::max(1, std::min(kCombineGrid, (d.tc_nitems + kBlock / kWave - 1) / (kBlock / kWave)));
  std::vector<double> cost;
  std::vector<int> body, off, list;
  for (const ColGroup& g : ts_items) {
    const int batch = kColWaves * (defect ? 16 : 8);
    cost.push_back(1.0 + (g.e1 - g.e0 + batch - 1) / batch);
    body.push_back(g.body);
  }
  balanced_lists(cost, body, bodies, d.g_ts, 1, off, list, true);
  d.ts_woff = dupload(owned, off);
  d.ts_list = dupload(owned, list);
}

// The split S list of `s_items`: items wider than `quads` quads cut into
// column chunks of at most that many quads (partial sums, combined in column
// order) into d (items, per-wavefront lists, combine items, partial buffer,
// grid); device buffers go to `owned`.
void build_split(DeviceState& d, const std::vector<RowItem>& s_items, int quads, int bodies, int rows_cap,
                 std::vector<void*>& owned) {
  std::vector<RowItem> ss_items;
  std::vector<RowCombine> sc_items;
  int parts = 0;
  for (const RowItem& it : s_items) {
    const int pieces = std::max(1, (it.quads + quads - 1) / quads);
    if (pieces == 1) {
      ss_items.push_back(it);
      continue;
    }
    sc_items.push_back(RowCombine{it.out_row, it.count, it.body, parts, pieces, {0, 0, 0}});
    for (int k = 0; k < pieces; ++k) {
      const int q0 = static_cast<int>(static_cast<long long>(it.quads) * k / pieces);
      const int q1 = static_cast<int>(static_cast<long long>(it.quads) * (k + 1) / pieces);
      RowItem c = it;
      c.offset = it.offset + 4LL * q0;
      c.col0 = it.col0 + 4 * q0;
      c.quads = q1 - q0;
      c.part = parts + k;
      ss_items.push_back(c);
    }
    parts += pieces;
  }
  d.split_quads = quads;
  d.ss_items = dupload(owned, ss_items);
  d.ss_nitems = static_cast<int>(ss_items.size());
  d.sc_items = dupload(owned, sc_items);
  d.sc_nitems = static_cast<int>(sc_items.size());
  d.s_part = dalloc<unsigned char>(owned, static_cast<size_t>(std::max(parts, 1)) * 32 *
                                              (d.f32 ? sizeof(float) : sizeof(double)));
  d.g_ss = std::min(product_grid((d.ss_nitems + 3) / 4), rows_cap >= 8 ? rows_cap & ~7 : rows_cap);
  std::vector<double> cost;
  std::vector<int> body, off, list;
  for (const RowItem& it : ss_items) {
    cost.push_back(0.25 + (it.quads + kWave - 1) / kWave);
    body.push_back(it.body);
  }
  balanced_lists(cost, body, bodies, d.g_ss * (kBlock / kWave), kBlock / kWave, off, list, true);
  d.ss_woff = dupload(owned, off);
  d.ss_list = dupload(owned, list);
}

// The product shapes tuned on this solver's own factor. Activity levels:
// prefixes of the bodies holding all, 1/2, ..., 1/16 of the column groups
// (one level for a single body). For each product every candidate cut (S^T:
// rows per slice; S: quads per column chunk) is timed (HIP events, warm
// launches) at every level; the candidate with the least time summed over
// the levels is kept, and its list runs below the largest active-group count
// at which it beat the uncut list. A candidate must win by 5% to replace
// the default, so timing noise rarely changes the choice (and with it the
// products' summation order).
void tune_products(dphd_solver* s, const std::vector<RowItem>& s_items, const std::vector<ColGroup>& t_items,
                   int rows_cap, int cols_cap, bool defect) {
  DeviceState& d = s->d;
  const int B = s->B;
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<int> tcb(B + 1, 0);
  for (const ColGroup& g : t_items) ++tcb[g.body + 1];
  for (int b = 0; b < B; ++b) tcb[b + 1] += tcb[b];
  const int total = tcb[B];
  std::vector<void*> tmp;
  const int levels = B == 1 ? 1 : 5;
  std::vector<const int*> act(levels);
  std::vector<int> act_groups(levels);
  for (int lv = 0; lv < levels; ++lv) {
    const int target = std::max(1, total >> lv);
    std::vector<int> a(B, 0);
    int groups = 0;
    for (int b = 0; b < B && groups < target; ++b) {
      a[b] = 1;
      groups += tcb[b + 1] - tcb[b];
    }
    act[lv] = dupload(tmp, a);
    act_groups[lv] = groups;
  }
  const int* modes = dupload(tmp, std::vector<int>{0, 1});
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  hipStream_t st = s->stream;
  auto timed = [&](auto launch) {
    launch();
    HIP_CHECK(hipEventRecord(e0, st));
    for (int r = 0; r < 3; ++r) launch();
    HIP_CHECK(hipEventRecord(e1, st));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    return static_cast<double>(ms) / 3.0;
  };
  // One product: the uncut list against each candidate cut; returns the
  // chosen cut and its switch point, leaving the lists built for it.
  auto tune = [&](const int* candidates, int count, auto build, auto launch) {
    std::vector<double> whole(levels);
    for (int lv = 0; lv < levels; ++lv) whole[lv] = timed([&] { launch(act[lv], modes); });
    int best = candidates[0];
    double best_sum = 0.0;
    std::vector<double> best_cut;
    for (int k = 0; k < count; ++k) {
      build(candidates[k], tmp);
      std::vector<double> cut(levels);
      double sum = 0.0;
      for (int lv = 0; lv < levels; ++lv) sum += (cut[lv] = timed([&] { launch(act[lv], modes + 1); }));
      if (k == 0 || sum < 0.95 * best_sum) {
        best = candidates[k];
        best_sum = sum;
        best_cut = cut;
      }
    }
    int threshold = 0;
    for (int lv = 0; lv < levels; ++lv)
      if (best_cut[lv] < whole[lv]) threshold = std::max(threshold, lv == 0 ? INT_MAX : act_groups[lv] + 1);
    return std::make_pair(best, threshold);
  };
  const int slices[] = {kSliceRows, 64, 128, 512};
  const auto t_choice = tune(
      slices, 4, [&](int rows, std::vector<void*>& o) { build_sliced(d, t_items, rows, B, cols_cap, defect, o); },
      [&](const int* active, const int* mode) { launch_cols(d, active, mode, st); });
  const int chunks[] = {kSplitQuads, 128, 512};
  const auto s_choice = tune(
      chunks, 3, [&](int quads, std::vector<void*>& o) { build_split(d, s_items, quads, B, rows_cap, o); },
      [&](const int* active, const int* mode) { launch_rows(d, active, mode, st); });
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  HIP_CHECK(hipStreamSynchronize(st));
  for (void* p : tmp) (void)hipFree(p);
  build_sliced(d, t_items, t_choice.first, B, cols_cap, defect, s->owned);
  build_split(d, s_items, s_choice.first, B, rows_cap, s->owned);
  d.t_switch = t_choice.second;
  d.s_switch = s_choice.second;
  s->info.slice_rows = t_choice.first;
  s->info.slice_switch = t_choice.second;
  s->info.split_quads = s_choice.first;
  s->info.split_switch = s_choice.second;
  s->info.tune_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void build_solver(dphd_solver* s, double dt) {
  const auto t0 = std::chrono::steady_clock::now();
  HIP_CHECK(hipSetDevice(s->device));
  if (!s->have_topology || !s->have_particles || !s->have_materials)
    throw std::invalid_argument("dphd_build: bodies, topology, particles and materials must be set first");
  if (!std::isfinite(dt) || dt <= 0.0) throw std::invalid_argument("dt must be finite and positive");
  s->release();
  const int B = s->B, N = s->N, E = s->E, K = s->K, Q = s->Q;
  const bool hex = s->cfg.element_kind == DPHD_HEX8;
  const bool pd = s->cfg.material == DPHD_PD;
  // Validation of DP/solver.py:464-489.
  for (int e = 0; e < E; ++e) {
    const double mu = s->materials[2 * e], lam = s->materials[2 * e + 1];
    if (!std::isfinite(mu) || !std::isfinite(lam) || mu <= 0.0 || lam < 0.0)
      throw std::invalid_argument("DiffPhD requires positive shear modulus and nonnegative Lame lambda");
  }
  std::vector<uint8_t> free(N);
  for (int i = 0; i < N; ++i) {
    const float w = s->inv_mass[i];
    if (!std::isfinite(w) || w < 0.0f) throw std::invalid_argument("Particle inverse masses must be finite and nonnegative");
    free[i] = w > 0.0f;
  }
  s->any_pinned = false;
  for (int i = 0; i < N; ++i) s->any_pinned |= !free[i];
  std::vector<int> vbody(N), ebody(E);
  for (int b = 0; b < B; ++b) {
    bool any_free = false;
    for (int i = s->voff[b]; i < s->voff[b + 1]; ++i) {
      vbody[i] = b;
      any_free |= free[i] != 0;
    }
    if (!any_free) throw std::invalid_argument("SolverDiffPhD needs at least one dynamic particle per body");
    for (int e = s->eoff[b]; e < s->eoff[b + 1]; ++e) {
      ebody[e] = b;
      for (int k = 0; k < K; ++k) {
        const int v = s->indices[static_cast<size_t>(e) * K + k];
        if (v < s->voff[b] || v >= s->voff[b + 1])
          throw std::invalid_argument("element indices must reference vertices of their own body");
      }
    }
  }
  // Geometry per body.
  std::vector<double> grads;  // Tet4 [E*12]
  std::vector<double> hexg(hex ? static_cast<size_t>(B) * 192 : 0);
  std::vector<double> mean_volume(B), svol(B), mass(B), inertia(B);
  if (!hex) {
    grads.resize(static_cast<size_t>(E) * 12);
    std::vector<double> vols(E);
    tet_geometry(E, s->poses.data(), grads.data(), vols.data());
    for (int b = 0; b < B; ++b) {
      double total = 0.0;
      const int e0 = s->eoff[b], e1 = s->eoff[b + 1];
      for (int e = e0; e < e1; ++e) {
        if (!std::isfinite(vols[e]) || vols[e] <= 0.0)
          throw std::invalid_argument("tetrahedra must have nondegenerate rest shapes");
        total += vols[e];
      }
      mean_volume[b] = total / (e1 - e0);
    }
  } else {
    for (int b = 0; b < B; ++b) {
      std::string err;
      const int e0 = s->eoff[b], e1 = s->eoff[b + 1];
      const double mv = hex_mean_volume(e1 - e0, s->indices.data() + static_cast<size_t>(e0) * 8, s->rest.data(), &err);
      if (!err.empty()) throw std::invalid_argument(err);
      mean_volume[b] = mv;
      hex_gradients(std::pow(mv, 1.0 / 3.0), hexg.data() + 192 * static_cast<size_t>(b));
    }
  }
  for (int b = 0; b < B; ++b) {
    svol[b] = mean_volume[b] / Q;
    mass[b] = s->density[b] * mean_volume[b];
    inertia[b] = mass[b] / (dt * dt);
  }
  // Element weights (DP/kernels.py:112-133, DP/hex.py:95-112) and Jacobian weights.
  std::vector<double> wc(E, 0.0), wv(E, 0.0), wn(E, 0.0), jwc(E, 0.0), jwv(E, 0.0), jwn(E, 0.0), nh(3 * B);
  for (int b = 0; b < B; ++b) {
    const int e0 = s->eoff[b], e1 = s->eoff[b + 1];
    double smu = 0.0, slam = 0.0;
    for (int e = e0; e < e1; ++e) {
      smu += s->materials[2 * e];
      slam += s->materials[2 * e + 1];
    }
    const double mmu = smu / (e1 - e0), mlam = slam / (e1 - e0);
    nh[3 * b] = mmu;
    nh[3 * b + 1] = mlam;
    nh[3 * b + 2] = 2.0 * mmu;
    for (int e = e0; e < e1; ++e) {
      const double mu = s->materials[2 * e], lam = s->materials[2 * e + 1];
      if (pd) {
        wc[e] = 2.0 * mu * svol[b];
        wv[e] = lam * svol[b];
      } else {
        wn[e] = 2.0 * mu * svol[b];
      }
      if (s->cfg.jacobian_mode == 0) {
        jwc[e] = pd ? 2.0 * mmu * svol[b] : 0.0;
        jwv[e] = pd ? mlam * svol[b] : 0.0;
        jwn[e] = pd ? 0.0 : 2.0 * mmu * svol[b];
      } else {
        jwc[e] = wc[e];
        jwv[e] = wv[e];
        jwn[e] = wn[e];
      }
    }
  }
  // Fibres (DP/solver.py:429-462).
  const bool muscles = !s->mstiff_in.empty();
  std::vector<double> mdir, mw;
  if (muscles) {
    mdir.assign(static_cast<size_t>(E) * 3, 0.0);
    mw.assign(E, 0.0);
    for (int e = 0; e < E; ++e) {
      const double k = s->mstiff_in[e];
      if (!(k > 0.0)) continue;
      const double* m = &s->mdir_in[3 * static_cast<size_t>(e)];
      const double len = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
      if (len <= 2.220446049250313e-16) throw std::invalid_argument("Active muscle directions must have nonzero length");
      for (int a = 0; a < 3; ++a) mdir[3 * static_cast<size_t>(e) + a] = m[a] / len;
      mw[e] = k * svol[ebody[e]];
    }
  }
  // Assembly, factor and device layout per body, in parallel over bodies.
  std::vector<FactorLayout> layouts(B);
  std::vector<int64_t> factor_entries(B, 0);
  std::vector<Csr> amats(B);
  std::vector<std::string> errors(B);
  {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nthreads = static_cast<int>(std::min<unsigned>(hw, static_cast<unsigned>(B)));
    std::vector<std::thread> pool;
    std::atomic<int> next{0};
    for (int t = 0; t < nthreads; ++t)
      pool.emplace_back([&]() {
        for (int b = next++; b < B; b = next++) {
          if (s->identical_bodies && b > 0) continue;  // copies of body 0, filled in below
          try {
            const int v0 = s->voff[b], v1 = s->voff[b + 1], e0 = s->eoff[b], e1 = s->eoff[b + 1];
            std::vector<int> local(static_cast<size_t>(e1 - e0) * K);
            for (size_t t2 = 0; t2 < local.size(); ++t2) local[t2] = s->indices[static_cast<size_t>(e0) * K + t2] - v0;
            std::vector<double> weights(e1 - e0);
            for (int e = e0; e < e1; ++e) weights[e - e0] = (wc[e] + wv[e]) + wn[e];
            AssemblyInput in;
            in.vertices = v1 - v0;
            in.elements = e1 - e0;
            in.corners = K;
            in.samples = Q;
            in.indices = local.data();
            in.grads = hex ? hexg.data() + 192 * static_cast<size_t>(b) : grads.data() + static_cast<size_t>(e0) * 12;
            in.shared_grads = hex;
            in.weights = weights.data();
            in.fibre_dir = muscles ? mdir.data() + 3 * static_cast<size_t>(e0) : nullptr;
            in.fibre_w = muscles ? mw.data() + e0 : nullptr;
            in.free = free.data() + v0;
            in.inertia = inertia[b];
            std::vector<int> plane_local;
            std::vector<double> plane_w;
            for (size_t j = 0; j < s->planar_idx.size(); ++j)
              if (s->planar_idx[j] >= v0 && s->planar_idx[j] < v1) {
                plane_local.push_back(s->planar_idx[j] - v0);
                plane_w.push_back(s->planar_w[j]);
              }
            in.plane_vertices = plane_local.data();
            in.plane_stiffness = plane_w.data();
            in.plane_count = static_cast<int>(plane_local.size());
            amats[b] = assemble_matrix(in);
            const Factor factor = build_factor(amats[b]);
            factor_entries[b] = static_cast<int64_t>(factor.s.col.size());
            layouts[b] = factor_layout(amats[b], factor);
          } catch (const std::exception& ex) {
            errors[b] = ex.what();
          }
        }
      });
    for (auto& th : pool) th.join();
  }
  for (int b = 0; b < B; ++b)
    if (!errors[b].empty()) throw std::invalid_argument(errors[b]);
  if (s->identical_bodies)
    for (int b = 1; b < B; ++b) {
      amats[b] = amats[0];
      factor_entries[b] = factor_entries[0];
      layouts[b] = layouts[0];
    }
  // Global factor blocks: body b's postorder slots are offset by slot0[b]
  // (a multiple of 4), its values by the values before it.
  int64_t znnz = 0, annz = 0;
  s->body_nnz.assign(B, 0);
  std::vector<int> slot0(B + 1, 0);
  for (int b = 0; b < B; ++b) {
    s->body_nnz[b] = factor_entries[b];
    znnz += factor_entries[b];
    annz += static_cast<int64_t>(amats[b].col.size());
    if (static_cast<int64_t>(slot0[b]) + layouts[b].slots > INT32_MAX - 8)
      throw std::invalid_argument("postorder slots exceed int32 indexing");
    slot0[b + 1] = slot0[b] + layouts[b].slots;
  }
  const int slots = slot0[B];
  std::vector<int> gpost(N), gorig(slots, -1);
  std::vector<double> fval;
  std::vector<RowItem> s_items;
  std::vector<ColGroup> t_items;
  std::vector<ColEntry> t_entries;
  std::vector<int> tcb(B + 1, 0);  // column groups of each body
  for (int b = 0; b < B; ++b) {
    tcb[b] = static_cast<int>(t_items.size());
    const int v0 = s->voff[b], n = s->voff[b + 1] - v0, c0 = slot0[b];
    FactorLayout& L = layouts[b];
    for (int o = 0; o < n; ++o) {
      gpost[v0 + o] = c0 + L.post[o];
      gorig[c0 + L.post[o]] = v0 + o;
    }
    const long long voff_b = static_cast<long long>(fval.size());
    for (const FactorBlock& fb : L.blocks) {
      const long long width = 4 * static_cast<long long>(fb.quads);
      // Narrow blocks: 2^sh lanes per row, kPackedPasses x (64 >> sh) rows per item.
      int sh = 0;
      while ((1 << sh) < fb.quads) ++sh;
      const bool packed = fb.quads <= kPackedQuads;
      const int per = packed ? kPackedPasses * (kWave >> sh) : kRowsPerGroup;
      for (int r = 0; r < fb.rows; r += per) {
        RowItem it{};
        it.offset = voff_b + fb.offset + width * r;
        it.col0 = c0 + fb.col0;
        it.quads = fb.quads;
        it.out_row = c0 + fb.row0 + r;
        it.count = std::min(per, fb.rows - r);
        it.body = b;
        it.part = -1;
        it.stride = fb.quads;
        it.packed = packed ? sh + 1 : 0;
        s_items.push_back(it);
      }
    }
    fval.insert(fval.end(), L.val.begin(), L.val.end());
    // Column groups of 256 slots and, per group, the rows (from the group's
    // first column on) of the blocks that cover it, block by block.
    const int groups = (L.slots + kColsPerGroup - 1) / kColsPerGroup;
    std::vector<std::vector<int>> cover(groups);
    for (size_t k = 0; k < L.blocks.size(); ++k) {
      const FactorBlock& fb = L.blocks[k];
      const int g0 = fb.col0 / kColsPerGroup, g1 = (fb.row0 + fb.rows - 1) / kColsPerGroup;
      for (int g = g0; g <= g1; ++g) cover[g].push_back(static_cast<int>(k));
    }
    for (int g = 0; g < groups; ++g) {
      const int gc0 = g * kColsPerGroup;  // body-local first column of the group
      const int e0 = static_cast<int>(t_entries.size());
      for (const int k : cover[g]) {
        const FactorBlock& fb = L.blocks[k];
        const long long width = 4 * static_cast<long long>(fb.quads);
        // Rows before the group's first column hold none of its columns.
        for (int r = std::max(fb.row0, gc0); r < fb.row0 + fb.rows; ++r) {
          ColEntry en{};
          en.rowoff = voff_b + fb.offset + width * (r - fb.row0);
          en.col0 = c0 + fb.col0;
          en.quads = fb.quads;
          en.yrow = c0 + r;
          t_entries.push_back(en);
        }
      }
      if (t_entries.size() > static_cast<size_t>(INT32_MAX)) throw std::invalid_argument("too many factor rows");
      t_items.push_back(ColGroup{c0 + gc0, e0, static_cast<int>(t_entries.size()), c0 + L.slots, b, -1,
                                 static_cast<int>(t_items.size()), 0});
    }
    L = FactorLayout{};
  }
  tcb[B] = static_cast<int>(t_items.size());
  std::vector<int> a_off, a_col;
  std::vector<double> a_val;
  const bool defect = s->cfg.factor_precision == 32;
  if (defect) {
    a_off.assign(N + 1, 0);
    for (int b = 0; b < B; ++b) {
      const int v0 = s->voff[b];
      const Csr& a = amats[b];
      for (int r = 0; r < a.n; ++r) {
        for (int p = a.off[r]; p < a.off[r + 1]; ++p) {
          a_col.push_back(a.col[p] + v0);
          a_val.push_back(a.val[p]);
        }
        a_off[v0 + r + 1] = static_cast<int>(a_col.size());
      }
    }
  }
  // Chunks and incidence.
  std::vector<Chunk> vchunks, echunks;
  std::vector<int> vcb(B + 1, 0), ecb(B + 1, 0);
  const int eper = hex ? kBlock / 8 : kBlock;
  for (int b = 0; b < B; ++b) {
    vcb[b] = static_cast<int>(vchunks.size());
    ecb[b] = static_cast<int>(echunks.size());
    for (int i = s->voff[b]; i < s->voff[b + 1]; i += kBlock) vchunks.push_back(Chunk{b, i, std::min(i + kBlock, s->voff[b + 1])});
    for (int e = s->eoff[b]; e < s->eoff[b + 1]; e += eper) echunks.push_back(Chunk{b, e, std::min(e + eper, s->eoff[b + 1])});
  }
  vcb[B] = static_cast<int>(vchunks.size());
  ecb[B] = static_cast<int>(echunks.size());
  if (B + 1 > 8192) throw std::invalid_argument("at most 8191 bodies per solver (LDS work-list prefix)");
  std::vector<int> inc_off, inc_ent;
  build_incidence(N, E, K, s->indices.data(), inc_off, inc_ent);
  std::vector<uint8_t> pinmask(E, 0);
  for (int e = 0; e < E; ++e)
    for (int k = 0; k < K; ++k)
      if (!free[s->indices[static_cast<size_t>(e) * K + k]]) pinmask[e] |= static_cast<uint8_t>(1u << k);
  // Local-step staging: the distinct vertices of each element chunk
  // (ascending) and each corner's position in its chunk's list. A chunk
  // holds at most 4 kBlock corners, so a position fits 16 bits.
  std::vector<int> lv_off(1, 0), lv_list;
  std::vector<uint16_t> lidx(static_cast<size_t>(E) * K);
  int lv_max = 0;
  for (const Chunk& c : echunks) {
    const size_t a0 = static_cast<size_t>(c.begin) * K, a1 = static_cast<size_t>(c.end) * K;
    std::vector<int> verts(s->indices.begin() + a0, s->indices.begin() + a1);
    std::sort(verts.begin(), verts.end());
    verts.erase(std::unique(verts.begin(), verts.end()), verts.end());
    for (size_t a = a0; a < a1; ++a)
      lidx[a] = static_cast<uint16_t>(std::lower_bound(verts.begin(), verts.end(), s->indices[a]) - verts.begin());
    lv_list.insert(lv_list.end(), verts.begin(), verts.end());
    lv_off.push_back(static_cast<int>(lv_list.size()));
    lv_max = std::max(lv_max, static_cast<int>(verts.size()));
  }
  // Sliced-ELL pattern of the adjoint operator: row i holds itself and every
  // vertex that shares an element with it, ascending.
  const int nvc_host = static_cast<int>(vchunks.size());
  std::vector<int> kwidth(nvc_host), kbase(nvc_host);
  std::vector<int> kcol;
  std::vector<double> kdiag;
  // Per operator block (i, j), in block order: the element blocks K_e[k][l]
  // (index (e*K + k)*K + l) with idx[e][k] = i and idx[e][l] = j, in the
  // incidence order of i and then l ascending.
  std::vector<int> kmap_off(1, 0), kmap_src;
  if (static_cast<int64_t>(E) * K * K > INT32_MAX) throw std::invalid_argument("element blocks exceed int32 indexing");
  {
    std::vector<std::vector<int>> pattern(kBlock);
    int64_t blocks = 0;
    for (int ci = 0; ci < nvc_host; ++ci) {
      const Chunk c = vchunks[ci];
      int width = 0;
      for (int i = c.begin; i < c.end; ++i) {
        std::vector<int>& row = pattern[i - c.begin];
        row.assign(1, i);
        for (int a = inc_off[i]; a < inc_off[i + 1]; ++a) {
          const int e = inc_ent[a] / K;
          for (int l = 0; l < K; ++l) row.push_back(s->indices[static_cast<size_t>(e) * K + l]);
        }
        std::sort(row.begin(), row.end());
        row.erase(std::unique(row.begin(), row.end()), row.end());
        width = std::max(width, static_cast<int>(row.size()));
      }
      if (blocks + static_cast<int64_t>(width) * kBlock > INT32_MAX / 9)
        throw std::invalid_argument("adjoint operator exceeds int32 block indexing");
      kwidth[ci] = width;
      kbase[ci] = static_cast<int>(blocks);
      blocks += static_cast<int64_t>(width) * kBlock;
      kcol.resize(blocks, -1);
      if (defect) kdiag.resize(blocks, 0.0);
      // Element blocks of the chunk's operator blocks (local block index
      // t * kBlock + lane), counted, then placed.
      std::vector<int> local(static_cast<size_t>(width) * kBlock + 1, 0);
      for (int pass = 0; pass < 2; ++pass) {
        for (int i = c.begin; i < c.end; ++i) {
          const std::vector<int>& row = pattern[i - c.begin];
          const size_t lane = static_cast<size_t>(i - c.begin);
          for (int a = inc_off[i]; a < inc_off[i + 1]; ++a) {
            const int ent = inc_ent[a], e = ent / K;
            for (int l = 0; l < K; ++l) {
              const int j = s->indices[static_cast<size_t>(e) * K + l];
              const size_t t = static_cast<size_t>(std::lower_bound(row.begin(), row.end(), j) - row.begin());
              const size_t lb = t * kBlock + lane;
              if (pass == 0)
                ++local[lb + 1];
              else
                kmap_src[kmap_off[kbase[ci] + lb] + local[lb]++] = ent * K + l;
            }
          }
        }
        if (pass == 0) {
          for (size_t lb = 0; lb + 1 < local.size(); ++lb) kmap_off.push_back(kmap_off.back() + local[lb + 1]);
          kmap_src.resize(kmap_off.back());
          std::fill(local.begin(), local.end(), 0);
        }
      }
      for (int i = c.begin; i < c.end; ++i) {
        const std::vector<int>& row = pattern[i - c.begin];
        const size_t lane = static_cast<size_t>(i - c.begin);
        for (size_t t = 0; t < row.size(); ++t) kcol[kbase[ci] + t * kBlock + lane] = row[t];
        if (defect) {
          for (int p = a_off[i]; p < a_off[i + 1]; ++p) {
            const auto it = std::lower_bound(row.begin(), row.end(), a_col[p]);
            if (it == row.end() || *it != a_col[p]) throw std::logic_error("matrix entry outside the element pattern");
            kdiag[kbase[ci] + static_cast<size_t>(it - row.begin()) * kBlock + lane] = a_val[p];
          }
        }
      }
    }
  }

  // Upload.
  auto& o = s->owned;
  DeviceState& d = s->d;
  d = DeviceState{};
  d.B = B;
  d.N = N;
  d.E = E;
  d.K = K;
  d.Q = Q;
  d.idx = dupload(o, s->indices);
  if (!hex) d.poses = dupload(o, s->poses);
  if (hex) d.hexg = dupload(o, hexg);
  d.pinmask = dupload(o, pinmask);
  d.lv_off = dupload(o, lv_off);
  d.lv_list = dupload(o, lv_list);
  d.lidx = dupload(o, lidx);
  d.lv_max = lv_max;
  {
    std::vector<int> pin_chunks;
    for (int ci = 0; ci < static_cast<int>(echunks.size()); ++ci)
      for (int e = echunks[ci].begin; e < echunks[ci].end; ++e)
        if (pinmask[e]) {
          pin_chunks.push_back(ci);
          break;
        }
    d.pin_chunks = dupload(o, pin_chunks);
    d.npin_chunks = static_cast<int>(pin_chunks.size());
  }
  d.ebody = dupload(o, ebody);
  d.vbody = dupload(o, vbody);
  d.free = dupload(o, free);
  d.inc_off = dupload(o, inc_off);
  d.inc_ent = dupload(o, inc_ent);
  d.wc = dupload(o, wc);
  d.wv = dupload(o, wv);
  d.wn = dupload(o, wn);
  d.jwc = dupload(o, jwc);
  d.jwv = dupload(o, jwv);
  d.jwn = dupload(o, jwn);
  if (muscles) {
    d.mdir = dupload(o, mdir);
    d.mw = dupload(o, mw);
  }
  d.inertia = dupload(o, inertia);
  d.mass = dupload(o, mass);
  d.gravity = dupload(o, s->gravity);
  d.nh = dupload(o, nh);
  d.svol = dupload(o, svol);
  d.slots = slots;
  d.gpost = dupload(o, gpost);
  d.gorig = dupload(o, gorig);
  d.f32 = defect;
  const int64_t factor_values = static_cast<int64_t>(fval.size());
  if (defect) {
    const std::vector<float> v32(fval.begin(), fval.end());
    d.fval = dupload(o, v32);
  } else {
    d.fval = dupload(o, fval);
  }
  std::vector<double>().swap(fval);
  d.s_items = dupload(o, s_items);
  d.s_nitems = static_cast<int>(s_items.size());
  d.t_items = dupload(o, t_items);
  d.t_nitems = static_cast<int>(t_items.size());
  d.t_entries = dupload(o, t_entries);
  int cols_capacity = 0, rows_capacity = 0;
  {
    // Grids of the products: kProductOversubscription times the workgroups
    // the device holds at once. The static lists are then shorter than one
    // per resident workgroup and the dispatcher starts the next workgroup
    // wherever one finished, which evens out the error of the lists' cost
    // estimates (all-active launch 184 -> 178 us on config 4; 2x to 6x
    // measured alike, 1x slower).
    int rows_cap = 0, cols_cap = 0;
    HIP_CHECK(product_capacity(defect, &rows_cap, &cols_cap));
    rows_cap *= kProductOversubscription;
    cols_cap *= kProductOversubscription;
    d.g_s = std::min(product_grid((d.s_nitems + 3) / 4), rows_cap >= 8 ? rows_cap & ~7 : rows_cap);
    d.g_t = std::min(product_grid(d.t_nitems), cols_cap >= 8 ? cols_cap & ~7 : cols_cap);
    // The sliced list while the active bodies' groups cannot fill the device
    // (replaced by the tuned shapes unless slice_rows is fixed).
    cols_capacity = cols_cap;
    rows_capacity = rows_cap;
    d.t_switch = cols_cap / kProductOversubscription;
    d.s_switch = cols_cap / kProductOversubscription;
    build_sliced(d, t_items, s->cfg.slice_rows > 0 ? s->cfg.slice_rows : kSliceRows, B, cols_cap, defect, o);
    build_split(d, s_items, kSplitQuads, B, rows_cap, o);
  }
  // The list choices every product launch reads: allocated here, before any
  // product runs (the frictional boundary's setup applies the inverse);
  // the tuning at the end of the build rewrites the all-bodies entries.
  {
    const int all_cut = static_cast<int>(t_items.size()) < cols_capacity / kProductOversubscription ? 1 : 0;
    d.t_mode = dupload(o, std::vector<int>{0, all_cut, 0, all_cut});
  }
  {
    // Cost of an item in memory round trips of its wavefront: an S item
    // streams its rows in passes of 64 quads, a wavefront of an S^T group
    // takes its share of the group's rows in batches of 16 (FP32 values) or
    // 8 (FP64); plus the item's reduction and store.
    std::vector<double> cost;
    std::vector<int> body, off, list;
    for (const RowItem& it : s_items) {
      cost.push_back(0.25 + (it.quads + kWave - 1) / kWave);
      body.push_back(it.body);
    }
    balanced_lists(cost, body, B, d.g_s * (kBlock / kWave), kBlock / kWave, off, list);
    d.s_woff = dupload(o, off);
    d.s_list = dupload(o, list);
    cost.clear();
    body.clear();
    for (const ColGroup& g : t_items) {
      const int batch = kColWaves * (defect ? 16 : 8);
      cost.push_back(1.0 + (g.e1 - g.e0 + batch - 1) / batch);
      body.push_back(g.body);
    }
    balanced_lists(cost, body, B, d.g_t, 1, off, list);
    d.t_woff = dupload(o, off);
    d.t_list = dupload(o, list);
  }
  if (defect) {
    d.a_off = dupload(o, a_off);
    d.a_col = dupload(o, a_col);
    d.a_val = dupload(o, a_val);
    d.defect = 1;
  }
  // 4-wide S-product vectors over the postorder slots (float in the FP32
  // mode); gap slots stay zero.
  const size_t v4bytes = (static_cast<size_t>(slots) + 4) * 4 * (d.f32 ? sizeof(float) : sizeof(double));
  d.rhs4 = dzeros<unsigned char>(o, v4bytes);
  d.y4 = dzeros<unsigned char>(o, v4bytes);
  const size_t v3 = static_cast<size_t>(N) * 3;
  d.q0 = dzeros<double>(o, v3);
  d.qt = dzeros<double>(o, v3);
  d.xd = dzeros<double>(o, v3);
  d.x = dzeros<double>(o, v3);
  d.xr = dzeros<double>(o, v3);
  d.p = dzeros<double>(o, v3);
  d.contrib = dzeros<double>(o, static_cast<size_t>(E) * K * 3);
  d.W = s->cfg.anderson_window;
  if (d.W > 0) {
    d.hg = dzeros<double>(o, static_cast<size_t>(d.W + 1) * v3);
    d.hr = dzeros<double>(o, static_cast<size_t>(d.W + 1) * v3);
  }
  d.aa = dzeros<int>(o, 4 * static_cast<size_t>(B));
  d.vchunks = dupload(o, vchunks);
  d.nvc = static_cast<int>(vchunks.size());
  d.vcb = dupload(o, vcb);
  d.echunks = dupload(o, echunks);
  d.nec = static_cast<int>(echunks.size());
  d.ecb = dupload(o, ecb);
  d.g_vert = grid_for(d.nvc);
  d.g_rows = std::max(1, std::min(kRowParts * d.nvc, 8192));
  if (d.g_rows >= 8) d.g_rows &= ~7;
  d.g_elem = grid_for(d.nec);
  d.gram_part = dzeros<double>(o, static_cast<size_t>(d.nvc) * kGramEntries);
  d.conv_part = dzeros<double>(o, static_cast<size_t>(d.nvc) * 4);
  d.clip_part2 = dzeros<double>(o, d.nvc);
  d.t_norm = dzeros<double>(o, std::max<size_t>(t_items.size(), 1));
  d.tcb = dupload(o, tcb);
  d.iter = dzeros<int>(o, B);
  d.active = dzeros<int>(o, B);
  d.status = dzeros<int>(o, 1);
  d.arhs = dzeros<double>(o, v3);
  d.sol = dzeros<double>(o, v3);
  d.aiter = dzeros<int>(o, B);
  d.aactive = dzeros<int>(o, B);
  d.pref_fv = dzeros<int>(o, B + 1);
  d.pref_fe = dzeros<int>(o, B + 1);
  d.pref_av = dzeros<int>(o, B + 1);
  d.fscale = dzeros<double>(o, static_cast<size_t>(B));
  d.route = dzeros<double>(o, 2 * static_cast<size_t>(B));
  d.gmat_tmp = dzeros<double>(o, 2 * static_cast<size_t>(E));
  d.gact_tmp = dzeros<double>(o, static_cast<size_t>(E));
  d.valid = dzeros<int>(o, 2 * static_cast<size_t>(B));
  d.ones = dupload(o, std::vector<int>(B, 1));
  d.eoff = dupload(o, s->eoff);
  d.jac = nullptr;  // allocated on the first backward, with ke and kval
  d.kwidth = dupload(o, kwidth);
  d.kbase = dupload(o, kbase);
  d.kcol = dupload(o, kcol);
  d.kmap_off = dupload(o, kmap_off);
  d.kmap_src = dupload(o, kmap_src);
  if (defect) d.kdiag = dupload(o, kdiag);
  d.kblocks = static_cast<int>(kcol.size());
  d.kmaxwidth = std::max(1, *std::max_element(kwidth.begin(), kwidth.end()));
  d.material = pd ? 0 : 1;
  d.has_muscles = muscles ? 1 : 0;
  if (!s->planar_idx.empty()) {
    std::vector<double> pw(N, 0.0);
    for (size_t j = 0; j < s->planar_idx.size(); ++j) pw[s->planar_idx[j]] = s->planar_w[j];
    d.planar_w = dupload(o, pw);
    for (int t = 0; t < 4; ++t) d.plane[t] = s->plane[t];
  }
  d.integrate = s->cfg.integrate_pinned_velocities ? 1 : 0;
  s->targets = muscles ? dupload(o, std::vector<double>(E, 1.0)) : nullptr;
  if (s->has_mesh_contact) build_mesh_contact(s);
  if (s->has_friction) build_friction(s);
  void* host = nullptr;
  HIP_CHECK(hipHostMalloc(&host, 2 * sizeof(int), hipHostMallocMapped));
  s->any_host = static_cast<int*>(host);
  s->any_host[0] = 0;
  s->any_host[1] = 0;
  void* dev = nullptr;
  HIP_CHECK(hipHostGetDevicePointer(&dev, host, 0));
  d.any_active = static_cast<int*>(dev);
  if (!s->stream) HIP_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  if (!s->ev_in) HIP_CHECK(hipEventCreateWithFlags(&s->ev_in, hipEventDisableTiming));
  if (!s->ev_out) HIP_CHECK(hipEventCreateWithFlags(&s->ev_out, hipEventDisableTiming));
  if (!s->ev_user) HIP_CHECK(hipEventCreateWithFlags(&s->ev_user, hipEventDisableTiming));
  HIP_CHECK(hipDeviceSynchronize());
  s->dt = dt;
  s->built = true;
  for (int m = 0; m < 3; ++m) {
    s->key_iter[m] = -1;
    s->last_iters[m] = 0;
  }
  // Jacobian, element operator blocks and the assembled operator.
  const int64_t jac_bytes = static_cast<int64_t>(E) * Q * 81 * 8 + static_cast<int64_t>(E) * K * K * 9 * 8 +
                            static_cast<int64_t>(d.kblocks) * 9 * 8;
  s->info.bodies = B;
  s->info.vertices = N;
  s->info.elements = E;
  s->info.samples_per_element = Q;
  s->info.corners = K;
  s->info.factor_nnz = znnz;
  s->info.matrix_nnz = annz;
  s->info.jacobian_bytes = jac_bytes;
  s->info.factor_values = factor_values;
  s->info.factor_slots = slots;
  s->info.slice_rows = d.slice_rows;
  s->info.slice_switch = d.t_switch;
  s->info.split_quads = d.split_quads;
  s->info.split_switch = d.s_switch;
  s->info.tune_seconds = 0.0;
  if (s->cfg.slice_rows == 0 && d.t_nitems > 0) {
    tune_products(s, s_items, t_items, rows_capacity, cols_capacity, defect);
    const int groups = static_cast<int>(t_items.size());
    const int all_cut[2] = {groups < d.t_switch ? 1 : 0, groups < d.s_switch ? 1 : 0};
    HIP_CHECK(hipMemcpy(d.t_mode + 1, &all_cut[0], sizeof(int), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(d.t_mode + 3, &all_cut[1], sizeof(int), hipMemcpyHostToDevice));
  }
  s->info.setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::fill(std::begin(s->loop_stats), std::end(s->loop_stats), 0);
}

// Device arrays of the mesh contact (MeshContact.__init__, mesh_contact.py:180-243).
void build_mesh_contact(dphd_solver* s) {
  auto& o = s->owned;
  const int n = s->N;
  MeshContactDev& mc = s->mc;
  mc = MeshContactDev{};
  mc.n = n;
  mc.radius = s->mc_params[0];
  mc.kf = s->mc_params[3];
  mc.mu = s->mc_params[4];
  auto unique_vertices = [](const std::vector<int>& faces) {
    std::vector<int> v(faces);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    return v;
  };
  const std::vector<int> va = unique_vertices(s->mc_faces_a), vb = unique_vertices(s->mc_faces_b);
  std::vector<int> slot_a(n, -1), slot_b(n, -1);
  for (size_t k = 0; k < va.size(); ++k) slot_a[va[k]] = static_cast<int>(k);
  for (size_t k = 0; k < vb.size(); ++k) slot_b[vb[k]] = static_cast<int>(k);
  // Per particle, its (face, corner) entries in face order.
  auto incidence = [n](const std::vector<int>& faces, std::vector<int>& off, std::vector<int>& ent) {
    off.assign(n + 1, 0);
    for (const int v : faces) ++off[v + 1];
    for (int i = 0; i < n; ++i) off[i + 1] += off[i];
    ent.assign(2 * faces.size(), 0);
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (size_t t = 0; t < faces.size(); ++t) {
      const int k = fill[faces[t]]++;
      ent[2 * k] = static_cast<int>(t / 3);
      ent[2 * k + 1] = static_cast<int>(t % 3);
    }
  };
  std::vector<int> ia_off, ia, ib_off, ib;
  incidence(s->mc_faces_a, ia_off, ia);
  incidence(s->mc_faces_b, ib_off, ib);
  const int fa = static_cast<int>(s->mc_faces_a.size() / 3), fb = static_cast<int>(s->mc_faces_b.size() / 3);
  const int* dva = dupload(o, va);
  const int* dvb = dupload(o, vb);
  const int* dfa = dupload(o, s->mc_faces_a);
  const int* dfb = dupload(o, s->mc_faces_b);
  auto side = [&](const int* verts, int nv, const int* faces, int nf, double kn) {
    MeshContactSide sd;
    sd.nv = nv;
    sd.nf = nf;
    sd.verts = verts;
    sd.faces = faces;
    sd.kn = kn;
    // A face's widened box spans at most two cells per axis (the cell edge is
    // at least the largest box edge), so it is binned into at most 8 cells.
    sd.grid = dzeros<double>(o, 11);
    sd.fcell_off = dzeros<int>(o, kGridCells + 1);
    sd.fcell = dzeros<int>(o, 8 * static_cast<size_t>(std::max(nf, 1)));
    sd.vcell_off = dzeros<int>(o, kGridCells + 1);
    sd.vcell = dzeros<int>(o, static_cast<size_t>(std::max(nv, 1)));
    sd.cursor = dzeros<int>(o, 2 * static_cast<size_t>(kGridCells));
    sd.vcell_of = dzeros<int>(o, static_cast<size_t>(std::max(nv, 1)));
    sd.unsorted = dzeros<int>(o, 8 * static_cast<size_t>(nf) + nv + 1);
    return sd;
  };
  mc.ab = side(dva, static_cast<int>(va.size()), dfb, fb, s->mc_params[1]);
  mc.ba = side(dvb, static_cast<int>(vb.size()), dfa, fa, s->mc_params[2]);
  mc.slot_a = dupload(o, slot_a);
  mc.slot_b = dupload(o, slot_b);
  mc.inc_a_off = dupload(o, ia_off);
  mc.inc_a = dupload(o, ia);
  mc.inc_b_off = dupload(o, ib_off);
  mc.inc_b = dupload(o, ib);
  mc.force = dzeros<double>(o, 3 * static_cast<size_t>(n));
  mc.pairs = dzeros<int>(o, 2);
}

// Device state of the frictional boundary and its panel A^{-1} e_j of the
// registered vertices (FrictionalContact.__init__ / build, contact.py:702-800).
void build_friction(dphd_solver* s) {
  if (s->B != 1) throw std::invalid_argument("a frictional boundary needs a single-body solver");
  auto& o = s->owned;
  const int K = static_cast<int>(s->fc_nodes.size()), n = s->N;
  const size_t k1 = std::max(K, 1);
  FrictionDev& fc = s->fc;
  fc = FrictionDev{};
  fc.kind = s->fc_kind;
  fc.K = K;
  fc.n = n;
  fc.friction = s->fc_friction;
  fc.restitution = s->fc_restitution;
  fc.params = dupload(o, std::vector<double>(s->fc_params, s->fc_params + 4));
  fc.nodes = dupload(o, s->fc_nodes.empty() ? std::vector<int>{0} : s->fc_nodes);
  fc.panel = dzeros<double>(o, k1 * n);
  fc.blockT = dzeros<double>(o, k1 * k1);
  fc.delassus = dzeros<double>(o, k1 * k1);
  fc.matrix = dzeros<double>(o, k1 * k1);
  fc.reg = dzeros<double>(o, k1);
  fc.normal = dzeros<double>(o, 3 * k1);
  fc.t1 = dzeros<double>(o, 3 * k1);
  fc.t2 = dzeros<double>(o, 3 * k1);
  fc.fl = dzeros<double>(o, 2 * k1);
  fc.mult = dzeros<double>(o, k1);
  fc.gap = dzeros<double>(o, k1);
  fc.omega = dzeros<double>(o, k1);
  fc.diag = dzeros<double>(o, k1);
  fc.resid = dzeros<double>(o, k1);
  fc.rhs = dzeros<double>(o, k1);
  fc.force = dzeros<double>(o, 3 * k1);
  fc.response = dzeros<double>(o, 3 * k1);
  fc.order = dzeros<int>(o, k1);
  fc.slot = dzeros<int>(o, k1);
  fc.nactive = dzeros<int>(o, 1);
  fc.prev_order = dzeros<int>(o, k1);
  fc.nprev = dzeros<int>(o, 1);
  fc.perm = dzeros<int>(o, k1);
  fc.compact = dzeros<double>(o, k1);
  fc.work = dzeros<double>(o, k1);
  fc.candidate = dzeros<int>(o, k1);
  fc.freec = dzeros<double>(o, 3 * k1);
  fc.fprev = dzeros<double>(o, 3 * k1);
  fc.defect = s->d.defect;
  if (s->d.defect) {
    std::vector<int> vslot(n, -1);
    for (int c = 0; c < K; ++c) vslot[s->fc_nodes[c]] = c;
    s->d.fc_vslot = dupload(o, vslot);
    s->d.fc_fprev = fc.fprev;
  }
  if (fc.kind == 2) {
    fc.radius = s->fc_radius;
    fc.split = s->fc_split;
    fc.left_count = s->fc_left;
    fc.face_count = static_cast<int>(s->fc_faces.size() / 3);
    fc.faces = dupload(o, s->fc_faces.empty() ? std::vector<int>{0, 0, 0} : s->fc_faces);
    fc.surface = dzeros<double>(o, 3 * static_cast<size_t>(n));
    fc.left_words = (fc.left_count + 63) / 64;
    fc.words = fc.left_words + (fc.face_count - fc.left_count + 63) / 64;
    fc.face_box = dzeros<double>(o, 6 * static_cast<size_t>(std::max(fc.words, 1)));
  }
  // Woodbury rows of the backward and the recorded contact state they read.
  s->rec_mult = dzeros<double>(o, k1);
  s->rec_omega = dzeros<double>(o, k1);
  s->rec_normal = dzeros<double>(o, 3 * k1);
  s->wb_mu = dzeros<double>(o, 3 * static_cast<size_t>(n));
  s->wb_z = dzeros<double>(o, 3 * static_cast<size_t>(n));
  WoodburyDev& w = s->wb;
  w = WoodburyDev{};
  w.K = K;
  w.n = n;
  w.tangents = s->fc_friction > 0.0 ? 1 : 0;
  w.R = w.tangents ? 3 * K : K;
  const size_t r1 = std::max(w.R, 1);
  w.nodes = fc.nodes;
  w.mult = s->rec_mult;
  w.omega = s->rec_omega;
  w.normal = s->rec_normal;
  w.row_node = dzeros<int>(o, r1);
  w.row_vec = dzeros<double>(o, 3 * r1);
  w.row_valid = dzeros<int>(o, r1);
  w.value = dzeros<double>(o, r1);
  w.order = dzeros<int>(o, r1);
  w.count = dzeros<int>(o, 1);
  w.gram = dzeros<double>(o, r1 * r1);
  w.coef = dzeros<double>(o, r1);
  w.perm = dzeros<int>(o, r1);
  w.work = dzeros<double>(o, r1);
  s->wb_capacity = 0;
  if (K == 0) return;
  double* rhs = dzeros<double>(o, 3 * static_cast<size_t>(n));
  double* out = dzeros<double>(o, 3 * static_cast<size_t>(n));
  double* r = dzeros<double>(o, 3 * static_cast<size_t>(n));
  double* corr = dzeros<double>(o, 3 * static_cast<size_t>(n));
  hipStream_t st = nullptr;
  for (int j0 = 0; j0 < K; j0 += 3) {
    launch_contact_units(fc, j0, rhs, st);
    // With the FP32 factor the columns are refined against the FP64 A. The
    // step's fixed point does not depend on them (the forces enter the
    // defect residual), only its convergence rate does.
    launch_refined_inverse(s->d, rhs, out, r, corr, 3, st);
    launch_contact_store(fc, j0, out, st);
  }
  launch_contact_block(fc, st);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipDeviceSynchronize());
}

// Work that a caller queues on its own stream and that shares this solver's
// device buffers (join_user before it, release_user after it): it runs after
// everything queued on the solver's stream, and the solver's next work on
// its stream waits for it (flush_user, at the start of step and backward:
// the event is recorded then, so consecutive calls on one stream cost no
// event each). Work pending on another caller stream is ordered first.
void flush_user(dphd_solver* s) {
  if (!s->user_pending) return;
  hipError_t e = hipEventRecord(s->ev_user, s->user_pending);
  if (e == hipSuccess) e = hipStreamWaitEvent(s->stream, s->ev_user, 0);
  s->user_pending = nullptr;
  if (e != hipSuccess) {
    // The caller's stream is gone: wait for the whole device instead.
    (void)hipGetLastError();
    HIP_CHECK(hipDeviceSynchronize());
  }
}
void join_user(dphd_solver* s, hipStream_t user) {
  if (s->user_pending && s->user_pending != user) flush_user(s);
  const hipError_t idle = hipStreamQuery(s->stream);
  if (idle == hipSuccess) return;
  if (idle != hipErrorNotReady) HIP_CHECK(idle);
  HIP_CHECK(hipEventRecord(s->ev_out, s->stream));
  HIP_CHECK(hipStreamWaitEvent(user, s->ev_out, 0));
}
void release_user(dphd_solver* s, hipStream_t user) { s->user_pending = user; }

StepScalars scalars(const dphd_solver* s) {
  return StepScalars{s->dt, s->cfg.abs_tol, s->cfg.rel_tol, s->cfg.max_iterations};
}

// Loop modes: the forward step, the backbone adjoint, and the Woodbury
// columns (adjoint iterations capped at min(max_iterations, 120) that do not
// report non-convergence, adjoint.py:1054-1060, 1028-1052).
enum LoopMode { kForward = 0, kAdjoint = 1, kWoodbury = 2 };

int loop_limit(const dphd_solver* s, int mode) {
  return mode == kWoodbury ? std::min(s->cfg.max_iterations, 120) : s->cfg.max_iterations;
}

hipGraphExec_t capture(dphd_solver* s, int mode) {
  StepScalars sc = scalars(s);
  sc.max_iterations = loop_limit(s, mode);
  sc.report = mode == kWoodbury ? 0 : 1;
  const int C = chunk_size(s, mode);
  hipGraph_t graph = nullptr;
  HIP_CHECK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  for (int c = 0; c < C; ++c) {
    if (mode != kForward)
      launch_adjoint_iteration(s->d, sc, s->stream);
    else
      launch_forward_iteration(s->d, sc, s->targets, s->has_friction ? &s->fc : nullptr, s->stream);
  }
  launch_any_active(s->d, mode != kForward ? s->d.aactive : s->d.active, mode != kForward ? s->d.aiter : s->d.iter,
                    s->stream);
  HIP_CHECK(hipStreamEndCapture(s->stream, &graph));
  hipGraphExec_t exec = nullptr;
  HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  HIP_CHECK(hipGraphDestroy(graph));
  return exec;
}

// Run the predicated loop to every body's convergence: launch enough chunks
// to cover the previous solve's iterations back to back, then poll.
int run_loop(dphd_solver* s, int mode) {
  hipGraphExec_t& exec = s->exec[mode];
  if (!exec || s->key_iter[mode] != s->cfg.max_iterations || s->key_abs[mode] != s->cfg.abs_tol ||
      s->key_rel[mode] != s->cfg.rel_tol) {
    if (exec) HIP_CHECK(hipGraphExecDestroy(exec));
    exec = capture(s, mode);
    s->key_iter[mode] = s->cfg.max_iterations;
    s->key_abs[mode] = s->cfg.abs_tol;
    s->key_rel[mode] = s->cfg.rel_tol;
  }
  const int C = chunk_size(s, mode);
  const int max_chunks = (loop_limit(s, mode) + C - 1) / C;
  const int last = s->last_iters[mode];
  // First burst: the chunks the previous solve needed (its most iterations of
  // any body, rounded up to whole chunks), so a solve that needs as many
  // waits once; one that converges sooner runs predicated-off chunks (every
  // kernel exits at its first flag read) instead of a host round trip.
  int first = std::max(1, std::min(max_chunks, (last + C - 1) / C));
  int launched = 0;
  while (launched < max_chunks) {
    const int burst = launched == 0 ? first : 1;
    for (int k = 0; k < burst && launched < max_chunks; ++k, ++launched) HIP_CHECK(hipGraphLaunch(exec, s->stream));
    HIP_CHECK(hipStreamSynchronize(s->stream));
    ++s->loop_stats[2 * mode + 1];
    if (!*reinterpret_cast<volatile int*>(s->any_host)) break;
  }
  s->loop_stats[2 * mode] += launched;
  s->last_iters[mode] = reinterpret_cast<volatile int*>(s->any_host)[1];
  if (mode != kForward) launch_adjoint_finalize(s->d, s->stream);
  return launched;
}

// The adjoint loop of dphd_backward_static: the kernels of the captured
// adjoint graph launched directly (a graph instantiation may allocate), a
// chunk of iterations at a time, polling the mapped flag after each chunk.
void run_adjoint_direct(dphd_solver* s, const DeviceState& d) {
  StepScalars sc = scalars(s);
  sc.max_iterations = loop_limit(s, kAdjoint);
  sc.report = 1;
  const int C = chunk_size(s, kAdjoint);
  const int max_chunks = (sc.max_iterations + C - 1) / C;
  for (int launched = 0; launched < max_chunks; ++launched) {
    for (int c = 0; c < C; ++c) launch_adjoint_iteration(d, sc, s->stream);
    launch_any_active(d, d.aactive, d.aiter, s->stream);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s->stream));
    s->loop_stats[2 * kAdjoint] += 1;
    s->loop_stats[2 * kAdjoint + 1] += 1;
    if (!*reinterpret_cast<volatile int*>(s->any_host)) break;
  }
  launch_adjoint_finalize(d, s->stream);
}

// The solver of the Woodbury columns: up to 64 copies of this single body
// (the reference solves its columns 64 at a time, adjoint.py:662-663), built
// once (woodbury_ready) and kept until this solver is rebuilt. It shares
// this solver's stream; its factor is computed once.
dphd_solver* woodbury_batch(dphd_solver* s) {
  if (s->wb_batch) return s->wb_batch;
  const int C = std::max(1, std::min(64, s->wb.R));
  const int N = s->N, E = s->E, K = s->K;
  auto* b = new dphd_solver;
  b->cfg = s->cfg;
  b->device = s->device;
  b->K = s->K;
  b->Q = s->Q;
  b->B = C;
  b->N = C * N;
  b->E = C * E;
  for (int c = 0; c <= C; ++c) {
    b->voff.push_back(c * N);
    b->eoff.push_back(c * E);
  }
  for (int c = 0; c < C; ++c) {
    for (const int v : s->indices) b->indices.push_back(v + c * N);
    b->poses.insert(b->poses.end(), s->poses.begin(), s->poses.end());
    b->inv_mass.insert(b->inv_mass.end(), s->inv_mass.begin(), s->inv_mass.end());
    b->gravity.insert(b->gravity.end(), s->gravity.begin(), s->gravity.begin() + 3);
    b->density.push_back(s->density[0]);
    b->materials.insert(b->materials.end(), s->materials.begin(), s->materials.end());
    b->mdir_in.insert(b->mdir_in.end(), s->mdir_in.begin(), s->mdir_in.end());
    b->mstiff_in.insert(b->mstiff_in.end(), s->mstiff_in.begin(), s->mstiff_in.end());
  }
  b->have_topology = b->have_particles = b->have_materials = true;
  b->identical_bodies = true;
  b->stream = s->stream;
  try {
    build_solver(b, s->dt);
  } catch (...) {
    b->stream = nullptr;
    delete b;
    throw;
  }
  s->wb_ckpt = dalloc<double>(s->owned, static_cast<size_t>(C) * 3 * N);
  s->wb_targets = s->d.has_muscles ? dalloc<double>(s->owned, static_cast<size_t>(C) * E) : nullptr;
  s->wb_batch = b;
  return b;
}

// The batch solver with its adjoint operator buffers allocated. Built with
// the factor of a solver with a frictional boundary (dphd_build), so that no
// backward pays for it; otherwise at the first backward that selects rows.
dphd_solver* woodbury_ready(dphd_solver* s) {
  dphd_solver* b = woodbury_batch(s);
  DeviceState& bd = b->d;
  if (!bd.jac) {
    bd.jac = dalloc<double>(b->owned, static_cast<size_t>(b->E) * b->Q * 81);
    bd.ke = dalloc<double>(b->owned, static_cast<size_t>(b->E) * b->K * b->K * 9);
    bd.kval = dalloc<double>(b->owned, static_cast<size_t>(bd.kblocks) * 9);
    for (int m = kAdjoint; m <= kWoodbury; ++m) {
      if (b->exec[m]) HIP_CHECK(hipGraphExecDestroy(b->exec[m]));
      b->exec[m] = nullptr;
    }
  }
  return b;
}

// Woodbury projection of the adjoint for the parameter gradients; returns z
// (Adjoint._woodbury, adjoint.py:1272-1354). The selected rows' columns
// W_T = (A - K)^{-1} J^T are solved in batches, one column per body of the
// batch solver, each stopping on its own gate after at most
// min(max_iterations, 120) iterations (adjoint.py:662-806, 966-1052).
const double* woodbury(dphd_solver* s, const double* checkpoint, const double* targets) {
  DeviceState& d = s->d;
  WoodburyDev& w = s->wb;
  hipStream_t st = s->stream;
  const size_t vb = sizeof(double) * 3 * static_cast<size_t>(s->N);
  HIP_CHECK(hipMemcpyAsync(s->wb_mu, d.sol, vb, hipMemcpyDeviceToDevice, st));
  launch_woodbury_rows(w, s->wb_mu, d.route, st);
  int count = 0;
  HIP_CHECK(hipMemcpyAsync(&count, w.count, sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  if (count > s->wb_capacity) {
    w.columns = dalloc<double>(s->owned, static_cast<size_t>(count) * 3 * s->N);
    s->wb_capacity = count;
  }
  if (count > 0) {
    dphd_solver* b = woodbury_ready(s);
    DeviceState& bd = b->d;
    const int C = b->B;
    // The tolerances may have changed between steps (dphd_set_tolerances).
    b->cfg.max_iterations = s->cfg.max_iterations;
    b->cfg.abs_tol = s->cfg.abs_tol;
    b->cfg.rel_tol = s->cfg.rel_tol;
    // The adjoint operator at the recorded step, once for every copy.
    for (int c = 0; c < C; ++c) {
      HIP_CHECK(hipMemcpyAsync(s->wb_ckpt + static_cast<size_t>(c) * 3 * s->N, checkpoint, vb,
                               hipMemcpyDeviceToDevice, st));
      if (s->wb_targets && targets)
        HIP_CHECK(hipMemcpyAsync(s->wb_targets + static_cast<size_t>(c) * s->E, targets, sizeof(double) * s->E,
                                 hipMemcpyDeviceToDevice, st));
    }
    StepScalars sc = scalars(b);
    launch_adjoint_setup(bd, sc, s->wb_ckpt, s->wb_targets && targets ? s->wb_targets : nullptr, nullptr, nullptr, st);
    for (int first = 0; first < count; first += C) {
      launch_woodbury_batch_seed(w, first, C, s->N, bd.arhs, bd.sol, bd.xr, bd.fscale, bd.aiter, bd.aactive, bd.aa,
                                 st);
      launch_active_prefix(bd, true, st);
      const int64_t chunks0 = b->loop_stats[2 * kWoodbury], waits0 = b->loop_stats[2 * kWoodbury + 1];
      run_loop(b, kWoodbury);
      s->loop_stats[2 * kWoodbury] += b->loop_stats[2 * kWoodbury] - chunks0;
      s->loop_stats[2 * kWoodbury + 1] += b->loop_stats[2 * kWoodbury + 1] - waits0;
      const int here = std::min(C, count - first);
      HIP_CHECK(hipMemcpyAsync(w.columns + static_cast<size_t>(first) * 3 * s->N, bd.sol,
                               vb * static_cast<size_t>(here), hipMemcpyDeviceToDevice, st));
    }
    // The columns' safe clips report in this solver's status (as the
    // reference's batched columns share the adjoint's status word).
    launch_status_merge(bd.status, d.status, st);
  }
  launch_woodbury_project(w, s->wb_mu, d.free, s->wb_z, d.status, st);
  return s->wb_z;
}

}  // namespace
}  // namespace dphd

using dphd::set_error;

#define DPHD_TRY(...)                                     \
  try {                                                   \
    __VA_ARGS__;                                          \
    return DPHD_OK;                                       \
  } catch (const std::invalid_argument& e) {              \
    set_error(e.what());                                  \
    return DPHD_EINVAL;                                   \
  } catch (const std::domain_error& e) {                  \
    set_error(e.what());                                  \
    return DPHD_ENOTIMPL;                                 \
  } catch (const std::bad_alloc&) {                       \
    set_error("out of memory");                           \
    return DPHD_ENOMEM;                                   \
  } catch (const std::exception& e) {                     \
    set_error(e.what());                                  \
    return DPHD_EHIP;                                     \
  }

extern "C" {

int dphd_create(const dphd_config* config, int device, dphd_solver** out) {
  if (!config || !out) {
    set_error("dphd_create: null argument");
    return DPHD_EINVAL;
  }
  const dphd_config& c = *config;
  if (c.element_kind != DPHD_TET4 && c.element_kind != DPHD_HEX8) {
    set_error("element_kind must be DPHD_TET4 or DPHD_HEX8");
    return DPHD_EINVAL;
  }
  if (c.material != DPHD_PD && c.material != DPHD_NEO_HOOKEAN) {
    set_error("material must be 'pd' or 'neo_hookean'");
    return DPHD_EINVAL;
  }
  if (c.anderson_window < 0 || c.anderson_window > dphd::kMaxWindow) {
    set_error("anderson_window must be an integer in [0, 8]");
    return DPHD_EINVAL;
  }
  if (c.max_iterations < 1) {
    set_error("max_iterations must be a positive integer");
    return DPHD_EINVAL;
  }
  if (!std::isfinite(c.abs_tol) || c.abs_tol < 0 || !std::isfinite(c.rel_tol) || c.rel_tol < 0) {
    set_error("abs_tol and rel_tol must be finite and nonnegative");
    return DPHD_EINVAL;
  }
  if (c.slice_rows < 0) {
    set_error("slice_rows must be 0 (tuned) or positive");
    return DPHD_EINVAL;
  }
  if (c.factor_precision != 64 && c.factor_precision != 32) {
    set_error("factor_precision must be 64 or 32");
    return DPHD_EINVAL;
  }
  DPHD_TRY({
    HIP_CHECK(hipSetDevice(device));
    dphd::scratch_setup(device);
    auto* s = new dphd_solver;
    s->cfg = c;
    s->device = device;
    s->K = c.element_kind == DPHD_TET4 ? 4 : 8;
    s->Q = c.element_kind == DPHD_TET4 ? 1 : 8;
    *out = s;
  })
}

void dphd_solver::drop_woodbury_batch() {
  if (!wb_batch) return;
  dphd_solver* b = wb_batch;
  wb_batch = nullptr;
  wb_ckpt = wb_targets = nullptr;  // owned by this solver, freed with its buffers
  (void)hipSetDevice(device);
  if (stream) (void)hipStreamSynchronize(stream);
  b->stream = nullptr;  // this solver's stream
  b->release();
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_out) (void)hipEventDestroy(b->ev_out);
  if (b->ev_user) (void)hipEventDestroy(b->ev_user);
  delete b;
}

void dphd_destroy(dphd_solver* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  // Caller-stream work (apply_inverse, mesh_contact_force) may still read
  // the buffers freed here.
  (void)hipDeviceSynchronize();
  s->release();
  if (s->ev_in) (void)hipEventDestroy(s->ev_in);
  if (s->ev_out) (void)hipEventDestroy(s->ev_out);
  if (s->ev_user) (void)hipEventDestroy(s->ev_user);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

int dphd_set_bodies(dphd_solver* s, int bodies, const int* vertex_offsets, const int* element_offsets) {
  DPHD_TRY({
    if (!s || bodies < 1 || !vertex_offsets || !element_offsets) throw std::invalid_argument("dphd_set_bodies: invalid arguments");
    if (vertex_offsets[0] != 0 || element_offsets[0] != 0) throw std::invalid_argument("offsets must start at 0");
    for (int b = 0; b < bodies; ++b)
      if (vertex_offsets[b + 1] <= vertex_offsets[b] || element_offsets[b + 1] <= element_offsets[b])
        throw std::invalid_argument("every body needs vertices and elements");
    s->B = bodies;
    s->voff.assign(vertex_offsets, vertex_offsets + bodies + 1);
    s->eoff.assign(element_offsets, element_offsets + bodies + 1);
    s->N = s->voff[bodies];
    s->E = s->eoff[bodies];
    s->have_topology = false;
    s->release();
  })
}

int dphd_set_tet4(dphd_solver* s, const int* indices, const float* poses) {
  DPHD_TRY({
    if (!s || !indices || !poses || s->B == 0) throw std::invalid_argument("dphd_set_tet4: set bodies first");
    if (s->cfg.element_kind != DPHD_TET4) throw std::invalid_argument("solver was created for Hex8");
    s->indices.assign(indices, indices + static_cast<size_t>(s->E) * 4);
    s->poses.assign(poses, poses + static_cast<size_t>(s->E) * 9);
    std::vector<int> used(s->N, 0);
    for (int v : s->indices) {
      if (v < 0 || v >= s->N) throw std::invalid_argument("tet indices out of range");
      used[v] = 1;
    }
    for (int i = 0; i < s->N; ++i)
      if (!used[i]) throw std::invalid_argument("Every particle must belong to a tetrahedron");
    s->have_topology = true;
    s->release();
  })
}

int dphd_set_hex8(dphd_solver* s, const int* indices, const double* rest_vertices) {
  DPHD_TRY({
    if (!s || !indices || !rest_vertices || s->B == 0) throw std::invalid_argument("dphd_set_hex8: set bodies first");
    if (s->cfg.element_kind != DPHD_HEX8) throw std::invalid_argument("solver was created for Tet4");
    s->indices.assign(indices, indices + static_cast<size_t>(s->E) * 8);
    s->rest.assign(rest_vertices, rest_vertices + static_cast<size_t>(s->N) * 3);
    std::vector<int> used(s->N, 0);
    for (int e = 0; e < s->E; ++e) {
      for (int k = 0; k < 8; ++k) {
        const int v = s->indices[8 * static_cast<size_t>(e) + k];
        if (v < 0 || v >= s->N) throw std::invalid_argument("hex indices out of range");
        for (int j = 0; j < k; ++j)
          if (s->indices[8 * static_cast<size_t>(e) + j] == v)
            throw std::invalid_argument("hex_mesh indices must have shape (element_count,8) with eight distinct valid vertices");
        used[v] = 1;
      }
    }
    for (int i = 0; i < s->N; ++i)
      if (!used[i]) throw std::invalid_argument("Every particle must belong to a hexahedron");
    s->have_topology = true;
    s->release();
  })
}

int dphd_set_particles(dphd_solver* s, const float* inverse_mass, const double* gravity, const double* density) {
  DPHD_TRY({
    if (!s || !inverse_mass || !gravity || !density || s->B == 0) throw std::invalid_argument("dphd_set_particles: invalid arguments");
    s->inv_mass.assign(inverse_mass, inverse_mass + s->N);
    s->gravity.assign(gravity, gravity + 3 * static_cast<size_t>(s->B));
    s->density.assign(density, density + s->B);
    for (double r : s->density)
      if (!std::isfinite(r) || r <= 0.0) throw std::invalid_argument("density must be finite and positive");
    s->have_particles = true;
    s->release();
  })
}

int dphd_set_materials(dphd_solver* s, const double* materials) {
  DPHD_TRY({
    if (!s || !materials || s->E == 0) throw std::invalid_argument("dphd_set_materials: invalid arguments");
    s->materials.assign(materials, materials + 2 * static_cast<size_t>(s->E));
    s->have_materials = true;
    s->release();
  })
}

int dphd_set_muscles(dphd_solver* s, const double* directions, const double* stiffness) {
  DPHD_TRY({
    if (!s || s->E == 0) throw std::invalid_argument("dphd_set_muscles: invalid arguments");
    if (!directions || !stiffness) {
      s->mdir_in.clear();
      s->mstiff_in.clear();
    } else {
      s->mdir_in.assign(directions, directions + 3 * static_cast<size_t>(s->E));
      s->mstiff_in.assign(stiffness, stiffness + s->E);
      for (int e = 0; e < s->E; ++e)
        if (!std::isfinite(s->mstiff_in[e]) || s->mstiff_in[e] < 0.0)
          throw std::invalid_argument("Muscle directions and stiffness must be finite, with nonnegative stiffness");
    }
    s->release();
  })
}

int dphd_set_planar(dphd_solver* s, int count, const int* vertices, const double* stiffness, const double* plane) {
  DPHD_TRY({
    if (!s || s->N == 0 || count < 0 || (count > 0 && (!vertices || !stiffness || !plane)))
      throw std::invalid_argument("dphd_set_planar: invalid arguments");
    std::vector<int> idx(vertices, vertices + count);
    std::vector<char> seen(s->N, 0);
    for (const int i : idx) {
      if (i < 0 || i >= s->N || seen[i]) throw std::invalid_argument("planar_energies needs distinct valid particle indices");
      seen[i] = 1;
    }
    std::vector<double> w(stiffness, stiffness + count);
    for (const double k : w)
      if (!std::isfinite(k) || k < 0.0) throw std::invalid_argument("planar_energies stiffness must be finite and nonnegative");
    double unit[4] = {0, 0, 1, 0};
    if (count > 0) {
      const double len = std::sqrt(plane[0] * plane[0] + plane[1] * plane[1] + plane[2] * plane[2]);
      bool finite = std::isfinite(len);
      for (int t = 0; t < 4; ++t) finite = finite && std::isfinite(plane[t]);
      if (!finite || len <= 1.0e-5)
        throw std::invalid_argument("planar_energies needs a finite (nx,ny,nz,offset) with nonzero normal");
      for (int t = 0; t < 4; ++t) unit[t] = plane[t] / len;
    }
    s->planar_idx = std::move(idx);
    s->planar_w = std::move(w);
    for (int t = 0; t < 4; ++t) s->plane[t] = unit[t];
    s->release();
  })
}

int dphd_set_mesh_contact(dphd_solver* s, double contact_radius, double kn_left, double kn_right, double kf,
                          double mu, int faces_a_count, const int* faces_a, int faces_b_count, const int* faces_b) {
  DPHD_TRY({
    if (!s || s->N == 0) throw std::invalid_argument("dphd_set_mesh_contact: invalid arguments");
    if (faces_a_count == 0 && faces_b_count == 0 && !faces_a && !faces_b) {
      s->has_mesh_contact = false;
      s->release();
      return DPHD_OK;
    }
    const double params[5] = {contact_radius, kn_left, kn_right, kf, mu};
    const char* names[5] = {"contact_radius", "kn_left", "kn_right", "kf", "mu"};
    for (int t = 0; t < 5; ++t)
      if (!std::isfinite(params[t]) || params[t] < 0.0)
        throw std::invalid_argument(std::string("mesh_contact ") + names[t] + " must be finite and nonnegative");
    if (faces_a_count <= 0 || faces_b_count <= 0 || !faces_a || !faces_b)
      throw std::invalid_argument("mesh_contact needs faces on both bodies");
    std::vector<int> fa(faces_a, faces_a + 3 * static_cast<size_t>(faces_a_count));
    std::vector<int> fb(faces_b, faces_b + 3 * static_cast<size_t>(faces_b_count));
    for (const auto* faces : {&fa, &fb})
      for (const int v : *faces)
        if (v < 0 || v >= s->N) throw std::invalid_argument("mesh_contact faces must reference existing particles");
    s->has_mesh_contact = true;
    for (int t = 0; t < 5; ++t) s->mc_params[t] = params[t];
    s->mc_faces_a = std::move(fa);
    s->mc_faces_b = std::move(fb);
    s->release();
  })
}

int dphd_set_frictional_boundary(dphd_solver* s, int kind, const double* parameters, int count, const int* vertices,
                                 double friction, double restitution, int split, int left_count, int face_count,
                                 const int* faces) {
  DPHD_TRY({
    if (!s || s->N == 0) throw std::invalid_argument("dphd_set_frictional_boundary: invalid arguments");
    if (kind < 0) {
      s->has_friction = false;
      s->release();
      return DPHD_OK;
    }
    if (kind > 2) throw std::invalid_argument("frictional boundary must be 'planar', 'spherical' or 'mesh'");
    if (!parameters) throw std::invalid_argument("frictional boundary parameters must be four finite values");
    double q[4];
    for (int t = 0; t < 4; ++t) {
      q[t] = parameters[t];
      if (!std::isfinite(q[t])) throw std::invalid_argument("frictional boundary parameters must be four finite values");
    }
    std::vector<int> f;
    if (kind == 2) {
      if (q[0] < 0.0) throw std::invalid_argument("contact_radius must be finite and nonnegative");
      if (face_count < 0 || left_count < 0 || left_count > face_count || (face_count > 0 && !faces))
        throw std::invalid_argument("mesh boundary parameters are (contact_radius, split_index, faces_left, faces_right)");
      f.assign(faces, faces + 3 * static_cast<size_t>(face_count));
      for (const int v : f)
        if (v < 0 || v >= s->N) throw std::invalid_argument("mesh boundary faces must reference existing particles");
    } else if (kind == 0 && std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) <= 2.220446049250313e-16) {
      throw std::invalid_argument("planar boundary needs a nonzero normal");
    }
    if (count < 0 || (count > 0 && !vertices)) throw std::invalid_argument("frictional vertices must be unique particle indices");
    std::vector<int> nodes(vertices, vertices + count);
    std::vector<char> seen(s->N, 0);
    for (const int v : nodes) {
      if (v < 0 || v >= s->N) throw std::invalid_argument("frictional vertices must reference existing particles");
      if (seen[v]) throw std::invalid_argument("frictional vertices must be unique particle indices");
      seen[v] = 1;
    }
    if (!std::isfinite(friction) || friction < 0.0) throw std::invalid_argument("friction must be finite and nonnegative");
    if (!std::isfinite(restitution) || restitution < 0.0)
      throw std::invalid_argument("restitution must be finite and nonnegative");
    s->has_friction = true;
    s->fc_kind = kind;
    for (int t = 0; t < 4; ++t) s->fc_params[t] = kind == 2 ? (t == 0 ? q[0] : 0.0) : q[t];
    s->fc_radius = kind == 2 ? q[0] : 0.0;
    s->fc_split = split;
    s->fc_left = left_count;
    s->fc_faces = std::move(f);
    s->fc_nodes = std::move(nodes);
    s->fc_friction = friction;
    s->fc_restitution = restitution;
    s->release();
  })
}

int dphd_boundary_parameters(dphd_solver* s, const double* set, double* get) {
  DPHD_TRY({
    if (!s || !s->built || !s->has_friction) throw std::invalid_argument("dphd_boundary_parameters: no frictional boundary");
    HIP_CHECK(hipSetDevice(s->device));
    if (set) {
      for (int t = 0; t < 4; ++t)
        if (!std::isfinite(set[t])) throw std::invalid_argument("frictional boundary parameters must be four finite values");
      // On the solver's stream, so the next step reads them after the copy.
      HIP_CHECK(hipMemcpyAsync(s->fc.params, set, sizeof(double) * 4, hipMemcpyHostToDevice, s->stream));
    }
    if (get) HIP_CHECK(hipMemcpyAsync(get, s->fc.params, sizeof(double) * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_CHECK(hipStreamSynchronize(s->stream));
  })
}

int dphd_contact_state(dphd_solver* s, const double* set_multiplier, double* multiplier, double* omega, double* normal,
                       int* candidate) {
  DPHD_TRY({
    if (!s || !s->built || !s->has_friction) throw std::invalid_argument("dphd_contact_state: no frictional boundary");
    HIP_CHECK(hipSetDevice(s->device));
    HIP_CHECK(hipDeviceSynchronize());
    const size_t K = s->fc.K;
    if (set_multiplier && K) HIP_CHECK(hipMemcpy(s->fc.mult, set_multiplier, sizeof(double) * K, hipMemcpyDefault));
    if (multiplier && K) HIP_CHECK(hipMemcpy(multiplier, s->fc.mult, sizeof(double) * K, hipMemcpyDefault));
    if (omega && K) HIP_CHECK(hipMemcpy(omega, s->fc.omega, sizeof(double) * K, hipMemcpyDefault));
    if (normal && K) HIP_CHECK(hipMemcpy(normal, s->fc.normal, sizeof(double) * 3 * K, hipMemcpyDefault));
    if (candidate && K) HIP_CHECK(hipMemcpy(candidate, s->fc.candidate, sizeof(int) * K, hipMemcpyDefault));
    // Device-to-device copies may still be running: finish them before the
    // solver's stream or the caller uses these arrays.
    HIP_CHECK(hipDeviceSynchronize());
  })
}

int dphd_contact_record(dphd_solver* s, const double* multiplier, const double* omega, const double* normal) {
  DPHD_TRY({
    if (!s || !s->built || !s->has_friction || !multiplier || !omega || !normal)
      throw std::invalid_argument("dphd_contact_record: needs a built solver with a frictional boundary");
    HIP_CHECK(hipSetDevice(s->device));
    const size_t K = s->fc.K;
    if (K) {
      // After the solver's queued work (a backward may still read the
      // recorded state), and complete before its next.
      HIP_CHECK(hipStreamSynchronize(s->stream));
      HIP_CHECK(hipMemcpy(s->rec_mult, multiplier, sizeof(double) * K, hipMemcpyDefault));
      HIP_CHECK(hipMemcpy(s->rec_omega, omega, sizeof(double) * K, hipMemcpyDefault));
      HIP_CHECK(hipMemcpy(s->rec_normal, normal, sizeof(double) * 3 * K, hipMemcpyDefault));
      HIP_CHECK(hipDeviceSynchronize());
    }
  })
}

int dphd_mesh_contact_force(dphd_solver* s, const double* position, const double* velocity, double* force,
                            void* stream) {
  DPHD_TRY({
    if (!s || !s->built || !s->has_mesh_contact || !position || !velocity || !force)
      throw std::invalid_argument("dphd_mesh_contact_force: needs a built solver with mesh contact");
    HIP_CHECK(hipSetDevice(s->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    dphd::join_user(s, st);
    dphd::launch_mesh_contact(s->mc, nullptr, nullptr, position, velocity, nullptr, nullptr, nullptr, 0.0, st);
    HIP_CHECK(hipMemcpyAsync(force, s->mc.force, sizeof(double) * 3 * static_cast<size_t>(s->N),
                             hipMemcpyDeviceToDevice, st));
    HIP_CHECK(hipGetLastError());
    dphd::release_user(s, st);
  })
}

int dphd_mesh_contact_pairs(dphd_solver* s, int* pairs) {
  DPHD_TRY({
    if (!s || !pairs) throw std::invalid_argument("dphd_mesh_contact_pairs: invalid arguments");
    *pairs = -1;
    if (!s->built || !s->has_mesh_contact) return DPHD_OK;
    HIP_CHECK(hipSetDevice(s->device));
    HIP_CHECK(hipDeviceSynchronize());
    int counts[2] = {0, 0};
    HIP_CHECK(hipMemcpy(counts, s->mc.pairs, sizeof(counts), hipMemcpyDeviceToHost));
    *pairs = counts[0] + counts[1];
  })
}

int dphd_mesh_contact_pair_counts(dphd_solver* s, int* counts) {
  DPHD_TRY({
    if (!s || !counts) throw std::invalid_argument("dphd_mesh_contact_pair_counts: invalid arguments");
    counts[0] = counts[1] = -1;
    if (!s->built || !s->has_mesh_contact) return DPHD_OK;
    HIP_CHECK(hipSetDevice(s->device));
    HIP_CHECK(hipDeviceSynchronize());
    HIP_CHECK(hipMemcpy(counts, s->mc.pairs, 2 * sizeof(int), hipMemcpyDeviceToHost));
  })
}

int dphd_set_tolerances(dphd_solver* s, int max_iterations, double abs_tol, double rel_tol) {
  DPHD_TRY({
    if (!s || max_iterations < 1 || !std::isfinite(abs_tol) || abs_tol < 0 || !std::isfinite(rel_tol) || rel_tol < 0)
      throw std::invalid_argument("invalid tolerances");
    s->cfg.max_iterations = max_iterations;
    s->cfg.abs_tol = abs_tol;
    s->cfg.rel_tol = rel_tol;
  })
}

int dphd_build(dphd_solver* s, double dt) {
  DPHD_TRY({
    if (!s) throw std::invalid_argument("null solver");
    // Caller-stream work may still read the buffers a rebuild frees.
    HIP_CHECK(hipSetDevice(s->device));
    HIP_CHECK(hipDeviceSynchronize());
    s->user_pending = nullptr;
    dphd::build_solver(s, dt);
    if (s->has_friction && s->wb.R > 0) {
      const auto t0 = std::chrono::steady_clock::now();
      dphd::woodbury_ready(s);
      s->info.setup_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    // The build's last launches (the tuner's, the batch solver's) finish
    // before any caller's work can reach the buffers they use.
    HIP_CHECK(hipDeviceSynchronize());
  })
}

int dphd_get_info(const dphd_solver* s, dphd_info* info) {
  if (!s || !info) {
    set_error("dphd_get_info: null argument");
    return DPHD_EINVAL;
  }
  *info = s->info;
  return DPHD_OK;
}

int dphd_loop_stats(const dphd_solver* s, int64_t* stats) {
  if (!s || !stats) {
    set_error("dphd_loop_stats: null argument");
    return DPHD_EINVAL;
  }
  std::copy(std::begin(s->loop_stats), std::end(s->loop_stats), stats);
  return DPHD_OK;
}

int dphd_body_nnz(const dphd_solver* s, int64_t* per_body) {
  if (!s || !per_body || !s->built) {
    set_error("dphd_body_nnz: build the solver first");
    return DPHD_ESTATE;
  }
  std::copy(s->body_nnz.begin(), s->body_nnz.end(), per_body);
  return DPHD_OK;
}

int dphd_step(dphd_solver* s, const float* q_in, const float* qd_in, const float* f_in, const float* dirichlet,
              const double* muscle_targets, float* q_out, float* qd_out, double* checkpoint, void* stream) {
  if (!s || !s->built) {
    set_error("dphd_step: call dphd_build first");
    return DPHD_ESTATE;
  }
  if (!q_in || !qd_in || !f_in || !q_out || !qd_out) {
    set_error("dphd_step: null state pointer");
    return DPHD_EINVAL;
  }
  if (q_in == q_out) {
    set_error("DiffPhD requires distinct input and output states");
    return DPHD_EINVAL;
  }
  DPHD_TRY({
    HIP_CHECK(hipSetDevice(s->device));
    hipStream_t user = static_cast<hipStream_t>(stream);
    dphd::flush_user(s);
    HIP_CHECK(hipEventRecord(s->ev_in, user));
    HIP_CHECK(hipStreamWaitEvent(s->stream, s->ev_in, 0));
    if (s->d.has_muscles) {
      if (muscle_targets)
        HIP_CHECK(hipMemcpyAsync(s->targets, muscle_targets, sizeof(double) * s->E, hipMemcpyDeviceToDevice, s->stream));
    }
    const dphd::StepScalars sc = dphd::scalars(s);
    dphd::launch_begin_step(s->d, sc, q_in, qd_in, f_in, dirichlet, s->stream);
    // q~ = q + h v + h^2/m (f_ext + ForwardStateForce(q, v)) (solver.py:762-770).
    if (s->has_mesh_contact)
      dphd::launch_mesh_contact(s->mc, q_in, qd_in, nullptr, nullptr, s->d.qt, s->d.vbody, s->d.mass, s->dt,
                                s->stream);
    if (s->has_friction) dphd::launch_contact_begin(s->fc, s->d.q0, s->dt, s->stream);
    dphd::run_loop(s, dphd::kForward);
    if (s->has_friction) {
      dphd::launch_contact_finish(s->fc, s->d.q0, qd_in, s->dt, s->d.x, s->stream);
      // The step's contact state, read by its backward unless replaced (StepRecord.contact).
      const size_t K = s->fc.K;
      if (K) {
        HIP_CHECK(hipMemcpyAsync(s->rec_mult, s->fc.mult, sizeof(double) * K, hipMemcpyDeviceToDevice, s->stream));
        HIP_CHECK(hipMemcpyAsync(s->rec_omega, s->fc.omega, sizeof(double) * K, hipMemcpyDeviceToDevice, s->stream));
        HIP_CHECK(
            hipMemcpyAsync(s->rec_normal, s->fc.normal, sizeof(double) * 3 * K, hipMemcpyDeviceToDevice, s->stream));
      }
    }
    dphd::launch_write_state(s->d, sc, q_out, qd_out, checkpoint, s->stream);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(s->ev_out, s->stream));
    HIP_CHECK(hipStreamWaitEvent(user, s->ev_out, 0));
  })
}

int dphd_backward(dphd_solver* s, const double* checkpoint, const double* muscle_targets, const float* gq_out,
                  const float* gqd_out, float* gq_in, float* gqd_in, float* gf_in, double* gmaterials,
                  float* gdirichlet, double* gactivations, void* stream) {
  if (!s || !s->built) {
    set_error("dphd_backward: call dphd_build first");
    return DPHD_ESTATE;
  }
  if (!checkpoint || !gq_in || !gqd_in || !gf_in) {
    set_error("dphd_backward: null pointer");
    return DPHD_EINVAL;
  }
  if (!s->planar_idx.empty()) {
    set_error("dphd_backward: planar vertex energies have no adjoint (the reference records none for Hex8)");
    return DPHD_ENOTIMPL;
  }
  if (gactivations && (s->Q != 1 || !s->d.has_muscles)) {
    set_error("dphd_backward: activation gradients need Tet4 elements with muscle fibres");
    return DPHD_EINVAL;
  }
  DPHD_TRY({
    HIP_CHECK(hipSetDevice(s->device));
    if (!s->d.jac) {
      s->d.jac = dphd::dalloc<double>(s->owned, static_cast<size_t>(s->E) * s->Q * 81);
      s->d.ke = dphd::dalloc<double>(s->owned, static_cast<size_t>(s->E) * s->K * s->K * 9);
      s->d.kval = dphd::dalloc<double>(s->owned, static_cast<size_t>(s->d.kblocks) * 9);
      for (int m = dphd::kAdjoint; m <= dphd::kWoodbury; ++m) {
        if (s->exec[m]) HIP_CHECK(hipGraphExecDestroy(s->exec[m]));
        s->exec[m] = nullptr;
      }
    }
    hipStream_t user = static_cast<hipStream_t>(stream);
    dphd::flush_user(s);
    HIP_CHECK(hipEventRecord(s->ev_in, user));
    HIP_CHECK(hipStreamWaitEvent(s->stream, s->ev_in, 0));
    if (s->d.has_muscles && muscle_targets)
      HIP_CHECK(hipMemcpyAsync(s->targets, muscle_targets, sizeof(double) * s->E, hipMemcpyDeviceToDevice, s->stream));
    const dphd::StepScalars sc = dphd::scalars(s);
    dphd::launch_adjoint_setup(s->d, sc, checkpoint, s->targets, gq_out, gqd_out, s->stream);
    dphd::run_loop(s, dphd::kAdjoint);
    dphd::launch_adjoint_route(s->d, s->stream);
    // Only parameter gradients use z; state gradients use mu (adjoint.py:1145-1148).
    const double* z = nullptr;
    if ((gmaterials || gactivations) && s->has_friction && s->fc.K > 0)
      z = dphd::woodbury(s, checkpoint, s->d.has_muscles ? s->targets : nullptr);
    dphd::launch_adjoint_outputs(s->d, sc, checkpoint, gqd_out, gq_in, gqd_in, gf_in, gmaterials, gactivations,
                                 gdirichlet, z, s->any_pinned, s->stream);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(s->ev_out, s->stream));
    HIP_CHECK(hipStreamWaitEvent(user, s->ev_out, 0));
  })
}

int dphd_backward_static(dphd_solver* s, const double* checkpoint, const double* muscle_targets, const float* gq_out,
                         const float* gqd_out, float* gq_in, float* gqd_in, float* gf_in, double* gmaterials,
                         float* gdirichlet, double* gactivations, void* stream) {
  if (!s || !s->built) {
    set_error("dphd_backward_static: call dphd_build first");
    return DPHD_ESTATE;
  }
  if (!checkpoint || !gq_in || !gqd_in || !gf_in) {
    set_error("dphd_backward_static: null pointer");
    return DPHD_EINVAL;
  }
  if (!s->planar_idx.empty()) {
    set_error("dphd_backward_static: planar vertex energies have no adjoint (the reference records none for Hex8)");
    return DPHD_ENOTIMPL;
  }
  if (gactivations && (s->Q != 1 || !s->d.has_muscles)) {
    set_error("dphd_backward_static: activation gradients need Tet4 elements with muscle fibres");
    return DPHD_EINVAL;
  }
  if ((gmaterials || gactivations) && s->has_friction && s->fc.K > 0) {
    set_error("dphd_backward_static: the Woodbury projection of parameter gradients with a frictional boundary "
              "allocates its columns; use dphd_backward");
    return DPHD_ENOTIMPL;
  }
  DPHD_TRY({
    HIP_CHECK(hipSetDevice(s->device));
    // Jacobian, element blocks and assembled operator, each from a 256-byte boundary.
    auto round = [](size_t count) { return (count + 31) / 32 * 32; };
    const size_t jac_n = round(static_cast<size_t>(s->E) * s->Q * 81);
    const size_t ke_n = round(static_cast<size_t>(s->E) * s->K * s->K * 9);
    const size_t kval_n = round(static_cast<size_t>(s->d.kblocks) * 9);
    const size_t need = (jac_n + ke_n + kval_n) * sizeof(double);
    if (need > dphd::kScratchBytes)
      throw std::invalid_argument("dphd_backward_static: this solver's adjoint needs " + std::to_string(need) +
                                  " bytes of scratch (Jacobian, element blocks, assembled operator), above the " +
                                  std::to_string(dphd::kScratchBytes) + "-byte static buffer; use dphd_backward");
    double* base = dphd::scratch_address(s->device);
    dphd::DeviceState d = s->d;
    d.jac = base;
    d.ke = base + jac_n;
    d.kval = base + jac_n + ke_n;
    hipStream_t user = static_cast<hipStream_t>(stream);
    dphd::flush_user(s);
    HIP_CHECK(hipEventRecord(s->ev_in, user));
    HIP_CHECK(hipStreamWaitEvent(s->stream, s->ev_in, 0));
    if (s->d.has_muscles && muscle_targets)
      HIP_CHECK(hipMemcpyAsync(s->targets, muscle_targets, sizeof(double) * s->E, hipMemcpyDeviceToDevice, s->stream));
    const dphd::StepScalars sc = dphd::scalars(s);
    dphd::launch_adjoint_setup(d, sc, checkpoint, s->targets, gq_out, gqd_out, s->stream);
    dphd::run_adjoint_direct(s, d);
    dphd::launch_adjoint_route(d, s->stream);
    dphd::launch_adjoint_outputs(d, sc, checkpoint, gqd_out, gq_in, gqd_in, gf_in, gmaterials, gactivations,
                                 gdirichlet, nullptr, s->any_pinned, s->stream);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(s->ev_out, s->stream));
    HIP_CHECK(hipStreamWaitEvent(user, s->ev_out, 0));
  })
}

int dphd_read_status(dphd_solver* s, int clear, int* bits) {
  DPHD_TRY({
    if (!s || !bits || !s->built) throw std::invalid_argument("dphd_read_status: invalid arguments");
    HIP_CHECK(hipSetDevice(s->device));
    HIP_CHECK(hipMemcpyAsync(bits, s->d.status, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    // The clear on the solver's stream, ahead of the next step's kernels.
    if (clear) HIP_CHECK(hipMemsetAsync(s->d.status, 0, sizeof(int), s->stream));
    HIP_CHECK(hipStreamSynchronize(s->stream));
  })
}

int dphd_iterations(dphd_solver* s, int adjoint, int* per_body) {
  DPHD_TRY({
    if (!s || !per_body || !s->built) throw std::invalid_argument("dphd_iterations: invalid arguments");
    HIP_CHECK(hipSetDevice(s->device));
    HIP_CHECK(hipStreamSynchronize(s->stream));
    HIP_CHECK(hipMemcpy(per_body, adjoint ? s->d.aiter : s->d.iter, sizeof(int) * s->B, hipMemcpyDeviceToHost));
  })
}

int dphd_synchronize(dphd_solver* s) {
  DPHD_TRY({
    if (!s) throw std::invalid_argument("null solver");
    HIP_CHECK(hipSetDevice(s->device));
    if (s->stream) HIP_CHECK(hipStreamSynchronize(s->stream));
  })
}

int dphd_apply_inverse(dphd_solver* s, const double* rhs, double* out, void* stream) {
  DPHD_TRY({
    if (!s || !s->built || !rhs || !out) throw std::invalid_argument("dphd_apply_inverse: invalid arguments");
    HIP_CHECK(hipSetDevice(s->device));
    hipStream_t user = static_cast<hipStream_t>(stream);
    dphd::join_user(s, user);
    dphd::launch_apply_inverse(s->d, rhs, out, user);
    HIP_CHECK(hipGetLastError());
    dphd::release_user(s, user);
  })
}

int dphd_deformation_determinants(dphd_solver* s, const float* q, double* out, void* stream) {
  if (s && s->Q != 8) {
    set_error("deformation_determinants currently requires hex_mesh");
    return DPHD_ENOTIMPL;
  }
  DPHD_TRY({
    if (!s || !s->built || !q || !out) throw std::invalid_argument("dphd_deformation_determinants: invalid arguments");
    HIP_CHECK(hipSetDevice(s->device));
    dphd::launch_hex_determinants(s->d, q, out, static_cast<hipStream_t>(stream));
    HIP_CHECK(hipGetLastError());
  })
}

int dphd_eval_projections(int count, const double* f, const double* directions, double mu, double lame,
                          double penalty, double fibre_target, double* sigma, double* reconstruction,
                          double* rotation, double* volume, double* neo_hookean, double* fibre, void* stream) {
  DPHD_TRY({
    if (count < 0 || (count && (!f || !directions || !sigma || !reconstruction || !rotation || !volume ||
                                !neo_hookean || !fibre)))
      throw std::invalid_argument("dphd_eval_projections: invalid arguments");
    dphd::launch_eval_projections(count, f, directions, mu, lame, penalty, fibre_target, sigma, reconstruction,
                                  rotation, volume, neo_hookean, fibre, static_cast<hipStream_t>(stream));
    HIP_CHECK(hipGetLastError());
  })
}

int dphd_eval_jvps(int count, const double* f, const double* df, const double* directions, double mu, double lame,
                   double fibre_target, double* rotation, double* volume, double* neo_hookean, double* fibre,
                   void* stream) {
  DPHD_TRY({
    if (count < 0 || (count && (!f || !df || !directions || !rotation || !volume || !neo_hookean || !fibre)))
      throw std::invalid_argument("dphd_eval_jvps: invalid arguments");
    dphd::launch_eval_jvps(count, f, df, directions, mu, lame, fibre_target, rotation, volume, neo_hookean, fibre,
                           static_cast<hipStream_t>(stream));
    HIP_CHECK(hipGetLastError());
  })
}

}  // extern "C"
